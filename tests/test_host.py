"""Host-side logic: config, the mixture constant pipeline, the mechanism generator, the C ABI's exports.

None of these launch anything; the library is loaded (it needs only the HIP runtime library)
and queried for its layout and compiled mechanisms.
"""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import support as sp
from peregrine_amd import build, mechgen, native
from peregrine_amd.config import Config
from peregrine_amd.mixture import fits
from peregrine_amd.mixture.constants import Ru, Tref
from peregrine_amd.mixture.fits import Nasa7
from peregrine_amd.simulator import get_simulator

ROOT = Path(__file__).resolve().parent.parent


def test_config_refuses_an_unknown_key():
    c = Config()
    with pytest.raises(KeyError):
        c["mixture"]["specis"] = "x"
    with pytest.raises(KeyError):
        c["nonsense"] = {}
    c["mixture"]["species"] = "h2o2.yaml"


def test_mechanism_parse():
    sim = get_simulator(sp.config("h2o2", chem=True))
    rx = sim.mixture.reactions
    assert len(rx) == 28
    kinds = [r["type"] for r in rx]
    assert kinds.count("elementary") == 22 and kinds.count("three-body") == 5 and kinds.count("troe") == 1
    # O + H2 <=> H + OH: A = 3.87e4 cm^3/mol/s = 38.7 m^3/kmol/s, Ea = 6260 cal/mol
    r = next(r for r in rx if r["equation"] == "O + H2 <=> H + OH")
    A, b, EaR = r["rate"]
    assert abs(A - 38.7) < 1e-12 and b == 2.7 and abs(EaR - 6260 * 4184.0 / Ru) < 1e-9
    # a three-body rate carries one more concentration: cm^6/mol^2/s -> m^6/kmol^2/s
    r = next(r for r in rx if r["equation"] == "2 O + M <=> O2 + M")
    assert abs(r["rate"][0] - 1.2e17 * 1e-6) < 1e-3
    assert r["efficiencies"] == {"H2": 2.4, "H2O": 15.4}
    troe = next(r for r in rx if r["type"] == "troe")
    assert troe["troe"] == [0.7346, 94.0, 1756.0, 5182.0]


@pytest.mark.parametrize("gas", ["air-tpg", "h2o2"])
def test_the_refits_follow_nasa7(gas):
    sim = get_simulator(sp.config(gas))
    mix = sim.mixture
    k = mix.constants()
    T = np.linspace(300.0, 3500.0, 97)
    u = np.log(T)
    for n, s in enumerate(mix.species.values()):
        n7 = s["NASA7"]
        # at the fit's own samples: within the refit tolerance, or the error of the highest degree allowed
        Ts, cps = mix.eos.caloric._sample(s)
        assert np.abs(np.polyval(np.flip(s["cppoly"]), np.log(Ts)) / cps - 1).max() <= max(1e-3, s["cpfiterr"]) * (1 + 1e-9)
        fit = np.polyval(np.flip(s["cppoly"]), u)
        # h's polynomial is cp's e^u integral: h(T2) - h(T1) = integral of cp dT, here by quadrature
        hR = T * np.polyval(np.flip(s["hpoly"]), u) + s["h298"]
        dh = np.diff(hR)
        quad = 0.5 * (fit[1:] + fit[:-1]) * np.diff(T)
        assert np.abs(dh - quad).max() < 1e-3 * np.abs(quad).max()
        # the reference enthalpy lands on NASA-7's at 298.15 K
        hRref = Tref * np.polyval(np.flip(s["hpoly"]), np.log(Tref)) + s["h298"]
        assert abs(hRref - Nasa7.hRT(n7, Tref) * Tref) < 1e-9 * max(1.0, abs(hRref))
        # g/RT from the fits against NASA-7's own, within the cp fit's error integrated
        g = mix.eos.gibbs(mix.species, T)[n]
        gn = np.array([Nasa7.hRT(n7, t) - Nasa7.sR(n7, t) for t in T])
        assert np.abs(g - gn).max() < 0.05
    assert k["cppolyR"].shape == (mix.ns, 7) and k["gpoly"].shape == (mix.ns, 8)


def test_expintegral_is_the_exponential_integral():
    P = np.array([1.5, -0.3, 0.07, 0.002])
    Q = fits.expintegral(P)
    # d/du (e^u Q) = e^u (Q + Q') = e^u P
    Qd = np.flip(np.polyder(np.flip(Q)))
    assert np.allclose(Q[: len(Qd)] + Qd, P[: len(Qd)]) and np.isclose(Q[-1], P[-1])


def test_the_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "peregrine_amd.h").read_text()
    names = re.findall(r"^\s*(?:int32_t|int64_t|const char \*)\s*(pga_\w+)\(", header, re.M)
    assert len(names) >= 20
    lib = ctypes.CDLL(str(native.LIBPATH))
    for name in names:
        assert hasattr(lib, name), name
    assert native.lib().pga_abi_version() == native.ABI_VERSION == 4


def test_the_constant_layout():
    layout, size = native.const_layout(9)
    assert layout["MW"] == (0, 9) and layout["cppolyR"][1] == 63 and layout["gpoly"][1] == 72
    spans = sorted(layout.values())
    for (o1, c1), (o2, _) in zip(spans, spans[1:]):
        assert o1 + c1 == o2
    assert spans[-1][0] + spans[-1][1] == size
    assert native.lib().pga_const_offset(9, b"nonsense", None) == -1


def test_the_compiled_mechanism_is_the_case_mechanism():
    sim = get_simulator(sp.config("h2o2", chem=True))
    mix = sim.mixture
    records, scalars = mix.chemistry.records(mix.species)
    sets = mix.chemistry.thirdbodies(mix.species)
    sig = mechgen.signature(records, sets, mix.ns)
    (name, libsig, size), = native.mechanisms()
    assert name == "h2o2" and libsig == sig
    assert size == len(mechgen.pack(records, sets, scalars, mix.ns))
    # the generated header in the tree is what the generator makes now
    text = (build.GEN / "mech_h2o2.hpp").read_text()
    assert text == mechgen.render("h2o2", records, sets, mix.ns)


def test_a_case_without_a_compiled_mechanism_is_refused_by_name():
    sim = get_simulator(sp.config("h2o2", chem=True))
    mix = sim.mixture
    records, _ = mix.chemistry.records(mix.species)
    sets = mix.chemistry.thirdbodies(mix.species)
    records[0].b = 0.0 if records[0].b else 1.0
    sig = mechgen.signature(records, sets, mix.ns)
    assert native.lib().pga_mechanism_find(sig.encode()) == -1


def test_case_create_checks_its_arguments():
    lib = native.lib()
    desc = native.CaseDesc()
    desc.ns, desc.eos, desc.mechanism = 0, 1, -1
    h = ctypes.c_void_p()
    # no device memory is touched: the pointers are only recorded
    assert lib.pga_case_create(ctypes.byref(desc), 8, None, ctypes.byref(h)) == -1
    desc.ns, desc.mechanism = 2, 0
    assert lib.pga_case_create(ctypes.byref(desc), 8, 8, ctypes.byref(h)) == -1
    desc.ns, desc.mechanism = 9, 0
    assert lib.pga_case_create(ctypes.byref(desc), 8, 8, ctypes.byref(h)) == 0
    lib.pga_case_destroy(h)
    assert b"species" in lib.pga_last_error() or lib.pga_last_error() is not None


def test_transport_grads_checks_its_arguments():
    """pga_transport_grads refuses before any launch: more than 8 boxes, a case without transport, a
    gradient box whose stencil leaves Q, a box past its arrays; an empty call launches nothing."""
    lib = native.lib()
    desc = native.CaseDesc()
    desc.ns, desc.eos, desc.mechanism, desc.trans = 1, 1, -1, 0
    h = ctypes.c_void_p()
    assert lib.pga_case_create(ctypes.byref(desc), 8, None, ctypes.byref(h)) == 0
    shape = (10, 10, 10)
    # views over no memory: every refusal here comes before a kernel could touch one
    Q, st, tr, hsp = (native.view(8, *shape, nc) for nc in (5, 8, 3, 1))
    Jinv, gr = native.view(8, *shape, 9), native.view(8, *shape, 12)
    inner = native.box((1, 1, 1), (8, 8, 8))
    boxes = (native.Box * 9)(*([inner] * 9))
    assert lib.pga_transport_grads(h, Q, st, tr, hsp, boxes, 9, Jinv, gr, inner, None) == native.PGA_EINVAL
    past = (native.Box * 1)(native.box((5, 5, 5), (8, 8, 8)))
    assert lib.pga_transport_grads(h, Q, st, tr, hsp, past, 1, Jinv, gr, inner, None) == native.PGA_ERANGE
    edge = native.box((0, 0, 0), (8, 8, 8))  # the gradients' stencil needs one layer round the box
    assert lib.pga_transport_grads(h, Q, st, tr, hsp, boxes, 0, Jinv, gr, edge, None) == native.PGA_ERANGE
    assert b"stencil" in lib.pga_last_error()
    empty = native.box((1, 1, 1), (0, 8, 8))
    assert lib.pga_transport_grads(h, Q, st, tr, hsp, boxes, 0, Jinv, gr, empty, None) == 0
    lib.pga_case_destroy(h)
    desc.trans = -1
    assert lib.pga_case_create(ctypes.byref(desc), 8, None, ctypes.byref(h)) == 0
    assert lib.pga_transport_grads(h, Q, st, tr, hsp, boxes, 1, Jinv, gr, inner, None) == native.PGA_EUNSUP
    lib.pga_case_destroy(h)


def test_integrator_weights_give_the_amplification_polynomials():
    from peregrine_amd.integrators import RK1, RK2, RK3, RK4, RK34, FixedController

    coeffs = {RK1: [1, 1], RK2: [1, 1, 1 / 2], RK3: [1, 1, 1 / 2, 1 / 6], RK34: [1, 1, 1 / 2, 1 / 6, 1 / 48],
              RK4: [1, 1, 1 / 2, 1 / 6, 1 / 24]}
    for cls, want in coeffs.items():
        integ = cls(FixedController({"simulation": {"dt": 1.0}}))
        regs = tuple(range(integ.nregs))
        for z in (0.3, -0.7, 1.1):
            v = np.zeros(integ.nregs)
            v[0] = 1.0
            # dQ = z Q at unit dt, so the step's dt factor is one
            for n in range(integ.nstages):
                Q = integ.stageregs(regs)[n].Q
                v[regs[-1]] = z * v[Q]
                for c in integ.stagekerns(n, regs):
                    acc = c.wout * v[c.out] if c.own else 0.0
                    for r, w in zip(c.X, c.w):
                        acc += w * v[r]
                    v[c.out] = acc
            assert abs(v[0] - np.polyval(np.flip(want), z)) < 1e-14


@pytest.mark.parametrize("nk,n", [(20, 3), (128, 4), (8, 8), (5, 9), (1, 1), (7, 2)])
def test_stage_slabs_cover_each_cell_and_face_once_before_it_changes(nk, n):
    """The fused stage's k-slabs: the cells partition 0..nk, the k faces 0..nk, each slab's stage has
    its own faces, and every face whose stencil (cells k - 2 to k + 1) holds a slab's cell is computed
    no later than that slab (its cells change in place after it)."""
    from peregrine_amd.solver import stage_slabs

    slabs = stage_slabs(nk, n)
    cells = [k for k0, k1, _, _ in slabs for k in range(k0, k1)]
    faces = [f for _, _, f0, f1 in slabs for f in range(f0, f1 + 1)]
    assert cells == list(range(nk)) and faces == list(range(nk + 1))
    done = set()
    for k0, k1, f0, f1 in slabs:
        done |= set(range(f0, f1 + 1))
        assert {k0, k1} <= done
        for k in range(k0, k1):
            assert set(range(max(0, k - 1), min(nk, k + 2) + 1)) <= done


# --- interface orientations (multiblock/interface/base.py:95-136, partition/orient.py:14-70,
# tests/communication/test_orientation.py) -------------------------------------------------

class _Lattice:
    """The multiblock a mesher fills, without device arrays."""

    def __init__(self, ng):
        self.blocks, self.ng = [], ng

    def add_blk(self):
        from peregrine_amd.solver import Block

        blk = Block(len(self.blocks), self)
        self.blocks.append(blk)
        return blk


def _reoriented_pair(S, ng):
    mb = _Lattice(ng)
    sp.ReorientedPair(S, mbdims=[2, 1, 1], blkdims=[6, 4, 3], lengths=[2.0, 1.3, 0.7], defaultbc="wall").fill(mb)
    return mb


@pytest.mark.parametrize("S", sp.orientations())
def test_a_reoriented_neighbour_places_each_halo_cell_where_it_lies(S):
    """Block 1 of a pair relabelled by each of the 24 orientations: the interfaces meet back with
    inverse digits, and every halo cell the exchange places (ng = 2 and 1 deep) is the neighbour's
    cell at the same place: the halo's centre from its own extrapolated nodes is the source cell's
    centre (the mapping stepped as the device steps it)."""
    from peregrine_amd import grid
    from peregrine_amd.solver import placement

    ng = 2
    mb = _reoriented_pair(S, ng)
    b0, b1 = mb.blocks
    f0 = next(f for f in b0.faces.values() if f.nbr == 1)
    f1 = next(f for f in b1.faces.values() if f.nbr == 0)
    assert f0.orient == S and f1.orient == f0.nbrorient and f0.nbrnface == f1.nface and f1.nbrnface == f0.nface
    centres = [grid.cellcenters(grid.halo_nodes(b.interior_nodes, ng, [False] * 3, b.period)) for b in mb.blocks]
    for blk, f, nbr in ((b0, f0, b1), (b1, f1, b0)):
        for depth in (ng, 1):
            dst, src, ext, orient = placement(blk, f, nbr, depth)
            a = np.stack(np.meshgrid(*map(np.arange, ext), indexing="ij"), axis=-1)
            s = np.broadcast_to(np.array(src), a.shape).copy()
            for m, code in enumerate(orient):
                A, back = (code - 1) % 3, code > 3
                s[..., A] += -a[..., m] if back else a[..., m]
            d = a + np.array(dst)
            got = centres[blk.nblki][d[..., 0], d[..., 1], d[..., 2]]
            want = centres[nbr.nblki][s[..., 0], s[..., 1], s[..., 2]]
            assert np.allclose(got, want, rtol=0, atol=1e-12), (S, blk.nblki, depth)
            # and the source cells are interior cells of the neighbour
            for A in range(3):
                assert s[..., A].min() >= ng and s[..., A].max() < ng + nbr.ncell[A]


def test_a_block_joined_to_itself_is_not_reoriented():
    from peregrine_amd import mesher
    from peregrine_amd.partition import Conditioner

    mb = _Lattice(2)
    mesher.CubeMesher(blkdims=[5, 5, 5], periodic=[True, False, False], defaultbc="wall").fill(mb)
    with pytest.raises(ValueError, match="joined to itself"):
        Conditioner().reorient(mb, mb.blocks[0], [1, 0, 2], [False, False, True])


# --- rotational periodics (mesher/annulus.py, tests/communication/test_xchg.py:142-213) --------

def _annulus(ng, mbdims=(1, 1, 2), sweep=90.0):
    from peregrine_amd import mesher

    mb = _Lattice(ng)
    mesher.AnnulusMesher(mbdims=list(mbdims), blkdims=[4, 4, 4], p1=[0, 0, 0], p2=[1, 0, 0], p3=[0, 0.5, 0],
                         sweep=sweep, thickness=0.5, periodic=True, defaultbc="wall").fill(mb)
    return mb


@pytest.mark.parametrize("ng", [1, 2])
def test_a_rotational_periodics_halo_continues_the_sweep(ng):
    """A 90-degree sector in two blocks of three 15-degree cells: the halo nodes across its ends are
    the far end's turned by the sweep, so the cell centres step by 15 degrees through halo and
    interior; the ends turn (y, z) -> (z, -y) coming in at the low end."""
    from peregrine_amd import grid
    from peregrine_amd.solver import block_nodes

    mb = _annulus(ng)
    b0, b1 = mb.blocks
    low, high = b0.faces[5], b1.faces[6]
    assert low.rotperiodic and high.rotperiodic and low.nbr == 1 and high.nbr == 0
    assert np.allclose(low.rot @ [1.0, 2.0, 3.0], [1.0, 3.0, -2.0], atol=1e-15)
    assert not b0.faces[6].periodic and b0.faces[6].nbr == 1
    nodes = block_nodes(mb.blocks, ng)
    for blk, first in ((b0, -7.5 - 15 * (ng - 1)), (b1, 45.0 - 7.5 - 15 * (ng - 1))):
        x = grid.cellcenters(nodes[blk.nblki])[ng, ng]
        angles = np.degrees(np.arctan2(x[:, 2], x[:, 1]))
        want = first + 15 * np.arange(3 + 2 * ng)
        assert np.allclose(angles, np.where(want > 180, want - 360, want), atol=1e-12)


def test_a_full_turn_joins_its_ends_as_an_interior_interface():
    from peregrine_amd import grid
    from peregrine_amd.solver import block_nodes

    mb = _annulus(2, mbdims=(1, 1, 4), sweep=360.0)
    ends = mb.blocks[0].faces[5], mb.blocks[3].faces[6]
    assert all(f.nbr is not None and not f.periodic and f.rot is None for f in ends)
    nodes = block_nodes(mb.blocks, 2)
    x = grid.cellcenters(nodes[0])[2, 2]
    assert np.allclose(np.degrees(np.arctan2(x[:, 2], x[:, 1])), -45.0 + 30 * np.arange(7), atol=1e-12)


def test_a_physics_plugin_is_claimed_for_its_physics_only():
    """plugins.viscsponge is Navier-Stokes' (plugins/simulator/viscsponge.py:9-10): refused on an
    Euler case saying which it is for, refused without its line, and an unknown name refused."""
    from types import SimpleNamespace

    from peregrine_amd.plugins import ViscSponge, claim

    sect = {"origin": [0, 0, 0], "ending": [2.0, 0, 0], "multiplier": 4.0}
    ns, euler = (SimpleNamespace(sim=SimpleNamespace(name=n)) for n in ("navier-stokes", "euler"))
    solver_plugins, physics = claim({"viscsponge": sect, "nancheck": {}}, ns)
    assert [p.name for p in solver_plugins] == ["nancheck"] and isinstance(physics[0], ViscSponge)
    assert list(physics[0].line) == [1.0, 0.0, 0.0, 0.0, 2.0, 4.0]
    with pytest.raises(ValueError, match="is for navier-stokes, not euler"):
        claim({"viscsponge": sect}, euler)
    with pytest.raises(ValueError, match="needs"):
        claim({"viscsponge": {"origin": [0, 0, 0]}}, ns)
    with pytest.raises(ValueError, match="no plugin of that name"):
        claim({"catalyst": {}}, ns)


def test_a_batch_is_independent_periodic_boxes():
    """The uneven batch's mesher: each block its own periodic box, joined to nothing else; its
    halo cells continue its own lattice across each end."""
    from peregrine_amd import grid, mesher
    from peregrine_amd.solver import block_nodes

    mb = _Lattice(2)
    dims = [(16, 20, 24), (30, 17, 16), (16, 16, 96)]
    mesher.BatchMesher(dims, lengths=(0.01, 0.01, 0.01)).fill(mb)
    assert [blk.ncell for blk in mb.blocks] == dims
    for blk in mb.blocks:
        assert all(f.nbr == blk.nblki and f.periodic and f.rot is None for f in blk.faces.values())
    nodes = block_nodes(mb.blocks, 2)
    for blk in mb.blocks:
        x = grid.cellcenters(nodes[blk.nblki])
        for a in range(3):
            line = np.moveaxis(x[..., a], a, 0)[:, 3, 3]
            step = 0.01 / blk.ncell[a]
            assert np.allclose(np.diff(line), step, rtol=0, atol=1e-15)
    q = sp.h2_prims(grid.cellcenters(nodes[1]), mb.blocks[1].ncell, lengths=(0.01,) * 3)
    assert q.shape == mb.blocks[1].cellshape + (13,) and np.isfinite(q).all()
