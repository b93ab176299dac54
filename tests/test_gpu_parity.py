"""The HIP path against the oracle, through the C ABI, on the same seeded inputs.

Every kernel is fed the same arrays as the oracle; results are compared per conserved row,
normalised by the row's largest magnitude over the block as the reference's tests normalise
(tests/simulator/test_facefluxes.py:43-46). Tolerances are written at each check with the
measured value beside them when known.
"""

import ctypes

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd import native
from peregrine_amd.backend import Mi355xBackend
from peregrine_amd.simulator import NativeCase, get_simulator

pytestmark = pytest.mark.gpu


def _rowerr(got, want, axis0=True):
    got, want = np.asarray(got), np.asarray(want)
    g = got.reshape(got.shape[0], -1)
    w = want.reshape(want.shape[0], -1)
    scale = np.abs(w).max(axis=1)
    scale = np.where(scale > 0, scale, 1.0)
    return (np.abs(g - w).max(axis=1) / scale).max()


class Dev:
    """Device copies of host (comp, k, j, i) arrays and their ABI views."""

    def __init__(self, torch):
        self.torch = torch

    def put(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).cuda()

    @staticmethod
    def view(t):
        nc, nk, nj, ni = t.shape
        v = native.view(t.data_ptr(), ni, nj, nk, nc)
        # the view holds its tensor: a temporary stays allocated until the launch it is passed to
        v._keep = t
        return v

    def get(self, t):
        self.torch.cuda.synchronize()
        return t.cpu().numpy()


@pytest.fixture(scope="module")
def dev(gpu):
    import torch

    return Dev(torch)


def _case(gas, recon="constant", primary="kepaec", limiter=None, chem=False):
    cfg = sp.config(gas, recon=recon, primary=primary, limiter=limiter, chem=chem)
    sim = get_simulator(cfg)
    return cfg, sim, NativeCase(sim, Mi355xBackend(cfg)), O.OracleCase(sim.mixture, recon, limiter, primary)


def _prepared(sim, oc, hb, q):
    """Q and state from primitives by the oracle, halos filled, halo state by the oracle's EOS."""
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    ng = hb.ng
    for a in range(3):
        O.lib.pgo_fill_periodic(O.arr(hb.regs[0]), sim.ne, O.ints(hb.ncell), ng, ng, a)
        O.lib.pgo_fill_periodic(O.arr(hb.state), 8, O.ints(hb.ncell), ng, ng, a)


GASES = ["air-cpg", "air-tpg", "binary", "h2o2"]


@pytest.mark.parametrize("gas", GASES)
def test_fromprims_and_fromQ(dev, gas):
    cfg, sim, nc, oc = _case(gas)
    hb = sp.HostBlock((70, 9, 5), 1, sim.ne, 2, jostle=0.2)
    rng = np.random.default_rng(12)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = 101325.0 * rng.uniform(0.5, 3.0, hb.cshape)
    q[..., 1:4] = rng.uniform(-100, 100, hb.cshape + (3,))
    q[..., 4] = rng.uniform(350.0, 2500.0, hb.cshape)
    if sim.ns > 1:
        Y = rng.random(hb.cshape + (sim.ns,))
        Y /= Y.sum(axis=-1, keepdims=True)
        q[..., 5:] = Y[..., :-1]
    P = sp.to_dev(q)
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(P), O.arr(hb.regs[0]), O.arr(hb.state), full)
    dP, dQ, dS = dev.put(P), dev.put(np.zeros_like(hb.regs[0])), dev.put(np.zeros_like(hb.state))
    L = native.lib()
    native.check(L.pga_fromprims(nc.handle, dev.view(dP), dev.view(dQ), dev.view(dS), native.box((0, 0, 0), hb.cshape), None),
                 "pga_fromprims")
    # measured: bitwise for cpg; the tpg cp/h polynomials round apart by 1e-15
    assert _rowerr(dev.get(dQ), hb.regs[0]) < 1e-13
    assert _rowerr(dev.get(dS), hb.state) < 1e-13
    # fromQ from a state whose T is off by up to 5%, on two boxes of the array
    st = hb.state.copy()
    st[1] *= 1.0 + 0.05 * rng.uniform(-1, 1, st[1].shape)
    dS = dev.put(st)
    boxes = [((0, 0, 0), (72, 11, 3)), ((0, 0, 3), (72, 11, 4))]
    for b in boxes:
        O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(st), O.box(*b))
    arr = (native.Box * 2)(*[native.box(*b) for b in boxes])
    native.check(L.pga_fromQ(nc.handle, dev.view(dev.put(hb.regs[0])), dev.view(dS), arr, 2, None), "pga_fromQ")
    # the same Newton from the same guesses: ln T is taken by one method on both sides (elog)
    err = _rowerr(dev.get(dS), st)
    print(f"fromQ {gas}: {err:.3e}")
    assert err < 1e-13


FLUXES = [("kepaec", "constant", None), ("hllc", "constant", None), ("rusanov", "constant", None),
          ("hllc", "muscl", "koren"), ("rusanov", "muscl", "koren"), ("hllc", "muscl", "minmod"),
          ("rusanov", "muscl", "mc"), ("hllc", "muscl", "superbee"), ("hllc", "muscl", "vanleer")]


@pytest.mark.parametrize("primary,recon,limiter", FLUXES)
@pytest.mark.parametrize("gas", ["air-tpg", "binary", "h2o2"])
def test_face_flux_and_divergence(dev, gas, primary, recon, limiter):
    cfg, sim, nc, oc = _case(gas, recon, primary, limiter)
    hb = sp.HostBlock((67, 9, 6), sim.ng, sim.ne, 3, jostle=0.3, seed=5)
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=sim.ns)
    if gas == "h2o2":
        q[..., 5:] *= 0.1
    _prepared(sim, oc, hb, q)
    ng, L = hb.ng, native.lib()
    dQv, dSt = dev.put(hb.regs[0]), dev.put(hb.state)
    dF = []
    for a in range(3):
        ext = list(hb.ncell)
        ext[a] += 1
        O.lib.pgo_flux(oc.ref, a, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.S[a]), O.arr(hb.F[a]), O.box((ng,) * 3, ext))
        F = dev.put(np.zeros_like(hb.F[a]))
        native.check(L.pga_flux(nc.handle, a, dev.view(dQv), dev.view(dSt), dev.view(dev.put(hb.S[a])), dev.view(F),
                                native.box((ng,) * 3, ext), None), "pga_flux")
        err = _rowerr(dev.get(F), hb.F[a])
        print(f"flux {gas} {primary} {recon} {limiter} axis {a}: {err:.3e}")
        assert err < 1e-12
        dF.append(F)
    # the divergence on identical face fluxes is bitwise: one order of sums, one product
    for append in (0, 1):
        dq = np.random.default_rng(3).random(hb.regs[2].shape)
        want = dq.copy()
        O.lib.pgo_applyflux(sim.ne, O.arr(want), *(O.arr(f) for f in hb.F), O.arr(hb.detJinv), append,
                            O.box((ng,) * 3, hb.ncell))
        got = dev.put(dq)
        Fs = [dev.put(f) for f in hb.F]
        native.check(L.pga_applyflux(sim.ne, dev.view(got), *(dev.view(f) for f in Fs), dev.view(dev.put(hb.detJinv)),
                                     append, native.box((ng,) * 3, hb.ncell), None), "pga_applyflux")
        assert np.array_equal(dev.get(got), want)


def test_production_at_random_states(dev):
    cfg, sim, nc, oc = _case("h2o2", "muscl", "hllc", "koren", chem=True)
    mix = sim.mixture
    rng = np.random.default_rng(21)
    n = 4099
    T = rng.uniform(800.0, 3000.0, n)
    Y = rng.random((mix.ns, n)) ** 3
    Y /= Y.sum(axis=0)
    from chem_exact import density

    rho = np.array([density(mix, Y[:, i], T[i], 101325.0 * rng.uniform(0.3, 30)) for i in range(n)])
    want = np.stack([oc.production(rho[i], Y[:, i], T[i]) for i in range(n)], axis=1)
    t = dev.torch
    d_rho, d_Y, d_T = (t.from_numpy(np.ascontiguousarray(a)).cuda() for a in (rho, Y, T))
    d_w = t.zeros((mix.ns, n), dtype=t.float64, device="cuda")
    native.check(native.lib().pga_production(nc.handle, n, d_rho.data_ptr(), d_Y.data_ptr(), d_T.data_ptr(),
                                             d_w.data_ptr(), None), "pga_production")
    got = dev.get(d_w)
    # per point, against that point's largest rate: the forward and reverse rates round apart
    # at 1e-14 and near equilibrium cancel
    scale = np.abs(want).max(axis=0)
    err = (np.abs(got - want).max(axis=0) / scale).max()
    print(f"production at {n} states: {err:.3e} of each point's largest rate")
    assert err <= 1e-12


def test_source_on_a_hot_block(dev):
    cfg, sim, nc, oc = _case("h2o2", "muscl", "hllc", "koren", chem=True)
    hb = sp.HostBlock((40, 12, 10), sim.ng, sim.ne, 3, lengths=(0.01,) * 3)
    q = sp.h2_prims(hb.xcc, hb.ncell)
    # some radicals, so every reaction runs
    q[..., 6] = 1e-4
    q[..., 9] = 2e-4
    _prepared(sim, oc, hb, q)
    want = np.zeros_like(hb.regs[2])
    O.lib.pgo_source(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(want), O.box((2, 2, 2), hb.ncell))
    got = dev.put(np.zeros_like(want))
    native.check(native.lib().pga_source(nc.handle, dev.view(dev.put(hb.regs[0])), dev.view(dev.put(hb.state)),
                                         dev.view(got), native.box((2, 2, 2), hb.ncell), None), "pga_source")
    g = dev.get(got)
    assert np.array_equal(g[:5], want[:5])
    err = _rowerr(g[5:], want[5:])
    print(f"source on a hot block: {err:.3e}")
    assert err < 1e-12


def test_combine(dev):
    rng = np.random.default_rng(4)
    shape = (6, 7, 9, 70)
    X = [rng.random(shape) for _ in range(3)]
    out0 = rng.random(shape)
    L = native.lib()
    t = dev.torch
    dt = t.tensor([1e-3], dtype=t.float64, device="cuda")
    for own, nterms, step in ((1, 2, 1), (0, 3, 1), (1, 1, 0)):
        w = [0.75, 0.25, 1.0 / 3.0][:nterms]
        want = out0.copy()
        Xs = (O.Arr * 3)(*[O.arr(x) for x in X])
        O.lib.pgo_combine(6, O.arr(want), own, 1.0 / 3.0, nterms, Xs, (ctypes.c_double * 3)(*w), 1e-3, step,
                          O.box((1, 1, 1), (68, 7, 5)))
        got = dev.put(out0)
        dX = [dev.put(x) for x in X]
        Xv = (native.View * 3)(*[dev.view(x) for x in dX])
        native.check(L.pga_combine(6, dev.view(got), own, 1.0 / 3.0, nterms, Xv, (ctypes.c_double * 3)(*w),
                                   dt.data_ptr(), step, native.box((1, 1, 1), (68, 7, 5)), None), "pga_combine")
        # one rounding per operation in the same order: the same bits
        assert np.array_equal(dev.get(got), want)


def test_halo_copies(dev):
    rng = np.random.default_rng(8)
    ncell, ng = (70, 6, 5), 2
    shape = (4,) + tuple(n + 2 * ng for n in ncell)[::-1]
    A = rng.random(shape)
    want = A.copy()
    for a in range(3):
        O.lib.pgo_fill_periodic(O.arr(want), 4, O.ints(ncell), ng, ng, a)
    dA = dev.put(A)
    L = native.lib()
    for a in range(3):
        native.check(L.pga_fill_periodic(dev.view(dA), 4, native.ints(ncell), ng, ng, a, None), "pga_fill_periodic")
    assert np.array_equal(dev.get(dA), want)
    # the same halos by box copies, as the solver exchanges them
    dB = dev.put(A)
    for a in range(3):
        for low in (True, False):
            ext = list(ncell)
            ext[a] = ng
            dst, src = [ng] * 3, [ng] * 3
            dst[a] = 0 if low else ng + ncell[a]
            src[a] = ncell[a] if low else ng
            native.check(L.pga_copy_box(dev.view(dB), native.ints(dst), dev.view(dB), native.ints(src), native.ints(ext),
                                        4, None, None, -1, 0, None), "pga_copy_box")
    assert np.array_equal(dev.get(dB), want)


def _run_pair(gas, recon, primary, limiter, chem, integrator, ncell, steps, dt, prims, lengths, fused=True,
              trans=None, species=None, nstate=8, walls=(), substep=None, mixture=None, slabs=1, sponge=None,
              implicit=None):
    """The same case stepped by the solver on the device and by the oracle on the host; walls: the
    faces (0 i-low .. 5 k-high) that are adiabatic slip walls, the rest periodic; substep: the
    sub-stepped chemistry's (max-substeps, bulk-substeps, stiff-bins, equilibrium-tolerance), the
    oracle taking the one-pass form; implicit: the implicit chemistry's (parts, max-sweeps,
    refine-tolerance), the oracle taking the FP64 form."""
    import peregrine_amd as pg

    cfg = sp.config(gas, recon=recon, primary=primary, limiter=limiter, chem=chem, integrator=integrator, dt=dt,
                    trans=trans)
    if substep is not None:
        ci = cfg["chem-integration"]
        ci["kind"] = "sub-stepped"
        ci["max-substeps"], ci["bulk-substeps"], ci["stiff-bins"], ci["equilibrium-tolerance"] = substep
    if implicit is not None:
        ci = cfg["chem-integration"]
        ci["kind"] = "implicit"
        ci["implicit-parts"], ci["max-sweeps"], ci["refine-tolerance"] = implicit
    if species is not None:
        cfg["mixture"]["species"] = species
    for key, value in (mixture or {}).items():
        cfg["mixture"][key] = value
    # walls: faces that are adiabatic slip walls, or {face: (kind, values)} of any wall kind
    spec = walls if isinstance(walls, dict) else {f: ("slp-adia-wall", {}) for f in walls}
    for f, (kind, values) in spec.items():
        cfg["bcvals"][f"wall{f}"] = {"kind": kind, **values}
    oracle_walls = {f: (native.WALL_KINDS.index(kind), tuple(values.get(v, 0.0) for v in "uvw"), values.get("T", 0.0))
                    for f, (kind, values) in spec.items()}
    cfg["backend-mi355x"]["fused-stage"] = fused
    cfg["backend-mi355x"]["stage-slabs"] = slabs
    if sponge is not None:
        cfg["plugins"]["viscsponge"] = sponge
    backend = pg.get_backend("mi355x", cfg)
    periodic = [not (2 * a in walls or 2 * a + 1 in walls) for a in range(3)]
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1 for n in ncell], lengths=lengths, periodic=periodic,
                                bcnames={f + 1: f"wall{f}" for f in spec})
    mb = pg.get_solver(cfg, mesh, backend)
    sim = mb.sim
    oc = O.OracleCase(sim.mixture, recon, limiter, primary)
    nregs = 4 if integrator == "rk4" else 3
    hb = sp.HostBlock(ncell, sim.ng, sim.ne, nregs, lengths=lengths, nstate=nstate)
    q = prims(hb.xcc)
    mb.set_prims([q])
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    # the solver's set_prims ends as the reference's from_prims does (multiblock/solver.py:298-305):
    # the EOS over the interior, then the halo exchange and the EOS over the halos, which its first
    # stage then takes as ready; the oracle's first stage exchanges and takes the halo EOS itself
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box((hb.ng,) * 3, hb.ncell))
    b = sp.oracle_block(hb, integrator, chem, viscous=sim.viscous, alpha=getattr(sim, "alpha", 1.0),
                        walls=oracle_walls)
    if sponge is not None:
        # the oracle takes the line the plugin made of the section
        b.sponge = 1
        b.spongeline[:] = list(mb.sponges[0].line)
    if substep is not None:
        b.chemistry, b.subcap, b.subeqtol = 2, substep[0], substep[3]
        hb.counts = np.zeros(int(np.prod(ncell)), dtype=np.int32)
        b.substeps = hb.counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    if implicit is not None:
        b.chemistry, b.imparts = 3, implicit[0]
    for _ in range(steps):
        mb.step(dt)
        O.lib.pgo_step(oc.ref, ctypes.byref(b), dt)
    return mb, hb


def test_reacting_rhs_matches(dev):
    """One right-hand side of the reacting case (prerhs + rhs: halo exchange, halo EOS, source,
    three flux sweeps, divergence) against the oracle's: 1e-12 per conserved row."""
    import peregrine_amd as pg

    n, L = 24, (0.01,) * 3
    cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, dt=1e-8)
    mb = pg.get_solver(cfg, pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=L, periodic=[True] * 3),
                       pg.get_backend("mi355x", cfg))
    sim = mb.sim
    oc = O.OracleCase(sim.mixture, "muscl", "koren", "hllc")
    hb = sp.HostBlock((n,) * 3, sim.ng, sim.ne, 3, lengths=L)
    q = sp.reacting_flow(hb.xcc, n, L)
    mb.set_prims([q])
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box((hb.ng,) * 3, hb.ncell))
    sp.halo_eos(oc, hb, sim.ne)
    # the halos' EOS ran once, in set_prims: the rhs begins with the source
    regs = mb.integ.stageregs(mb.regs)[0]
    mb.rhs(Q=regs.Q, dQ=regs.dQ)
    ng = hb.ng
    dQ = hb.regs[2]
    O.lib.pgo_source(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(dQ), O.box((ng,) * 3, hb.ncell))
    for a in range(3):
        ext = list(hb.ncell)
        ext[a] += 1
        O.lib.pgo_flux(oc.ref, a, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.S[a]), O.arr(hb.F[a]), O.box((ng,) * 3, ext))
    O.lib.pgo_applyflux(sim.ne, O.arr(dQ), *(O.arr(f) for f in hb.F), O.arr(hb.detJinv), 1, O.box((ng,) * 3, hb.ncell))
    got = mb.blocks[0].regs[regs.dQ].get()[ng:-ng, ng:-ng, ng:-ng].transpose(3, 2, 1, 0)
    err = _rowerr(got, dQ[hb.interior()])
    print(f"reacting rhs: max row-normalised difference {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_reacting_steps_match(dev, fused):
    """H2/O2-N2, MUSCL + HLLC, coupled chemistry, tvd-rk3: the reacting case at 70 x 24 x 24 (tiles
    that overhang the block), 10 steps, by the fused stage and by the per-kernel sequence."""
    n = 24
    L = (0.01,) * 3
    mb, hb = _run_pair("h2o2", "muscl", "hllc", "koren", True, "tvd-rk3", (70, n, n), 10, 1e-8,
                       lambda x: sp.reacting_flow(x, n, L), (0.01 * 70 / 24, 0.01, 0.01), fused=fused)
    assert mb.fused == fused
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    want = hb.regs[0][hb.interior()]
    err = _rowerr(got, want)
    print(f"reacting 10 steps ({'fused' if fused else 'per-kernel'}): max row-normalised difference {err:.3e}")
    assert err < 1e-10


def test_fused_and_per_kernel_stages_agree_to_the_bit(dev):
    """The fused stage against the per-kernel sequence on the reacting case, RK3, RK2 and RK1: the same
    device functions in the same order give the same bits; and the march (the k sweep inside the
    stage's launch, in chunks of 7, 7 and 6 planes and of 20) gives the split k sweep's bits."""
    import peregrine_amd as pg

    n, L = 20, (0.01,) * 3
    for integ in ("tvd-rk3", "rk2", "rk1"):
        out = []
        # the fused stage whole, in three uneven k-slabs (7, 7 and 6 cells), marching, and the
        # per-kernel one
        for fused, slabs, march in ((True, 1, 0), (True, 3, 0), (True, 1, 7), (True, 1, 20), (False, 1, 0)):
            cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, integrator=integ, dt=1e-8)
            cfg["backend-mi355x"]["fused-stage"] = fused
            cfg["backend-mi355x"]["k-chunk"] = 3
            cfg["backend-mi355x"]["stage-slabs"] = slabs
            cfg["backend-mi355x"]["k-march"] = march
            mb = pg.get_solver(cfg, pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=L, periodic=[True] * 3),
                               pg.get_backend("mi355x", cfg))
            assert mb.fused == fused
            mb.set_prims([sp.reacting_flow(mb.blocks[0].xcc.get(), n, L)])
            for _ in range(3):
                mb.step(1e-8)
            out.append((mb.blocks[0].Q.get(), mb.blocks[0].state.get()))
        g = mb.ng
        inner = np.s_[g:-g, g:-g, g:-g]
        for other in out[1:]:
            assert np.array_equal(out[0][0][inner], other[0][inner]), integ
            assert np.array_equal(out[0][1][inner], other[1][inner]), integ


def test_the_march_refuses_to_write_its_input(dev):
    """pga_stage_march with Qout = Qin: refused before any launch (a chunk's last k face would read
    cells the next chunk has written)."""
    cfg, sim, nc, oc = _case("air-tpg", recon="muscl", primary="hllc", limiter="koren")
    L = native.lib()
    hb = sp.HostBlock((8, 8, 8), sim.ng, sim.ne, 1)
    Q, st = dev.put(hb.regs[0]), dev.put(hb.state)
    faces = [dev.put(f) for f in hb.F]
    S = [dev.put(x) for x in hb.S]
    det = dev.put(hb.detJinv)
    ddt = dev.put(np.array([1e-6]))
    box = native.box((hb.ng,) * 3, (8, 8, 8))
    rc = L.pga_stage_march(nc.handle, dev.view(Q), dev.view(st), dev.view(Q), dev.view(Q), dev.view(st),
                           dev.view(faces[0]), dev.view(faces[1]), dev.view(S[2]), dev.view(det), native.View(), 1,
                           native.ints([2]), (ctypes.c_double * 1)(1.0), native.ints([1]), ddt.data_ptr(), box, 4, None)
    assert rc == -1 and b"output is the input" in L.pga_last_error()


@pytest.mark.parametrize("gas", ["air-tpg", "air-cpg"])
def test_tgv_rk4_100_steps_match(dev, gas):
    """Taylor-Green at Mach 0.1 (air, thermally perfect as the 32^3 case states, and calorically
    perfect), KEEP, RK4, 100 steps at 16^3."""
    n = 16
    L = (2 * np.pi,) * 3
    R, cp = 8314.46261815324 / 28.96, 1005.0
    gamma = cp / (cp - R)
    c0 = np.sqrt(gamma * R * 300.0)
    V0 = 0.1 * c0
    rho0 = 101325.0 / (R * 300.0)

    def prims(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        q = np.zeros(x.shape[:3] + (5,))
        q[..., 0] = 101325.0 + rho0 * V0**2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
        q[..., 1] = V0 * np.sin(X) * np.cos(Y) * np.cos(Z)
        q[..., 2] = -V0 * np.cos(X) * np.sin(Y) * np.cos(Z)
        q[..., 4] = q[..., 0] / (rho0 * R)
        return q

    dt = 0.2 * (2 * np.pi / n) / (c0 + V0)
    mb, hb = _run_pair(gas, "constant", "kepaec", None, False, "rk4", (n,) * 3, 100, dt, prims, L)
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"tgv {gas} rk4 100 steps: max row-normalised difference {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("trans", [None, "constant"], ids=["euler", "navier-stokes"])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_blocks_of_a_lattice_step_as_one_block(dev, fused, trans):
    """2 x 2 x 1 blocks against one: the halo exchange between blocks is the periodic fill's equal;
    with the viscous terms, the gradients' exchange too (the fused step's transport and gradients
    merged in one launch per block)."""
    import peregrine_amd as pg

    n = 16
    L = (1.0,) * 3
    results = []
    field = None
    for mbdims in ((1, 1, 1), (2, 2, 1)):
        cfg = sp.config("air-tpg", recon="muscl", primary="hllc", limiter="koren", dt=2e-6, trans=trans)
        if trans:
            cfg["mixture"]["species"] = {"Air": {**sp.AIR_NASA7["Air"], "mu0": 1.8e-5, "kappa0": 0.026}}
        cfg["backend-mi355x"]["fused-stage"] = fused
        backend = pg.get_backend("mi355x", cfg)
        blkdims = [n // m + 1 for m in mbdims]
        mesh = pg.mesher.CubeMesher(mbdims=mbdims, blkdims=blkdims, lengths=L, periodic=[True] * 3)
        mb = pg.get_solver(cfg, mesh, backend)
        g = mb.ng
        if field is None:
            # the whole lattice's primitives with its periodic halo: each block's halo starts as its
            # partners' cells, so every halo EOS starts from the same T guess in both runs (the
            # Newton iterate's last bits depend on its start)
            inner = sp.smooth_prims(mb.blocks[0].xcc.get(), mb.sim.ne)[g:-g, g:-g, g:-g]
            field = np.pad(inner, [(g, g)] * 3 + [(0, 0)], mode="wrap")
        prims = []
        for blk in mb.blocks:
            i, j = blk.nblki % mbdims[0], (blk.nblki // mbdims[0]) % mbdims[1]
            ni, nj, nk = blk.ncell
            prims.append(field[i * ni:(i + 1) * ni + 2 * g, j * nj:(j + 1) * nj + 2 * g, :].copy())
        mb.set_prims(prims)
        for _ in range(3):
            mb.step(2e-6)
        whole = np.zeros((n, n, n, mb.sim.ne))
        for blk in mb.blocks:
            i, j = blk.nblki % mbdims[0], (blk.nblki // mbdims[0]) % mbdims[1]
            ni, nj, nk = blk.ncell
            whole[i * ni:(i + 1) * ni, j * nj:(j + 1) * nj, :] = blk.Q.get()[g:-g, g:-g, g:-g]
        results.append(whole)
    diff = np.abs(results[1] - results[0]).max(axis=(0, 1, 2)) / np.abs(results[0]).max(axis=(0, 1, 2))
    print(f"lattice 2x2x1 vs one block: max row-normalised difference {diff.max():.3e}")
    # each cell's arithmetic is the same whichever block holds it, and so is each halo EOS's
    # start: the same bits
    assert np.array_equal(results[1], results[0])


# --- the viscous terms -----------------------------------------------------------------------


@pytest.mark.parametrize("gas,trans,diffusion,bath", [
    ("air-tpg", "constant", "lewis", None), ("binary", "constant", "lewis", None),
    ("h2o2", "kinetic-theory", "lewis", None), ("h2o2", "kinetic-theory", "binary", None),
    ("h2o2", "kinetic-theory", "binary", ["H2", "O2", "H2O", "N2"])])
def test_transport_gradients_and_viscous_flux(dev, gas, trans, diffusion, bath):
    """pga_transport, pga_grads and pga_diffflux against the oracle on a jostled block: the same
    operations in the same order, so the same bits; Lewis numbers and binary diffusion, over every
    species and over a partial bath."""
    cfg = sp.config(gas, trans=trans)
    cfg["mixture"]["diffusion"], cfg["mixture"]["bath"] = diffusion, bath
    sim = get_simulator(cfg)
    nc = NativeCase(sim, Mi355xBackend(cfg))
    oc = O.OracleCase(sim.mixture)
    hb = sp.HostBlock((67, 9, 6), 1, sim.ne, 2, jostle=0.3, seed=5)
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=sim.ns)
    if gas == "h2o2":
        q[..., 5:] *= 0.1
    _prepared(sim, oc, hb, q)
    L = native.lib()
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp), full)
    dQv, dSt = dev.put(hb.regs[0]), dev.put(hb.state)
    dTr, dH = dev.put(np.zeros_like(hb.trstate)), dev.put(np.zeros_like(hb.hsp))
    boxes = (native.Box * 1)(native.box((0, 0, 0), hb.cshape))
    native.check(L.pga_transport(nc.handle, dev.view(dQv), dev.view(dSt), dev.view(dTr), dev.view(dH), boxes, 1, None),
                 "pga_transport")
    assert np.array_equal(dev.get(dTr), hb.trstate)
    assert np.array_equal(dev.get(dH), hb.hsp)
    inner = O.box((1, 1, 1), tuple(n - 2 for n in hb.cshape))
    O.lib.pgo_grads(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.Jinv), O.arr(hb.grads), inner)
    dG = dev.put(np.zeros_like(hb.grads))
    native.check(L.pga_grads(nc.handle, dev.view(dQv), dev.view(dSt), dev.view(dev.put(hb.Jinv)), dev.view(dG),
                             native.box((1, 1, 1), tuple(n - 2 for n in hb.cshape)), None), "pga_grads")
    assert np.array_equal(dev.get(dG), hb.grads)
    # the two in one launch, the transport's cells as two boxes (k < 2 and the rest)
    ni, nj, nk = hb.cshape
    two = (native.Box * 2)(native.box((0, 0, 0), (ni, nj, 2)), native.box((0, 0, 2), (ni, nj, nk - 2)))
    dTr2, dH2, dG2 = (dev.put(np.zeros_like(x)) for x in (hb.trstate, hb.hsp, hb.grads))
    native.check(L.pga_transport_grads(nc.handle, dev.view(dQv), dev.view(dSt), dev.view(dTr2), dev.view(dH2), two, 2,
                                       dev.view(dev.put(hb.Jinv)), dev.view(dG2),
                                       native.box((1, 1, 1), (ni - 2, nj - 2, nk - 2)), None), "pga_transport_grads")
    assert np.array_equal(dev.get(dTr2), hb.trstate)
    assert np.array_equal(dev.get(dH2), hb.hsp)
    assert np.array_equal(dev.get(dG2), hb.grads)
    dXc = dev.put(hb.xccd)
    for a in range(3):
        # faces whose two cells both hold gradients
        start, ext = [2, 2, 2], [n - 4 for n in hb.cshape]
        ext[a] += 1
        F0 = np.random.default_rng(a).random(hb.F[a].shape)
        want = F0.copy()
        O.lib.pgo_diffflux(oc.ref, a, 1.0, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp),
                           O.arr(hb.grads), O.arr(hb.xccd), O.arr(hb.xf[a]), O.arr(hb.S[a]), O.arr(want),
                           O.box(start, ext))
        dF = dev.put(F0)
        native.check(L.pga_diffflux(nc.handle, a, 1.0, dev.view(dQv), dev.view(dSt), dev.view(dTr), dev.view(dH),
                                    dev.view(dG), dev.view(dXc), dev.view(dev.put(hb.xf[a])),
                                    dev.view(dev.put(hb.S[a])), dev.view(dF), native.box(start, ext), None),
                     "pga_diffflux")
        got = dev.get(dF)
        err = _rowerr(got, want)
        print(f"viscous flux {gas} axis {a}: {err:.3e}")
        assert np.array_equal(got, want)


@pytest.mark.parametrize("gas", ["air-tpg", "air-cpg"])
def test_tgv_navier_stokes_rk4_100_steps_match(dev, gas):
    """The Taylor-Green case as stated for it: Navier-Stokes at Re 1600 (constant mu, Pr 0.71),
    Mach 0.1, KEEP, alpha damping, RK4, 100 steps at 32^3; 1e-10 per conserved row."""
    n = 32
    L = (2 * np.pi,) * 3
    R, cp = 8314.46261815324 / 28.96, 1005.0
    gamma = cp / (cp - R)
    c0 = np.sqrt(gamma * R * 300.0)
    V0 = 0.1 * c0
    rho0 = 101325.0 / (R * 300.0)
    base = sp.AIR_CPG if gas == "air-cpg" else sp.AIR_NASA7
    species = {k: dict(v, mu0=sp.TGV_MU, kappa0=sp.TGV_KAPPA) for k, v in base.items()}

    def prims(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        q = np.zeros(x.shape[:3] + (5,))
        q[..., 0] = 101325.0 + rho0 * V0**2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
        q[..., 1] = V0 * np.sin(X) * np.cos(Y) * np.cos(Z)
        q[..., 2] = -V0 * np.cos(X) * np.sin(Y) * np.cos(Z)
        q[..., 4] = q[..., 0] / (rho0 * R)
        return q

    dt = 0.2 * (2 * np.pi / n) / (c0 + V0)
    mb, hb = _run_pair(gas, "constant", "kepaec", None, False, "rk4", (n,) * 3, 100, dt, prims, L, trans="constant",
                       species=species)
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"tgv navier-stokes {gas} rk4 100 steps: max row-normalised difference {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_viscous_binary_muscl_steps_match(dev, fused):
    """Two species with unequal Lewis numbers, MUSCL + HLLC and the viscous flux, tvd-rk3, 10 steps
    on a 40 x 20 x 12 block, by the fused stage and by the per-kernel sequence."""
    ncell = (40, 20, 12)
    L = (1.0, 0.5, 0.3)
    mb, hb = _run_pair("binary", "muscl", "hllc", "koren", False, "tvd-rk3", ncell, 10, 2e-6,
                       lambda x: sp.smooth_prims(x, 6, ns=2), L, fused=fused, trans="constant")
    assert mb.fused == fused
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"viscous binary 10 steps ({'fused' if fused else 'per-kernel'}): {err:.3e}")
    assert err < 1e-10


# --- cases the main library is not compiled for: a library built for each ------------------------


@pytest.mark.timeout(600)
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_five_species_steps_match_on_a_library_built_for_them(dev, fused):
    """Five calorically perfect species with the viscous flux, MUSCL + HLLC, tvd-rk3, 10 steps: a
    species count the main library has no kernels for, run by a library built for it (cached from
    the CPU suite's build, else built here) and matched to the oracle."""
    ncell = (40, 20, 12)
    L = (1.0, 0.5, 0.3)
    mb, hb = _run_pair("binary", "muscl", "hllc", "koren", False, "tvd-rk3", ncell, 10, 2e-6,
                       lambda x: sp.smooth_prims(x, 9, ns=5), L, fused=fused, trans="constant", species=sp.FIVE_CPG)
    assert mb.case.L is not native.lib()
    assert mb.fused == fused
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"five species 10 steps ({'fused' if fused else 'per-kernel'}): {err:.3e}")
    assert err < 1e-10


@pytest.mark.timeout(600)
def test_a_mechanism_of_its_own_structure_steps_on_a_library_built_for_it(dev, tmp_path):
    """The H2/O2 mechanism cut to its first 21 reactions (a structure the main library is not
    compiled for), the reacting case by the fused stage, 10 steps against the oracle."""
    n = 24
    L = (0.01,) * 3
    mb, hb = _run_pair("h2o2", "muscl", "hllc", "koren", True, "tvd-rk3", (n, n, n), 10, 1e-8,
                       lambda x: sp.reacting_flow(x, n, L), L, species=sp.trimmed_mechanism(tmp_path))
    assert mb.case.L is not native.lib()
    assert len(mb.sim.mixture.chemistry.records(mb.sim.mixture.species)[0]) == 21
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"21-reaction mechanism 10 steps: {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_reacting_navier_stokes_with_binary_diffusion_matches(dev, fused):
    """H2/O2-N2 with kinetic-theory transport, binary (mixture-averaged) diffusion and coupled
    chemistry, MUSCL + HLLC, tvd-rk3: 5 steps on 20 x 12 x 12 by the solver against the oracle."""
    n = 12
    L = (0.01,) * 3
    mb, hb = _run_pair("h2o2", "muscl", "hllc", "koren", True, "tvd-rk3", (20, n, n), 5, 1e-8,
                       lambda x: sp.reacting_flow(x, n, L), (0.01 * 20 / 12, 0.01, 0.01), fused=fused,
                       trans="kinetic-theory", mixture={"diffusion": "binary"})
    assert mb.fused == fused and mb.sim.mixture.diffusion.name == "binary"
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"reacting navier-stokes, binary diffusion, 5 steps ({'fused' if fused else 'per-kernel'}): {err:.3e}")
    assert err < 1e-10


# --- the cubic real gas ----------------------------------------------------------------------

REAL = ["o2n2-srk", "o2n2-pr"]


def _transcritical(xcc, ne, seed=7):
    """60 bar, O2 at 120 K below N2 at 300 K across a tanh layer in y, with a shear and a small
    seeded perturbation: both sides of the pseudo-critical point in one block."""
    rng = np.random.default_rng(seed)
    y = xcc[..., 1]
    Ly = xcc[..., 1].max() + xcc[..., 1].min()
    s = np.tanh((y - 0.5 * Ly) / (0.1 * Ly))
    q = np.zeros(xcc.shape[:3] + (ne,))
    q[..., 0] = 6e6
    q[..., 1] = 10.0 * s + 0.5 * np.sin(2 * np.pi * xcc[..., 0] / Ly + rng.uniform(0, 6.3))
    q[..., 2] = 0.5 * np.sin(2 * np.pi * xcc[..., 2] / Ly + rng.uniform(0, 6.3))
    q[..., 4] = 210.0 - 90.0 * s
    q[..., 5] = 0.5 * (1.0 - s) * 0.999 + 0.0005
    return q


@pytest.mark.parametrize("gas", REAL)
def test_real_gas_fromprims_and_fromQ(dev, gas):
    cfg, sim, nc, oc = _case(gas)
    hb = sp.HostBlock((70, 9, 5), 1, sim.ne, 2, jostle=0.2, nstate=9, lengths=(0.7e-3, 0.09e-3, 0.05e-3))
    q = _transcritical(hb.xcc, sim.ne)
    P = sp.to_dev(q)
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(P), O.arr(hb.regs[0]), O.arr(hb.state), full)
    dP, dQ, dS = dev.put(P), dev.put(np.zeros_like(hb.regs[0])), dev.put(np.zeros_like(hb.state))
    L = native.lib()
    native.check(L.pga_fromprims(nc.handle, dev.view(dP), dev.view(dQ), dev.view(dS), native.box((0, 0, 0), hb.cshape), None),
                 "pga_fromprims")
    err = max(_rowerr(dev.get(dQ), hb.regs[0]), _rowerr(dev.get(dS), hb.state))
    print(f"fromprims {gas}: {err:.3e}")
    assert err < 1e-12
    st = hb.state.copy()
    st[1] *= 1.02
    dS = dev.put(st)
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(st), full)
    arr = (native.Box * 1)(native.box((0, 0, 0), hb.cshape))
    native.check(L.pga_fromQ(nc.handle, dev.view(dev.put(hb.regs[0])), dev.view(dS), arr, 1, None), "pga_fromQ")
    err = _rowerr(dev.get(dS), st)
    print(f"fromQ {gas}: {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("gas", REAL)
def test_real_gas_muscl_faces_and_chung_transport(dev, gas):
    cfg = sp.config(gas, recon="muscl", primary="rusanov", limiter="koren", trans="chung")
    sim = get_simulator(cfg)
    nc = NativeCase(sim, Mi355xBackend(cfg))
    oc = O.OracleCase(sim.mixture, "muscl", "koren", "rusanov")
    hb = sp.HostBlock((67, 9, 6), sim.ng, sim.ne, 3, jostle=0.3, seed=5, nstate=9, lengths=(0.67e-3, 0.09e-3, 0.06e-3))
    q = _transcritical(hb.xcc, sim.ne)
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    ng, L = hb.ng, native.lib()
    dQv, dSt = dev.put(hb.regs[0]), dev.put(hb.state)
    for a in range(3):
        ext = list(hb.ncell)
        ext[a] += 1
        O.lib.pgo_flux(oc.ref, a, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.S[a]), O.arr(hb.F[a]), O.box((ng,) * 3, ext))
        F = dev.put(np.zeros_like(hb.F[a]))
        native.check(L.pga_flux(nc.handle, a, dev.view(dQv), dev.view(dSt), dev.view(dev.put(hb.S[a])), dev.view(F),
                                native.box((ng,) * 3, ext), None), "pga_flux")
        err = _rowerr(dev.get(F), hb.F[a])
        print(f"real-gas flux {gas} axis {a}: {err:.3e}")
        assert err < 1e-12
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp), full)
    dTr, dH = dev.put(np.zeros_like(hb.trstate)), dev.put(np.zeros_like(hb.hsp))
    boxes = (native.Box * 1)(native.box((0, 0, 0), hb.cshape))
    native.check(L.pga_transport(nc.handle, dev.view(dQv), dev.view(dSt), dev.view(dTr), dev.view(dH), boxes, 1, None),
                 "pga_transport")
    err = max(_rowerr(dev.get(dTr), hb.trstate), _rowerr(dev.get(dH), hb.hsp))
    print(f"chung transport {gas}: {err:.3e}")
    assert err < 1e-12


@pytest.mark.timeout(600)
def test_a_three_species_real_gas_steps_on_a_library_built_for_it(dev):
    """O2/N2 with 2% water at 60 bar, Peng-Robinson, MUSCL-Koren + Rusanov, tvd-rk3, 10 steps on a
    48 x 24 x 8 block with slip walls across y: the cubic real gas at a species count the main
    library has it not for (1 and 2), run by a library built for it and matched to the oracle."""
    ncell = (48, 24, 8)
    L = tuple(1e-5 * n for n in ncell)

    def prims(x):
        q2 = _transcritical(x, 6)
        q = np.zeros(q2.shape[:3] + (7,))
        q[..., :5] = q2[..., :5]
        q[..., 5] = 0.98 * q2[..., 5]
        q[..., 6] = 0.98 * (1.0 - q2[..., 5])
        return q

    mb, hb = _run_pair("o2n2-pr", "muscl", "rusanov", "koren", False, "tvd-rk3", ncell, 10, 2e-9, prims, L,
                       nstate=9, walls=(2, 3), species=["O2", "N2", "H2O"])
    assert mb.case.L is not native.lib() and mb.sim.ns == 3
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"three-species real gas 10 steps: {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_transcritical_mixing_layer_steps_match(dev, fused):
    """O2/N2 at 60 bar, Peng-Robinson with Chung transport, MUSCL-Koren + Rusanov, Navier-Stokes,
    tvd-rk3, 10 steps on a 48 x 24 x 8 block (10 um cells), periodic in x and z with adiabatic
    slip walls at both y ends, as the mixing layer is stated; 1e-10 per conserved row."""
    ncell = (48, 24, 8)
    L = tuple(1e-5 * n for n in ncell)
    mb, hb = _run_pair("o2n2-pr", "muscl", "rusanov", "koren", False, "tvd-rk3", ncell, 10, 2e-9,
                       lambda x: _transcritical(x, 6), L, fused=fused, trans="chung", nstate=9, walls=(2, 3))
    assert mb.fused == fused
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"transcritical 10 steps ({'fused' if fused else 'per-kernel'}): {err:.3e}")
    # the cubic's root and Chung's factors take only IEEE operations, dcbrt, elog and texp: the
    # same bits on both sides (with the libraries' cbrt, acos, cos, pow and exp, which round apart
    # by an ulp, this case measured 3.9e-10 after 10 steps: the pseudo-critical layer amplifies)
    assert err < 1e-10


def test_transcritical_rhs_matches(dev):
    """One right-hand side of the mixing layer (the halo EOS, the walls' mirror, transport,
    gradients, advective, viscous and wall fluxes, divergence) against the oracle's: 1e-12 per row."""
    import peregrine_amd as pg

    ncell = (48, 24, 8)
    L = tuple(1e-5 * n for n in ncell)
    cfg = sp.config("o2n2-pr", recon="muscl", primary="rusanov", limiter="koren", trans="chung", dt=2e-9)
    cfg["bcvals"]["wall"] = {"kind": "slp-adia-wall"}
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1 for n in ncell], lengths=L, periodic=[True, False, True],
                                bcnames={3: "wall", 4: "wall"})
    mb = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))
    sim = mb.sim
    oc = O.OracleCase(sim.mixture, "muscl", "koren", "rusanov")
    hb = sp.HostBlock(ncell, sim.ng, sim.ne, 3, lengths=L, nstate=9)
    q = _transcritical(hb.xcc, sim.ne)
    mb.set_prims([q])
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box((hb.ng,) * 3, hb.ncell))
    b = sp.oracle_block(hb, "tvd-rk3", False, viscous=True, alpha=sim.alpha, walls=(2, 3))
    O.lib.pgo_rhs(oc.ref, ctypes.byref(b), 0, 2)
    regs = mb.integ.stageregs(mb.regs)[0]
    mb.rhs(Q=regs.Q, dQ=regs.dQ)
    g = hb.ng
    got = mb.blocks[0].regs[regs.dQ].get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[2][hb.interior()])
    print(f"transcritical rhs: {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_slip_walls_steps_match(dev, fused):
    """Two species, MUSCL-Koren + HLLC, Euler, tvd-rk3 between adiabatic slip walls at both j ends
    and both k ends, periodic in i: 10 steps, by the fused stage and the per-kernel sequence."""
    ncell = (40, 20, 12)
    L = (1.0, 0.5, 0.3)
    mb, hb = _run_pair("binary", "muscl", "hllc", "koren", False, "tvd-rk3", ncell, 10, 2e-6,
                       lambda x: sp.smooth_prims(x, 6, ns=2), L, fused=fused, walls=(2, 3, 4, 5))
    assert mb.fused == fused
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"slip walls 10 steps ({'fused' if fused else 'per-kernel'}): {err:.3e}")
    assert err < 1e-10


def test_cflmax_matches_to_the_bit(dev):
    """The CFL reduction on a jostled grid with a random state, through the C ABI, and its work
    size for a box that needs more than one workgroup."""
    from peregrine_amd import grid

    cfg, sim, nc, oc = _case("binary", recon="muscl", primary="hllc", limiter="koren")
    ncell = (37, 9, 6)
    hb = sp.HostBlock(ncell, sim.ng, sim.ne, 1, lengths=(0.37, 0.09, 0.06), jostle=0.3, seed=2)
    rng = np.random.default_rng(21)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = 101325.0 * rng.uniform(0.5, 2.0, hb.cshape)
    q[..., 1:4] = rng.uniform(-80, 80, hb.cshape + (3,))
    q[..., 4] = rng.uniform(250.0, 900.0, hb.cshape)
    q[..., 5] = rng.uniform(0.0, 1.0, hb.cshape)
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    d = sp.to_dev(grid.dijk(hb.nodes, ncell))
    g = hb.ng
    want = np.zeros(3)
    O.lib.pgo_cflmax(O.arr(hb.regs[0]), O.arr(hb.state), O.arr(d), *(O.arr(s) for s in hb.S), O.box((g,) * 3, ncell),
                     want.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    L = native.lib()
    box = native.box((g,) * 3, ncell)
    n = L.pga_cflmax_worksize(box)
    assert n == 3 * ((37 * 9 * 6 + 255) // 256)
    work, out = dev.put(np.full(n, np.nan)), dev.put(np.zeros(3))
    native.check(L.pga_cflmax(nc.handle, dev.view(dev.put(hb.regs[0])), dev.view(dev.put(hb.state)), dev.view(dev.put(d)),
                              *(dev.view(dev.put(s)) for s in hb.S), box, work.data_ptr(), n, out.data_ptr(), None),
                 "pga_cflmax")
    got = dev.get(out)
    print("cflmax", got, want)
    assert np.array_equal(got, want)
    # too small a work array is refused before any launch, as a size out of range
    rc = L.pga_cflmax(nc.handle, dev.view(dev.put(hb.regs[0])), dev.view(dev.put(hb.state)), dev.view(dev.put(d)),
                      *(dev.view(dev.put(s)) for s in hb.S), box, work.data_ptr(), n - 1, out.data_ptr(), None)
    assert rc == -4 and b"work holds" in L.pga_last_error()


def test_cfl_controlled_steps_match(dev):
    """Three steps under the CFL controller: each dt from the device's reduction, the oracle
    stepping with dt from its own; the solutions agree to the bit."""
    import peregrine_amd as pg
    from peregrine_amd import grid

    ncell, lengths = (16, 12, 8), (1.0, 0.75, 0.5)
    cfg = sp.config("air-tpg", recon="muscl", primary="hllc", limiter="koren")
    cfg["simulation"]["controller"], cfg["simulation"]["cfl"], cfg["simulation"]["dt-max"] = "cfl", 0.4, 1.0
    cfg["simulation"]["niter"] = 3
    mb = pg.get_solver(cfg, pg.mesher.CubeMesher(blkdims=[n + 1 for n in ncell], lengths=lengths, periodic=[True] * 3),
                       pg.get_backend("mi355x", cfg))
    sim = mb.sim
    oc = O.OracleCase(sim.mixture, "muscl", "koren", "hllc")
    hb = sp.HostBlock(ncell, sim.ng, sim.ne, 3, lengths=lengths)
    q = sp.smooth_prims(hb.xcc, sim.ne)
    mb.set_prims([q])
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box((hb.ng,) * 3, hb.ncell))
    b = sp.oracle_block(hb, "tvd-rk3", False)
    d = sp.to_dev(grid.dijk(hb.nodes, ncell))
    g = hb.ng
    dts = []
    for _ in range(3):
        speeds = np.zeros(3)
        O.lib.pgo_cflmax(O.arr(hb.regs[0]), O.arr(hb.state), O.arr(d), *(O.arr(s) for s in hb.S),
                         O.box((g,) * 3, ncell), speeds.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        dt = min(0.4 / speeds[2], 1.0)
        dts.append(dt)
        O.lib.pgo_step(oc.ref, ctypes.byref(b), dt)
    mb.run()
    assert mb.nrt == 3 and mb.tme.get() == sum(dts)
    assert np.array_equal(sp.to_dev(mb.blocks[0].Q.get())[hb.interior()], hb.regs[0][hb.interior()])


def _hot_and_cold(sim, ncell, radicals=2e-3):
    """Stoichiometric H2-air at 10 bar: the low-i half at 400 K, the high-i half at 2500 K with
    radicals, so one step is stiff for half the cells and settled for the other."""
    names = sim.mixture.spnames
    q = np.zeros(ncell + (sim.ne,))
    q[..., 0] = 10e5
    q[..., 4] = 400.0
    q[ncell[0] // 2:, ..., 4] = 2500.0
    Y = np.zeros(sim.ns)
    for name, v in sp.H2_AIR_Y.items():
        Y[names.index(name)] = v
    hot = Y.copy()
    for name in ("H", "O", "OH"):
        hot[names.index(name)] = radicals
    q[..., 5:] = Y[:-1]
    q[ncell[0] // 2:, ..., 5:] = hot[:-1]
    return q


@pytest.mark.parametrize("bulk,bins", [(200, 1), (1, 1), (2, 8)], ids=["one-pass", "bulk1-bin1", "bulk2-bins8"])
@pytest.mark.parametrize("begin_dQ", [True, False], ids=["dQ", "rates"])
def test_substep_passes_match_the_one_pass_oracle(dev, bulk, bins, begin_dQ):
    """pga_substep through the C ABI over a stiff and a settled half, twice in a row (the second
    call's bins keyed by the first's counts): the mean rates and the counts are the oracle's
    one-pass form's to the bit, however the cells were listed."""
    cfg, sim, nc, oc = _case("h2o2", chem=True)
    ncell = (16, 6, 5)
    q = _hot_and_cold(sim, ncell)
    hb = sp.HostBlock(ncell, 0, sim.ne, 1)
    full = O.box((0, 0, 0), ncell)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    cap, eqtol, dt = 200, 1e-7, 1e-5
    rows = sim.ne if begin_dQ else sim.ns - 1
    want = np.zeros((rows,) + hb.regs[0].shape[1:])
    wcounts = np.zeros(int(np.prod(ncell)), dtype=np.int32)
    O.lib.pgo_substep(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(want), int(begin_dQ),
                      wcounts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap, eqtol, dt, full)
    import torch

    L = native.lib()
    opts = native.SubstepOpts(cap, bulk, bins, 0, eqtol)
    box = native.box((0, 0, 0), ncell)
    nlist = L.pga_substep_listsize(ctypes.byref(opts), box)
    assert nlist == 1 + bins + bins * int(np.prod(ncell))
    dQv, dS = dev.put(hb.regs[0]), dev.put(hb.state)
    out = dev.put(np.full_like(want, np.nan))
    counts = torch.zeros(int(np.prod(ncell)), dtype=torch.int32, device="cuda")
    lst = torch.zeros(nlist, dtype=torch.int32, device="cuda")
    ddt = dev.put(np.array([dt]))
    for call in range(2):
        native.check(L.pga_substep(nc.handle, ctypes.byref(opts), dev.view(dQv), dev.view(dS), dev.view(out),
                                   int(begin_dQ), counts.data_ptr(), counts.numel(), lst.data_ptr(), nlist,
                                   ddt.data_ptr(), box, None), "pga_substep")
        got, gcounts = dev.get(out), counts.cpu().numpy()
        print(f"call {call}: counts {np.bincount(gcounts)[:3]}... max {gcounts.max()}, listed {int(lst[0])}")
        assert np.array_equal(gcounts, wcounts)
        assert np.array_equal(got, want)
    # the hot half ran to many substeps, the cold half one; a compacted call listed the hot half
    per = wcounts.reshape(ncell[::-1])
    assert (per[..., : ncell[0] // 2] == 1).all() and (per[..., ncell[0] // 2:] > 2).all()
    if bulk < cap:
        assert int(lst[0]) == int(np.prod(ncell)) // 2


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_substepped_reacting_steps_match(dev, fused):
    """The reacting case with sub-stepped chemistry (compacted: bulk 2 of 50 substeps, 8 bins, the
    equilibrium exit on), 4 steps of tvd-rk3 by the solver against the oracle's one-pass form."""
    n = 12
    L = (0.01,) * 3

    def stiff(x):
        # the reacting flow with its high-x half 1000 K hotter and seeded with radicals, so those
        # cells run past the bulk pass's two substeps and are listed
        q = sp.reacting_flow(x, n, L)
        hot = x[..., 0] > 0.5 * x[..., 0].max()
        q[hot, 4] += 1000.0
        for sp_idx in (1, 2, 4):  # H, O, OH
            q[hot, 5 + sp_idx] = 2e-3
        return q

    mb, hb = _run_pair("h2o2", "muscl", "hllc", "koren", True, "tvd-rk3", (20, n, n), 4, 2e-7, stiff,
                       (0.01 * 20 / 12, 0.01, 0.01), fused=fused, substep=(50, 2, 8, 1e-7))
    assert mb.fused == fused
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    want = hb.regs[0][hb.interior()]
    err = _rowerr(got, want)
    counts = mb.blocks[0].substeps.cpu().numpy()
    print(f"sub-stepped 4 steps ({'fused' if fused else 'per-kernel'}): {err:.3e}, counts up to {counts.max()}")
    assert np.array_equal(counts, hb.counts)
    assert counts.max() > 2
    assert err < 1e-10


WALL_CASES = [
    ("slp-adia-wall", {}), ("slp-isot-wall", {"T": 350.0}), ("no-slp-adia-wall", {}),
    ("no-slp-isot-wall", {"T": 350.0}), ("mov-adia-wall", {"u": 20.0, "v": 0.0, "w": -4.0}),
    ("mov-isot-wall", {"u": 20.0, "v": 0.0, "w": -4.0, "T": 350.0}),
]


@pytest.mark.parametrize("kind,values", WALL_CASES, ids=[k for k, _ in WALL_CASES])
def test_wall_halos_and_fluxes_match(dev, kind, values):
    """Each wall kind's mirror halo and wall flux through the C ABI against the oracle, on both sides
    of every axis of a jostled viscous block (kinetic-theory transport): to the bit."""
    cfg = sp.config("h2o2", trans="kinetic-theory")
    sim = get_simulator(cfg)
    nc = NativeCase(sim, Mi355xBackend(cfg))
    oc = O.OracleCase(sim.mixture)
    hb = sp.HostBlock((13, 9, 6), 1, sim.ne, 1, jostle=0.3, seed=4)
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=sim.ns)
    q[..., 5:] *= 0.1
    _prepared(sim, oc, hb, q)
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp),
                        O.box((0, 0, 0), hb.cshape))
    w = native.Wall()
    w.kind = native.WALL_KINDS.index(kind)
    for d, v in enumerate("uvw"):
        w.u[d] = values.get(v, 0.0)
    w.T = values.get("T", 0.0)
    wall = O.dbls([w.u[0], w.u[1], w.u[2]])
    L = native.lib()
    for axis in range(3):
        for side in (0, 1):
            Q, st = hb.regs[0].copy(), hb.state.copy()
            dQv, dSt = dev.put(Q), dev.put(st)
            O.lib.pgo_wall_mirror(oc.ref, axis, side, w.kind, wall, w.T, O.ints(hb.ncell), hb.ng, st.shape[0], O.arr(Q),
                                  O.arr(st), O.arr(hb.S[axis]))
            native.check(L.pga_wall_mirror(nc.handle, axis, side, ctypes.byref(w), native.ints(hb.ncell), hb.ng,
                                           dev.view(dQv), dev.view(dSt), dev.view(dev.put(hb.S[axis])), None),
                         "pga_wall_mirror")
            assert np.array_equal(dev.get(dQv), Q), (kind, axis, side)
            assert np.array_equal(dev.get(dSt), st), (kind, axis, side)
            F0 = np.random.default_rng(axis).random(hb.F[axis].shape)
            want = F0.copy()
            O.lib.pgo_wall_flux(oc.ref, axis, side, w.kind, wall, w.T, O.ints(hb.ncell), hb.ng, 1, O.arr(Q),
                                O.arr(st), O.arr(hb.trstate), O.arr(hb.xccd), O.arr(hb.xf[axis]), O.arr(hb.S[axis]),
                                O.arr(want))
            dF = dev.put(F0)
            native.check(L.pga_wall_flux(nc.handle, axis, side, ctypes.byref(w), native.ints(hb.ncell), hb.ng,
                                         dev.view(dQv), dev.view(dSt), dev.view(dev.put(hb.trstate)),
                                         dev.view(dev.put(hb.xccd)), dev.view(dev.put(hb.xf[axis])),
                                         dev.view(dev.put(hb.S[axis])), dev.view(dF), None), "pga_wall_flux")
            assert np.array_equal(dev.get(dF), want), (kind, axis, side)


@pytest.mark.parametrize("fused,slabs,kwalls", [(True, 1, False), (False, 1, False), (True, 3, True)],
                         ids=["fused", "per-kernel", "fused-slabs-kwalls"])
def test_channel_between_held_walls_steps_match(dev, fused, slabs, kwalls):
    """A viscous binary mixture between an isothermal no-slip wall (j low, 350 K) and a moving
    adiabatic wall (j high, 20 m/s along x), periodic in i and k (or walled in k by slip walls, the
    fused stage in three k-slabs), MUSCL + HLLC, tvd-rk3, 10 steps by the solver against the oracle."""
    ncell = (16, 12, 8)
    L = (0.4, 0.3, 0.2)
    walls = {2: ("no-slp-isot-wall", {"T": 350.0}), 3: ("mov-adia-wall", {"u": 20.0, "v": 0.0, "w": 0.0})}
    if kwalls:
        walls.update({4: ("slp-adia-wall", {}), 5: ("no-slp-adia-wall", {})})
    mb, hb = _run_pair("binary", "muscl", "hllc", "koren", False, "tvd-rk3", ncell, 10, 2e-6,
                       lambda x: sp.smooth_prims(x, 6, ns=2), L, fused=fused, trans="constant", walls=walls,
                       slabs=slabs)
    assert mb.fused == fused
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"channel 10 steps ({'fused' if fused else 'per-kernel'}): {err:.3e}")
    assert err < 1e-10


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_sods_tube_steps_match_and_meet_the_exact_solution(dev, fused):
    """Sod's tube (Toro test 1 as the reference numbers it 0), MUSCL-Koren + HLLC, on a walled line of
    200 cells one cell thick, 2000 steps of tvd-rk3 by the solver: the oracle's line to the bit, and
    the reference's measured mean density error over the fan (tests/simulator/test_flux.py:87)."""
    import peregrine_amd as pg
    import test_oracle_riemann as tr

    case, scheme = 0, "muscl-koren-hllc"
    x_o, rho_o, u_o, p_o, r, _ = tr.tube(case, scheme)
    c = pg.Config()
    c["simulation"]["simulator"] = "euler"
    c["mixture"]["species"], c["mixture"]["eos"] = tr.AIR, "cpg"
    a = c["adv-flux"]
    a["primary"], a["recon"], a["limiter"] = "hllc", "muscl", "koren"
    c["bcvals"]["walls"] = {"kind": "slp-adia-wall"}
    c["backend-mi355x"]["fused-stage"] = fused
    c.validate()
    n, L = 200, 1.0
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1, 2, 2], lengths=[L, 0.1, 0.1], periodic=[False] * 3,
                                bcnames=dict.fromkeys(range(1, 7), "walls"))
    mb = pg.get_solver(c, mesh, pg.get_backend("mi355x", c))
    blk = mb.blocks[0]
    x = blk.xcc.get()[..., 0] + (r.x0 - 0.5 * L)
    left = x <= r.x0
    q = np.zeros(blk.cellshape + (mb.sim.ne,))
    q[..., 0] = np.where(left, r.pL, r.pR)
    q[..., 1] = np.where(left, r.uL, r.uR)
    q[..., 4] = np.where(left, r.TL, r.TR)
    mb.set_prims([q])
    for _ in range(int(round(r.t / r.dt))):
        mb.step(r.dt)
    g = mb.ng
    Q = blk.Q.get()[g:-g, g, g]
    rho = Q[:, 0]
    diff = np.abs(rho - rho_o)
    print(f"sod device vs oracle: max |drho| {diff.max():.3e} at {int(diff.argmax())}, cells differing {int((diff > 0).sum())}")
    assert np.array_equal(rho, rho_o) and np.array_equal(Q[:, 1] / Q[:, 0], u_o)
    rho_err, _, _ = tr.errors(x_o, rho, u_o, p_o, r)
    print(f"sod on the device ({'fused' if fused else 'per-kernel'}): rho error {rho_err:.4e}")
    assert rho_err == pytest.approx(tr.TUBES[case]["rho"][scheme], rel=1e-3)


def test_the_viscous_sponge_matches(dev):
    """pga_viscsponge against the oracle on a jostled block's centres and an oblique line, to the
    bit; and the walled channel stepped with a sponge (the viscosity raised eightfold along a line
    across it) against the oracle's steps, which a sponge left out would not match."""
    rng = np.random.default_rng(11)
    hb = sp.HostBlock((12, 9, 7), 2, 6, 1, jostle=0.3, seed=2)
    tr = rng.uniform(1e-5, 2e-5, hb.cshape + (4,))
    line = O.dbls([0.6, 0.64, 0.48, 0.1, 0.7, 8.0])
    want = sp.to_dev(tr)
    box = O.box((1, 1, 1), tuple(n - 2 for n in hb.cshape))
    O.lib.pgo_viscsponge(O.arr(hb.xccd), O.arr(want), line, box)
    dtr = dev.put(sp.to_dev(tr))
    native.check(native.lib().pga_viscsponge(dev.view(dev.put(hb.xccd)), dev.view(dtr), line,
                                             (native.Box * 1)(native.box((1, 1, 1), tuple(n - 2 for n in hb.cshape))),
                                             1, None), "pga_viscsponge")
    assert np.array_equal(dev.get(dtr), want)
    assert not np.array_equal(want, sp.to_dev(tr))

    ncell = (16, 12, 8)
    L = (0.4, 0.3, 0.2)
    walls = {2: ("no-slp-isot-wall", {"T": 350.0}), 3: ("mov-adia-wall", {"u": 20.0, "v": 0.0, "w": 0.0})}
    runs = {}
    for name, sponge in (("sponge", {"origin": [0.1, 0.0, 0.0], "ending": [0.3, 0.1, 0.05], "multiplier": 8.0}),
                         ("none", None)):
        mb, hb = _run_pair("binary", "muscl", "hllc", "koren", False, "tvd-rk3", ncell, 10, 2e-6,
                           lambda x: sp.smooth_prims(x, 6, ns=2), L, trans="constant", walls=walls, sponge=sponge)
        g = hb.ng
        got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
        err = _rowerr(got, hb.regs[0][hb.interior()])
        print(f"channel with {name}: {err:.3e}")
        assert err < 1e-10
        runs[name] = got
    assert _rowerr(runs["sponge"], runs["none"]) > 1e-8


def test_taylor_green_at_its_full_size_navier_stokes_rk4_100_steps(dev):
    """The Taylor-Green case as the reference's CPU-runnable benchmark states it: one 32^3 periodic
    block on [0, 2 pi]^3, thermally perfect air, Mach 0.1, Re 1600 (mu = rho0 V0 L / 1600 with
    L = 1, kappa = mu cp / 0.71), KEEP, Navier-Stokes with alpha damping, RK4, dt = 0.2 dx / (c0 + V0),
    100 steps: the device against the oracle within 1e-10 of each conserved row's largest value."""
    n = 32
    L = (2 * np.pi,) * 3
    R = 8314.46261815324 / sp.AIR_NASA7["Air"]["MW"]
    cp = 1005.0
    gamma = cp / (cp - R)
    c0 = np.sqrt(gamma * R * 300.0)
    V0 = 0.1 * c0
    rho0 = 101325.0 / (R * 300.0)
    mu = rho0 * V0 * 1.0 / 1600.0
    species = {"Air": {**sp.AIR_NASA7["Air"], "mu0": mu, "kappa0": mu * cp / 0.71}}

    def prims(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        q = np.zeros(x.shape[:3] + (5,))
        q[..., 0] = 101325.0 + rho0 * V0**2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
        q[..., 1] = V0 * np.sin(X) * np.cos(Y) * np.cos(Z)
        q[..., 2] = -V0 * np.cos(X) * np.sin(Y) * np.cos(Z)
        q[..., 4] = q[..., 0] / (rho0 * R)
        return q

    dt = 0.2 * (2 * np.pi / n) / (c0 + V0)
    mb, hb = _run_pair("air-tpg", "constant", "kepaec", None, False, "rk4", (n,) * 3, 100, dt, prims, L,
                       trans="constant", species=species)
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    err = _rowerr(got, hb.regs[0][hb.interior()])
    print(f"taylor-green 32^3 navier-stokes rk4 100 steps: {err:.3e}")
    assert err < 1e-10
