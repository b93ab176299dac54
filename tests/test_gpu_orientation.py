"""Halo exchange across reoriented block interfaces on the device, after the reference's own
test (tests/communication/test_orientation.py:63-126): a pair of blocks, the second relabelled by
each of the 24 orientations a right-handed grid can meet another in, holding the same physical
field; after the exchange every array equals the aligned pair's, relabelled. The mapping itself is
pinned on the host by cell centres (test_host.py); here the device copies (pga_copy_box, pga_halo
with its EOS) are held to it bit for bit, and a few steps of a reoriented pair to the aligned
pair's within a rounding bound (the face area vectors and the divergence sum of a relabelled
block round differently)."""

import numpy as np
import pytest

import support as sp

pytestmark = pytest.mark.gpu

# the halo depth each reconstruction exchanges
RECONS = {"constant": ("kepaec", None), "muscl": ("hllc", "koren")}


def _solver(S, recon, gas="air-tpg", chem=False, fused=True, blkdims=(6, 3, 2), lengths=(2.0, 1.0, 1.0)):
    import peregrine_amd as pg

    primary, limiter = RECONS[recon]
    cfg = sp.config(gas, recon=recon, primary=primary, limiter=limiter, chem=chem, dt=1e-8)
    cfg["bcvals"]["wall"] = {"kind": "slp-adia-wall"}
    cfg["backend-mi355x"]["fused-stage"] = fused
    mesh = sp.ReorientedPair(S, mbdims=[2, 1, 1], blkdims=list(blkdims), lengths=list(lengths), defaultbc="wall")
    return pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))


def _traded(S, recon, seed):
    """Random values in the aligned layout (block 1 relabelled), then the Q exchange."""
    mb = _solver(S, recon)
    rng = np.random.default_rng(seed)
    shape = mb.blocks[0].Q.shape
    for blk in mb.blocks:
        values = rng.random(shape)
        blk.Q.set(values if blk.nblki == 0 else sp.reorient(values, S))
    mb.exchange(0)
    mb.synchronize()
    return mb, [blk.Q.get() for blk in mb.blocks]


_aligned = {}


@pytest.mark.parametrize("recon", list(RECONS))
@pytest.mark.parametrize("S", sp.orientations())
def test_the_exchange_places_a_reoriented_neighbours_cells(gpu, S, recon):
    seed = 20260901
    if recon not in _aligned:
        mb, ref = _traded("123", recon, seed)
        ng = mb.ng
        # a trade that wrote nothing would pass below: the landing itself first
        for g in range(ng):
            assert np.array_equal(ref[0][-ng + g, ng:-ng, ng:-ng], ref[1][ng + g, ng:-ng, ng:-ng])
            assert np.array_equal(ref[1][g, ng:-ng, ng:-ng], ref[0][-2 * ng + g, ng:-ng, ng:-ng])
        _aligned[recon] = ref
    ref = _aligned[recon]
    _, got = _traded(S, recon, seed)
    assert np.array_equal(got[0], ref[0]), (S, "block 0")
    assert np.array_equal(got[1], sp.reorient(ref[1], S)), (S, "block 1")


def _halo_passed(S):
    """The aligned pair's smooth state laid out in each block; the exchanged halo slabs' Q cleared
    and their state but T spoiled; then the fused stage's halo pass (the Q copy and the EOS over
    the halo, its Newton from the T there). Interior and exchanged slab, in the aligned layout."""
    ref = _solver("123", "muscl")
    prims = [sp.smooth_prims(blk.xcc.get(), ref.sim.ne) for blk in ref.blocks]
    mb = _solver(S, "muscl")
    mb.set_prims([prims[0], sp.reorient(prims[1], S)])
    mb.synchronize()
    g = mb.ng
    # block 0 meets block 1 at its i-high interface, block 1 (aligned) at its i-low one
    slabs = (np.s_[-g:, g:-g, g:-g], np.s_[:g, g:-g, g:-g])
    kept = (np.s_[g:, g:-g, g:-g], np.s_[:-g, g:-g, g:-g])
    there = (lambda a: a, lambda a: sp.reorient_back(a, S))
    back = (lambda a: a, lambda a: np.ascontiguousarray(sp.reorient(a, S)))
    for blk in mb.blocks:
        n = blk.nblki
        Q, st = there[n](blk.Q.get()).copy(), there[n](blk.state.get()).copy()
        Q[slabs[n]] = 0.0
        T = st[..., 1].copy()
        st[slabs[n]] = -1.0
        st[..., 1] = T
        blk.Q.set(back[n](Q))
        blk.state.set(back[n](st))
    mb._halo_pass(0)
    mb.synchronize()
    return [(there[n](blk.Q.get())[kept[n]], there[n](blk.state.get())[kept[n]]) for n, blk in enumerate(mb.blocks)]


@pytest.mark.parametrize("S", sp.orientations())
def test_the_halo_pass_places_and_closes_a_reoriented_neighbours_cells(gpu, S):
    want = _halo_passed("123")
    got = _halo_passed(S)
    for n in range(2):
        assert np.array_equal(got[n][0], want[n][0]), (S, n, "Q")
        assert np.array_equal(got[n][1], want[n][1]), (S, n, "state")
        assert (got[n][1][..., 0] > 0).all()


# the bound on the relabelled pair: a calorically perfect gas rounds the same physics to within a
# few ulps per step; the thermally perfect gas's e(T) in ln T polynomials turns a last-bit change
# of a face's T into about 1e-11 of e, so the reacting H2/O2 pair is held to 1e-9
CASES = {"air-cpg": (False, 1e-12), "h2o2": (True, 1e-9)}


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
@pytest.mark.parametrize("S", sp.orientations()[::5])
@pytest.mark.parametrize("gas", list(CASES))
def test_a_reoriented_pair_steps_as_the_aligned_pair(gpu, gas, S, fused):
    """MUSCL-Koren + HLLC, tvd-rk3, 4 steps of a walled pair of 8 x 6 x 5-cell blocks (H2/O2 with
    coupled chemistry, or calorically perfect air): the relabelled pair against the aligned one
    within the gas's bound of each component's largest value (the identity, "123", to the bit)."""
    chem, bound = CASES[gas]
    runs = []
    for digits in ("123", S):
        mb = _solver(digits, "muscl", gas=gas, chem=chem, fused=fused, blkdims=(9, 7, 6), lengths=(0.02, 0.006, 0.005))
        if not runs:
            prims = [sp.smooth_prims(blk.xcc.get(), mb.sim.ne, ns=mb.sim.ns) for blk in mb.blocks]
            if chem:
                for q in prims:
                    q[..., 4] += 900.0
        mb.set_prims([prims[0], sp.reorient(prims[1], digits)])
        for _ in range(4):
            mb.step(1e-8)
        g = mb.ng
        runs.append([blk.Q.get()[g:-g, g:-g, g:-g] for blk in mb.blocks])
    (a0, a1), (b0, b1) = runs
    b1 = sp.reorient_back(b1, S)
    worst = 0.0
    for want, got in ((a0, b0), (a1, b1)):
        if S == "123":
            assert np.array_equal(got, want)
        scale = np.abs(want).reshape(-1, want.shape[-1]).max(axis=0)
        worst = max(worst, (np.abs(got - want).reshape(-1, want.shape[-1]).max(axis=0) / scale).max())
    print(f"reoriented {gas} {S} ({'fused' if fused else 'per-kernel'}): {worst:.3e}")
    assert worst <= bound, S
