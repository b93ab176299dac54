"""Implicit chemistry on the MI355X through the C ABI against the oracle's FP64 form.

The device factors each cell's matrix I - h J in FP32 on the matrix cores (16 cells per wave) and
refines every solve against the FP64 matrix with FP64 residuals; the oracle solves the same FP64
systems by an FP64 LU. The bar is the survey's for the implicit form: 1e-8 relative on species and
temperature against the FP64 form. The measured gap is printed per result."""

import ctypes

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd import native
from test_gpu_parity import _case, _hot_and_cold, _rowerr, _run_pair, dev  # noqa: F401

pytestmark = pytest.mark.gpu


def _systems(n, nv, seed):
    """n systems of nv unknowns, condition about 1e3 to 1e4, scaled rows as a chemical matrix has."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, nv, nv))
    A += nv * np.eye(nv)
    A *= np.exp(rng.uniform(-3.0, 3.0, size=(n, nv, 1)))
    b = rng.standard_normal((n, nv))
    return A, b


@pytest.mark.parametrize("nv", [1, 3, 4, 5, 9, 10, 12])
def test_batched_lu_solve_refines_to_fp64(dev, nv):
    """pga_lu_solve on 1000 systems (a partial last group of 16): one FP32 solve is good to ~1e-6,
    the refined solve to the FP64 solve's accuracy, every row's residual within the tolerance."""
    import torch

    n = 1000
    A, b = _systems(n, nv, nv)
    want = np.linalg.solve(A, b[..., None])[..., 0]
    L = native.lib()
    work = torch.zeros(int(L.pga_lu_worksize(n, nv)), dtype=torch.float64, device="cuda")
    dA, db = torch.from_numpy(A).cuda(), torch.from_numpy(b).cuda()
    errs = {}
    for sweeps, tol in ((1, 0.0), (8, 1e-14)):
        dx = torch.full((n, nv), float("nan"), dtype=torch.float64, device="cuda")
        cnt = torch.zeros(n, dtype=torch.int32, device="cuda")
        native.check(L.pga_lu_solve(n, nv, dA.data_ptr(), db.data_ptr(), dx.data_ptr(), sweeps, tol, cnt.data_ptr(),
                                    work.data_ptr(), work.numel(), None), "pga_lu_solve")
        x = dx.cpu().numpy()
        errs[sweeps] = (np.abs(x - want).max(axis=1) / np.abs(want).max(axis=1)).max()
        counts = cnt.cpu().numpy()
    # the device stops at a residual within 1e-14 of each row's scale in its own order of sums; numpy's
    # residual of the same x, summed in another order, holds it to a few times that
    r = b - np.einsum("nij,nj->ni", A, x)
    scale = np.abs(b) + np.einsum("nij,nj->ni", np.abs(A), np.abs(x))
    print(f"nv {nv}: one FP32 solve {errs[1]:.2e}, refined {errs[8]:.2e}, corrections {np.bincount(counts)}, "
          f"backward error {(np.abs(r) / scale).max():.1e}")
    assert (np.abs(r) <= 1e-13 * scale).all()
    assert errs[8] <= 1e-12
    assert (counts >= 1).all() and (counts <= 8).all()
    if nv > 1:
        assert errs[1] > 1e-10 and (counts >= 2).any()


def _device_y(q_regs, sim):
    """The part's start as the device reads it from Q and the state: the carried mass fractions
    clamped to [0, 1] and scaled when they sum past one, the last species the rest (massfrac)."""
    Q, T = q_regs
    rinv = 1.0 / Q[0]
    Y, last, total = [], np.ones_like(Q[0]), np.zeros_like(Q[0])
    for n in range(sim.ns - 1):
        Yn = np.maximum(np.minimum(Q[5 + n] * rinv, 1.0), 0.0)
        Y.append(Yn)
        last, total = last - Yn, total + Yn
    scale = 1.0 / np.maximum(total, 1.0)
    return np.stack([v * scale for v in Y] + [np.maximum(last, 0.0), T]), Q[0]


def test_the_matrices_are_the_oracles(dev):
    """The FP64 matrices I - h J the device assembles (a species column's rates from the cache of the
    T-only values, the T column and f(y) whole) read back from the work array, against
    pgo_implicit_matrix at each cell's y rounded to FP32: within 1e-12 of each column's largest entry
    (the same rates, the same steps). A cached value taken for the wrong reaction or species, or at
    the wrong T, moves a column by far more."""
    import torch

    cfg, sim, nc, oc = _case("h2o2", chem=True)
    ncell = (20, 6, 5)
    q = _hot_and_cold(sim, ncell, radicals=1e-3)
    hb = sp.HostBlock(ncell, 0, sim.ne, 1)
    full = O.box((0, 0, 0), ncell)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    L = native.lib()
    box = native.box((0, 0, 0), ncell)
    cells, nv, dt = int(np.prod(ncell)), sim.ns + 1, 1e-7
    groups = -(-cells // 16)
    nwork = int(L.pga_implicit_worksize(nc.handle, box))
    work = torch.zeros(nwork, dtype=torch.float64, device="cuda")
    sweeps = torch.zeros(cells, dtype=torch.int32, device="cuda")
    out = dev.put(np.zeros((sim.ns - 1,) + hb.regs[0].shape[1:]))
    ddt = dev.put(np.array([dt]))
    opts = native.ImplicitOpts(1, 8, 1e-13)
    native.check(L.pga_implicit(nc.handle, ctypes.byref(opts), dev.view(dev.put(hb.regs[0])), dev.view(dev.put(hb.state)),
                                dev.view(out), 0, sweeps.data_ptr(), cells, work.data_ptr(), nwork, ddt.data_ptr(), box,
                                None), "pga_implicit")
    w = work.cpu().numpy()
    A = w[:groups * nv * nv * 16].reshape(groups, nv, nv, 16).transpose(0, 3, 1, 2).reshape(-1, nv, nv)[:cells]
    y, rho = _device_y((hb.regs[0].reshape(sim.ne, -1), hb.state[1].reshape(-1)), sim)
    worst = 0.0
    for c in range(cells):
        want = oc.implicit_matrix(rho[c], y[:, c].astype(np.float32).astype(np.float64), dt)
        scale = np.abs(want).max(axis=0)
        worst = max(worst, float((np.abs(A[c] - want).max(axis=0) / scale).max()))
    print(f"device matrices against the oracle's: {worst:.2e} of a column's largest entry")
    assert worst <= 1e-12


@pytest.mark.parametrize("parts", [1, 3])
@pytest.mark.parametrize("begin_dQ", [True, False], ids=["dQ", "rates"])
def test_implicit_source_matches_the_fp64_form(dev, parts, begin_dQ):
    """pga_implicit over a block of a settled (400 K) and a stiff (2500 K, radicals) half, twice in a
    row: the mean rates against pgo_implicit's FP64 form. One part differs only by the solves (the
    refined FP32-factored one against the FP64 LU): within 1e-11 of each row's largest (measured
    1.3e-13). Later parts start from states that differ in their last bits, whose matrices (taken at
    the FP32 rounding of the state) can then round apart: within the implicit form's bar, 1e-8
    (measured 3.8e-10 with three parts)."""
    import torch

    cfg, sim, nc, oc = _case("h2o2", chem=True)
    ncell = (20, 6, 5)
    q = _hot_and_cold(sim, ncell, radicals=1e-3)
    hb = sp.HostBlock(ncell, 0, sim.ne, 1)
    full = O.box((0, 0, 0), ncell)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    dt = 1e-7
    rows = sim.ne if begin_dQ else sim.ns - 1
    want = np.zeros((rows,) + hb.regs[0].shape[1:])
    O.lib.pgo_implicit(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(want), int(begin_dQ), parts, dt, full)
    L = native.lib()
    box = native.box((0, 0, 0), ncell)
    opts = native.ImplicitOpts(parts, 8, 1e-13)
    nwork = int(L.pga_implicit_worksize(nc.handle, box))
    work = torch.zeros(nwork, dtype=torch.float64, device="cuda")
    cells = int(np.prod(ncell))
    sweeps = torch.zeros(cells, dtype=torch.int32, device="cuda")
    dQv, dS = dev.put(hb.regs[0]), dev.put(hb.state)
    ddt = dev.put(np.array([dt]))
    for call in range(2):
        out = dev.put(np.full_like(want, np.nan))
        native.check(L.pga_implicit(nc.handle, ctypes.byref(opts), dev.view(dQv), dev.view(dS), dev.view(out),
                                    int(begin_dQ), sweeps.data_ptr(), cells, work.data_ptr(), nwork, ddt.data_ptr(),
                                    box, None), "pga_implicit")
        got = dev.get(out)
        counts = sweeps.cpu().numpy()
        err = _rowerr(got[-(sim.ns - 1):], want[-(sim.ns - 1):])
        print(f"parts {parts} call {call}: rates within {err:.2e} of the FP64 form, corrections {np.bincount(counts)}")
        assert np.isfinite(got).all()
        if begin_dQ:
            assert not got[:5].any()
        assert err <= (1e-11 if parts == 1 else 1e-8)
        assert (counts >= parts).all() and (counts < 1000).all()


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "per-kernel"])
def test_implicit_reacting_steps_match(dev, fused):
    """The reacting case with implicit chemistry (two parts per step), 4 steps of tvd-rk3 by the solver
    against the oracle's FP64 form: species and T within 1e-8 of each row's largest."""
    n = 12
    L = (0.01,) * 3

    def stiff(x):
        q = sp.reacting_flow(x, n, L)
        hot = x[..., 0] > 0.5 * x[..., 0].max()
        q[hot, 4] += 1000.0
        for sp_idx in (1, 2, 4):  # H, O, OH
            q[hot, 5 + sp_idx] = 2e-3
        return q

    mb, hb = _run_pair("h2o2", "muscl", "hllc", "koren", True, "tvd-rk3", (20, n, n), 4, 2e-8, stiff,
                       (0.01 * 20 / 12, 0.01, 0.01), fused=fused, implicit=(2, 8, 1e-13))
    assert mb.fused == fused
    g = hb.ng
    got = mb.blocks[0].Q.get()[g:-g, g:-g, g:-g].transpose(3, 2, 1, 0)
    want = hb.regs[0][hb.interior()]
    err = _rowerr(got, want)
    Tgot = mb.blocks[0].state.get()[g:-g, g:-g, g:-g, 1]
    Twant = hb.state[1][hb.interior()[1:]].transpose(2, 1, 0)
    errT = np.abs(Tgot - Twant).max() / np.abs(Twant).max()
    counts = mb.blocks[0].sweeps.cpu().numpy()
    print(f"implicit 4 steps ({'fused' if fused else 'per-kernel'}): Q {err:.3e}, T {errT:.3e}, "
          f"corrections {np.bincount(counts)}")
    assert err < 1e-8 and errT < 1e-8
    assert (counts < 1000).all()


def test_an_uneven_batch_with_implicit_chemistry_steps_as_its_cases_one_by_one(gpu):
    """Three periodic H2/O2 boxes of different sizes, hot and seeded with radicals, stepped with implicit
    chemistry by one solver whose phases are each one launch over the three (pga_batch_implicit among
    them), replayed from one graph per step: each box to the bit of the box stepped alone."""
    import peregrine_amd as pg

    dims = [(16, 12, 20), (24, 16, 12), (12, 20, 16)]
    L = (0.01,) * 3

    def hot(b):
        q = sp.h2_prims(b.xcc.get(), b.ncell, lengths=L)
        q[..., 4] += 800.0
        q[..., 5 + 1] = 1e-3
        return q

    def solver(dlist, graphs):
        cfg = sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, dt=2e-8)
        cfg["chem-integration"]["kind"] = "implicit"
        cfg["backend-mi355x"]["graphs"] = graphs
        mb = pg.get_solver(cfg, pg.mesher.BatchMesher(dlist, lengths=L), pg.get_backend("mi355x", cfg))
        mb.set_prims([hot(b) for b in mb.blocks])
        return mb

    together = solver(dims, True)
    alone = [solver([d], False) for d in dims]
    assert together.batched and not alone[0].batched
    for _ in range(3):
        together.step(2e-8)
        for mb in alone:
            mb.step(2e-8)
    assert together.replays == 2
    for blk, mb in zip(together.blocks, alone):
        assert np.array_equal(blk.Q.get(), mb.blocks[0].Q.get())
        assert np.array_equal(blk.state.get(), mb.blocks[0].state.get())
