"""The allocation-free CFL reduction (pga_cflmax_noalloc, partials in the library's own device array)
against pga_cflmax with a caller's work array, through the C ABI: two calls in a row, the second
on a smaller box (so the array still holds the first call's partials beyond the second's), both
equal to the bit; the oracle's reduction pins the first. A box beyond the array is refused before
any launch."""

import ctypes

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd import grid, native
from test_gpu_parity import _case, dev  # noqa: F401

pytestmark = pytest.mark.gpu


def _block(sim, oc, ncell, seed):
    hb = sp.HostBlock(ncell, sim.ng, sim.ne, 1, lengths=tuple(0.01 * n for n in ncell), jostle=0.3, seed=seed)
    rng = np.random.default_rng(seed)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = 101325.0 * rng.uniform(0.5, 2.0, hb.cshape)
    q[..., 1:4] = rng.uniform(-80, 80, hb.cshape + (3,))
    q[..., 4] = rng.uniform(250.0, 900.0, hb.cshape)
    q[..., 5] = rng.uniform(0.0, 1.0, hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))
    return hb, sp.to_dev(grid.dijk(hb.nodes, ncell))


def test_noalloc_reduction_equals_the_work_array_one(dev):
    cfg, sim, nc, oc = _case("binary", recon="muscl", primary="hllc", limiter="koren")
    L = native.lib()
    for call, ncell in enumerate(((61, 37, 9), (13, 7, 5))):
        hb, d = _block(sim, oc, ncell, seed=3 + call)
        g = hb.ng
        views = [dev.view(dev.put(a)) for a in (hb.regs[0], hb.state, d, *hb.S)]
        box = native.box((g,) * 3, ncell)
        n = L.pga_cflmax_worksize(box)
        work, want, got = dev.put(np.full(n, np.nan)), dev.put(np.zeros(3)), dev.put(np.full(3, np.nan))
        native.check(L.pga_cflmax(nc.handle, *views, box, work.data_ptr(), n, want.data_ptr(), None), "pga_cflmax")
        native.check(L.pga_cflmax_noalloc(nc.handle, *views, box, got.data_ptr(), None), "pga_cflmax_noalloc")
        oracle = np.zeros(3)
        O.lib.pgo_cflmax(O.arr(hb.regs[0]), O.arr(hb.state), O.arr(d), *(O.arr(s) for s in hb.S),
                         O.box((g,) * 3, ncell), oracle.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        print(f"call {call} {ncell}: {dev.get(got)} {dev.get(want)} {oracle} ({n // 3} partials)")
        assert np.array_equal(dev.get(got), dev.get(want))
        assert np.array_equal(dev.get(got), oracle)


def test_a_box_beyond_the_array_is_refused(dev):
    """2^26 + 1 cells need one partial more than the array holds: refused before the views are read."""
    cfg, sim, nc, oc = _case("binary", recon="muscl", primary="hllc", limiter="koren")
    L = native.lib()
    out = dev.put(np.zeros(3))
    empty = native.View()
    rc = L.pga_cflmax_noalloc(nc.handle, empty, empty, empty, empty, empty, empty,
                              native.box((0, 0, 0), (65537, 1024, 1)), out.data_ptr(), None)
    assert rc == -4 and b"scratch array holds" in L.pga_last_error()
