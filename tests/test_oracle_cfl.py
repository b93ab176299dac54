"""The oracle's CFL reduction (euler/kernels/cflmax.mako:3-33) against closed forms, and the host
side of the CFL controller and the plugins' cadence (integrators/controllers.py:19-29,
plugins/solver/cadence.py:4-25)."""

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd import grid
from peregrine_amd.integrators import CFLController, FixedController, get_integrator
from peregrine_amd.plugins import Cadence
from peregrine_amd.simulator import get_simulator


def _block(ncell, lengths, jostle=0.0):
    cfg = sp.config("air-cpg")
    sim = get_simulator(cfg)
    oc = O.OracleCase(sim.mixture)
    hb = sp.HostBlock(ncell, sim.ng, sim.ne, 1, lengths=lengths, jostle=jostle, seed=5)
    d = sp.to_dev(grid.dijk(hb.nodes, ncell))
    return sim, oc, hb, d


def _fill(oc, hb, q):
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))


def _cfl(hb, d):
    out = np.zeros(3)
    g = hb.ng
    O.lib.pgo_cflmax(O.arr(hb.regs[0]), O.arr(hb.state), O.arr(d), *(O.arr(s) for s in hb.S),
                     O.box((g, g, g), hb.ncell), out.ctypes.data_as(O.C.POINTER(O.C.c_double)))
    return out


def test_uniform_flow_on_a_uniform_grid():
    ncell, lengths = (8, 5, 4), (0.8, 1.0, 0.2)
    sim, oc, hb, d = _block(ncell, lengths)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0], q[..., 1], q[..., 2], q[..., 3], q[..., 4] = 101325.0, 30.0, -12.0, 5.0, 300.0
    _fill(oc, hb, q)
    c = hb.state[7].max()
    dx = [L / n for L, n in zip(lengths, ncell)]
    acoustic = c * np.sqrt(sum(1.0 / h**2 for h in dx))
    convective = 30.0 / dx[0] + 12.0 / dx[1] + 5.0 / dx[2]
    got = _cfl(hb, d)
    assert np.allclose(got, [acoustic, convective, acoustic + convective], rtol=1e-13, atol=0)


def test_an_axis_one_cell_thick_sets_no_limit():
    ncell, lengths = (6, 5, 1), (0.6, 0.5, 0.1)
    sim, oc, hb, d = _block(ncell, lengths)
    assert np.isinf(grid.dijk(hb.nodes, ncell)[..., 2]).all()
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0], q[..., 1], q[..., 3], q[..., 4] = 101325.0, 10.0, 1000.0, 300.0
    _fill(oc, hb, q)
    c = hb.state[7].max()
    got = _cfl(hb, d)
    # w along the thin axis contributes nothing, nor does that axis's acoustic length
    assert np.isclose(got[0], c * np.sqrt(2.0) / 0.1, rtol=1e-13)
    assert np.isclose(got[1], 10.0 / 0.1, rtol=1e-13)


def test_the_largest_cell_wins_on_a_jostled_grid():
    ncell, lengths = (7, 6, 5), (0.7, 0.6, 0.5)
    sim, oc, hb, d = _block(ncell, lengths, jostle=0.3)
    rng = np.random.default_rng(4)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = 101325.0 * rng.uniform(0.5, 2.0, hb.cshape)
    q[..., 1:4] = rng.uniform(-50, 50, hb.cshape + (3,))
    q[..., 4] = rng.uniform(250.0, 600.0, hb.cshape)
    _fill(oc, hb, q)
    # every cell's speeds from the definition, in numpy, and their maxima
    g = hb.ng
    x = sp.to_host(d)[g:-g, g:-g, g:-g]
    Q, st = sp.to_host(hb.regs[0])[g:-g, g:-g, g:-g], sp.to_host(hb.state)[g:-g, g:-g, g:-g]
    vel = Q[..., 1:4] / Q[..., :1]
    conv = 0.0
    for a in range(3):
        S = sp.to_host(hb.S[a])
        lo = S[g:g + ncell[0], g:g + ncell[1], g:g + ncell[2]]
        sl = [slice(g, g + n) for n in ncell]
        sl[a] = slice(g + 1, g + 1 + ncell[a])
        hi = S[tuple(sl)]
        n = 0.5 * (lo / np.linalg.norm(lo, axis=-1, keepdims=True) + hi / np.linalg.norm(hi, axis=-1, keepdims=True))
        conv = conv + np.abs((n * vel).sum(axis=-1)) / x[..., a]
    ac = st[..., 7] * np.sqrt((1.0 / x**2).sum(axis=-1))
    want = [ac.max(), conv.max(), (ac + conv).max()]
    assert np.allclose(_cfl(hb, d), want, rtol=1e-13, atol=0)


class _Clock:
    def __init__(self, t):
        self.t = t

    def get(self):
        return self.t


class _Plugin:
    def __init__(self, cfg):
        self.cadence = Cadence(cfg)

    def tnext(self, solver):
        return self.cadence.next(solver)


class _Solver:
    def __init__(self, tme, combined, plugins=(), nrt=0, dt=0.0):
        self.tme, self.dt = _Clock(tme), _Clock(dt)
        self.plugins, self.nrt = list(plugins), nrt
        self._combined = combined

    def cflmax(self):
        return {"acoustic": 0.0, "convective": 0.0, "combined": self._combined}


def test_the_cfl_controller():
    cfg = sp.config("air-cpg")
    cfg["simulation"]["controller"], cfg["simulation"]["cfl"], cfg["simulation"]["dt-max"] = "cfl", 0.5, 1e-3
    ctl = get_integrator(cfg).controller
    assert isinstance(ctl, CFLController)
    assert ctl.stepsize(_Solver(0.0, 2e4)) == 0.5 / 2e4
    # dt-max binds
    assert ctl.stepsize(_Solver(0.0, 10.0)) == 1e-3
    # shortened to land on the next output time of a dt-out plugin
    p = _Plugin({"dt-out": 1e-4})
    assert ctl.stepsize(_Solver(0.99e-4, 10.0, [p])) == pytest.approx(1e-6, rel=1e-9)
    cfg["simulation"]["controller"] = "fixed"
    assert isinstance(get_integrator(cfg).controller, FixedController)
    cfg["simulation"]["controller"] = "pid"
    with pytest.raises(ValueError):
        get_integrator(cfg)


def test_the_cadence():
    c = Cadence({"niter-out": 3})
    assert [c.due(_Solver(0.0, 1.0, nrt=n)) for n in range(1, 7)] == [False, False, True, False, False, True]
    assert c.next(_Solver(0.0, 1.0)) is None
    c = Cadence({"dt-out": 0.1})
    # a step that lands a few ulps short of 0.3 counts as reaching it
    assert c.due(_Solver(0.1 + 0.1 + 0.1 - 1e-17, 1.0, dt=0.05))
    assert not c.due(_Solver(0.26, 1.0, dt=0.05))
    assert c.next(_Solver(0.26, 1.0)) == pytest.approx(0.3)
    assert Cadence({}).due(_Solver(0.0, 1.0, nrt=5))
