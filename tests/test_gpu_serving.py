"""The serving loop on the MI355X: each request's result the same alone and in the mix (to the bit),
admission held to the HBM budget, and a reactor's ignition against the oracle's FP64 implicit form."""

import numpy as np
import pytest

import support as sp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _stream():
    from peregrine_amd import serving

    reqs = [serving.ignition(T0, p0, phi, tend=3e-5) for T0, p0, phi in
            ((1100.0, 5e5, 1.0), (1300.0, 2e6, 0.6), (950.0, 4e6, 1.8), (1400.0, 1e5, 1.2), (1200.0, 1e6, 0.9))]
    reqs += [serving.tube(200), serving.tube(333), serving.taylor_green(16, steps=3), serving.taylor_green(20, steps=2)]
    for n, r in enumerate(reqs):
        r.rid, r.arrival = n, 0.002 * n
    return reqs


def _same(a, b):
    for k in a:
        if k in ("delay", "sweeps"):
            continue
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if not np.array_equal(x, y):
            return k
    return None


def test_each_result_is_the_same_alone_and_in_the_mix(gpu):
    import peregrine_amd as pg
    from peregrine_amd import serving

    backend = pg.get_backend("mi355x", sp.config("air-cpg"))
    mix = serving.Server(backend, reactors=16)
    done, wall = mix.run(_stream())
    assert len(done) == 9 and all(r.finished >= r.arrival for r in done)
    p50, p99 = serving.percentiles(done)
    print(f"mix: {len(done)} requests in {wall:.3f} s, p50 {p50 * 1e3:.1f} ms, p99 {p99 * 1e3:.1f} ms, "
          f"ticks {mix.ticks}")
    alone = serving.Server(backend, reactors=1, solvers=1)
    for r in done:
        one = serving.Request(r.kind, dict(r.params), rid=r.rid)
        (d,), _ = alone.run([one])
        assert _same(r.result, d.result) is None, (r.kind, r.params, _same(r.result, d.result))
        if r.kind == "ignition":
            assert r.result["delay"] == d.result["delay"] or (np.isnan(r.result["delay"]) and np.isnan(d.result["delay"]))


def test_admission_holds_the_budget(gpu):
    """A budget that holds one 333-cell tube and no more: the solvers never overlap, and their results are
    those of an unlimited server."""
    import peregrine_amd as pg
    from peregrine_amd import serving

    backend = pg.get_backend("mi355x", sp.config("air-cpg"))
    reqs = [serving.tube(333), serving.tube(300), serving.tube(250)]
    need = serving.solver_footprint("tube", 333)
    pool = serving.ReactorPool.footprint(4) * 4
    tight = serving.Server(backend, reactors=4, budget=pool + need)
    seen = []
    orig = tight._admit

    def admit(req, now):
        ok = orig(req, now)
        seen.append(len(tight.solvers))
        return ok

    tight._admit = admit
    done, _ = tight.run([serving.Request(r.kind, dict(r.params), rid=n) for n, r in enumerate(reqs)])
    assert max(seen) == 1 and len(done) == 3
    free = serving.Server(backend, reactors=4)
    done2, _ = free.run([serving.Request(r.kind, dict(r.params), rid=n) for n, r in enumerate(reqs)])
    by = {r.rid: r for r in done2}
    for r in done:
        assert np.array_equal(r.result["rho"], by[r.rid].result["rho"])


def test_a_reactor_follows_the_fp64_implicit_form(gpu):
    """One reactor (1500 K, 20 bar, stoichiometric) served over 300 steps of 0.1 us against the oracle's
    FP64 linearly implicit form stepped the same way (mean rates over dt, rho Y advanced, the EOS): Y and
    T within 1e-8."""
    import peregrine_amd as pg
    from peregrine_amd import serving

    backend = pg.get_backend("mi355x", sp.config("air-cpg"))
    srv = serving.Server(backend, reactors=2)
    req = serving.ignition(1500.0, 2e6, 1.0, tend=3e-5)
    (got,), _ = srv.run([req])
    sim = srv.pool.sim
    oc = O.OracleCase(sim.mixture)
    hb = sp.HostBlock((1, 1, 1), 0, sim.ne, 1)
    q = np.zeros((1, 1, 1, sim.ne))
    q[..., 0], q[..., 4], q[..., 5:] = 2e6, 1500.0, serving.h2_air(1.0, sim.mixture.spnames)[:-1]
    full = O.box((0, 0, 0), (1, 1, 1))
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    rates = np.zeros((sim.ns - 1, 1, 1, 1))
    dt = 1e-7
    for _ in range(300):
        O.lib.pgo_implicit(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(rates), 0, 1, dt, full)
        hb.regs[0][5:] += dt * rates
        O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), full)
    Q = hb.regs[0][:, 0, 0, 0]
    Y = Q[5:] / Q[0]
    want_Y, want_T = np.append(Y, 1.0 - Y.sum()), hb.state[1, 0, 0, 0]
    dY = np.abs(got.result["Y"] - want_Y).max()
    dT = abs(got.result["T"] - want_T) / want_T
    print(f"reactor 300 steps: |dY| {dY:.2e}, dT/T {dT:.2e}, T {got.result['T']:.2f} K")
    assert dY <= 1e-8 and dT <= 1e-8
    assert abs(got.result["T"] - 1500.0) > 1.0
