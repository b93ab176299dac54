"""The oracle's sub-stepped chemistry (chemint/kernels/substep.mako, bulk.mako) against what fixes
it, after the reference's own tests (tests/chemint/test_integration.py:7-46,
test_equilibrium.py:10-55): one substep inside the bound is the finite-rate source to the bit,
substeps keep every mass fraction in [0, 1] where the explicit source overshoots, a cell ignites
when a stiff reactor integration of the textbook rates says it does (Cantera's part in the
reference's test, absent here), a settled cell leaves by the exit in one relaxed move, a
right-hand side over no time is the finite-rate source, and the exit never changes an unsettled
cell's result."""

import ctypes

import numpy as np
import pytest

import support as sp
from chem_exact import density
from oracle import oracle as O
from peregrine_amd.simulator import get_simulator


@pytest.fixture(scope="module")
def h2():
    cfg = sp.config("h2o2", chem=True)
    cfg["chem-integration"]["kind"] = "sub-stepped"
    sim = get_simulator(cfg)
    return sim, O.OracleCase(sim.mixture)


def stoichiometric(sim, radicals=0.0):
    names = sim.mixture.spnames
    Y = np.zeros(len(names))
    for name, v in sp.H2_AIR_Y.items():
        Y[names.index(name)] = v
    for name in ("H", "O", "OH"):
        Y[names.index(name)] = radicals
    Y[-1] = 1.0 - Y[:-1].sum()
    return Y


def cell(sim, oc, p, T, Y):
    """One cell at (p, T, Y), at rest: Q and state by the oracle's fromprims."""
    hb = sp.HostBlock((1, 1, 1), 0, sim.ne, 2)
    q = np.zeros((1, 1, 1, sim.ne))
    q[..., 0], q[..., 4], q[..., 5:] = p, T, Y[:-1]
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), (1, 1, 1)))
    return hb


def substep(sim, oc, hb, cap, eqtol, dt):
    """The cell's sub-stepped source (ns - 1 carried rows) and its substep count."""
    out = np.zeros((sim.ns - 1, 1, 1, 1))
    counts = np.zeros(1, dtype=np.int32)
    O.lib.pgo_substep(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(out), 0,
                      counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cap, eqtol, dt, O.box((0, 0, 0), (1, 1, 1)))
    return out[:, 0, 0, 0], int(counts[0])


def explicit(sim, oc, hb):
    """The coupled (finite-rate) source of the cell's carried species, from the same mass
    fractions the substeps start from."""
    dQ = np.zeros_like(hb.regs[0])
    O.lib.pgo_source(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(dQ), O.box((0, 0, 0), (1, 1, 1)))
    return dQ[5:, 0, 0, 0]


def massfrac(hb):
    """The cell's mass fractions as the kernels read them (solution.mako:25-35)."""
    Q = sp.to_host(hb.regs[0])[0, 0, 0]
    Y = np.clip(Q[5:] / Q[0], 0.0, 1.0)
    Y = Y / max(Y.sum(), 1.0)
    return Q[0], np.append(Y, max(1.0 - Y.sum(), 0.0))


def test_one_substep_within_the_bound_is_the_finite_rate_source(h2):
    sim, oc = h2
    rng = np.random.default_rng(5)
    Y = rng.random(sim.ns)
    Y /= Y.sum()
    hb = cell(sim, oc, 10e5, 1800.0, Y)
    rho, Yc = massfrac(hb)
    omega = oc.production(rho, Yc, 1800.0)
    # the fastest species' way to its bound, in time: half of it
    headroom = np.where(omega < 0, Yc, 1.0 - Yc)
    with np.errstate(divide="ignore"):
        dt = 0.5 * (rho * headroom / np.abs(omega)).min()
    src, n = substep(sim, oc, hb, 1, 1e-7, dt)
    assert n == 1
    assert np.array_equal(src, explicit(sim, oc, hb))


@pytest.mark.parametrize("cap", [1, 10, 50])
def test_substeps_keep_the_state_in_bounds(h2, cap):
    sim, oc = h2
    dt = 1e-5
    hb = cell(sim, oc, 10e5, 2500.0, stoichiometric(sim, radicals=2e-3))
    Q = sp.to_host(hb.regs[0])[0, 0, 0]
    after = (Q[5:] + explicit(sim, oc, hb) * dt) / Q[0]
    assert after.min() < -0.1 or after.max() > 1.1
    src, n = substep(sim, oc, hb, cap, 1e-7, dt)
    after = (Q[5:] + src * dt) / Q[0]
    assert np.isfinite(src).all() and n == cap
    assert after.min() > -1e-12 and after.sum() < 1.0 + 1e-12
    # the rates of every species sum to nothing in every substep; the carried ones' mean sum is
    # what the last species' would balance, which here (N2, inert) is nothing too
    assert abs(src.sum()) <= 1e-12 * np.abs(src).max()


def _reactor_delay(mix, oc, p0, T0, Y0, rise):
    """The time a constant-volume reactor takes to rise by `rise` K, by an implicit Radau
    integration with the internal energy fixed. The rates are the oracle's production, which
    test_oracle.py pins to the textbook mass action; what this integration pins is the
    sub-stepping."""
    from scipy.integrate import solve_ivp

    k = mix.constants()
    cppoly, hpoly, h298, Rsp = (np.asarray(k[n]) for n in ("cppoly", "hpoly", "h298", "Rsp"))
    rho = density(mix, Y0, T0, p0)

    def f(t, y):
        Y, T = y[:-1], y[-1]
        u = np.log(T)
        omega = oc.production(rho, Y, T)
        cpR = np.array([np.polyval(np.flip(c), u) for c in cppoly])
        hRT = np.array([np.polyval(np.flip(c), u) for c in hpoly]) + h298 / T
        e = (hRT - 1.0) * T * Rsp
        cv = ((cpR - 1.0) * Rsp * Y).sum()
        return np.append(omega / rho, -(e * omega).sum() / (rho * cv))

    def risen(t, y):
        return y[-1] - (T0 + rise)

    risen.terminal, risen.direction = True, 1
    sol = solve_ivp(f, (0.0, 1e-2), np.append(Y0, T0), method="Radau", rtol=1e-9, atol=1e-14, events=risen)
    return sol.t_events[0][0]


def _our_delay(sim, oc, p0, T0, Y0, rise, dt, cap, eqtol, until):
    """Steps one walled cell under tvd-rk3 with sub-stepped chemistry until T rises by `rise`."""
    hb = sp.HostBlock((1, 1, 1), 1, sim.ne, 3)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0], q[..., 4], q[..., 5:] = p0, T0, Y0[:-1]
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))
    b = sp.oracle_block(hb, "tvd-rk3", 2, walls=(0, 1, 2, 3, 4, 5))
    counts = np.zeros(1, dtype=np.int32)
    b.subcap, b.subeqtol, b.substeps = cap, eqtol, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    g, t, before = hb.ng, 0.0, T0
    while t < until:
        O.lib.pgo_step(oc.ref, ctypes.byref(b), dt)
        t += dt
        T = hb.state[1, g, g, g]
        if T > T0 + rise:
            # the crossing by linear interpolation within the step
            return t - dt * (T - (T0 + rise)) / (T - before), hb.regs[0].copy()
        before = T
    return None, hb.regs[0].copy()


@pytest.mark.parametrize("T0", [1200.0, 1500.0])
def test_ignition_follows_the_reactor(h2, T0):
    """20 bar, stoichiometric H2-air: the time to a 400 K rise within 1% of the stiff
    integration's (the reference measured 0.2% against Cantera at this dt and cap), with the
    equilibrium exit and without."""
    sim, oc = h2
    p0, Y0 = 20e5, stoichiometric(sim)
    theirs = _reactor_delay(sim.mixture, oc, p0, T0, Y0, 400.0)
    # a step well inside the induction time: the cap of ten substeps would throttle a longer one
    dt = theirs / 2000.0
    for eqtol in (0.0, 1e-7):
        ours, _ = _our_delay(sim, oc, p0, T0, Y0, 400.0, dt, 10, eqtol, 4 * theirs)
        print("delay", T0, eqtol, ours, theirs)
        assert ours is not None and abs(ours / theirs - 1.0) < 0.01


def test_a_settled_cell_finishes_in_one_relaxed_move(h2):
    sim, oc = h2
    hb = cell(sim, oc, 10e5, 400.0, stoichiometric(sim))
    omega = explicit(sim, oc, hb)
    src, n = substep(sim, oc, hb, 200, 1e-7, 1e-6)
    assert n == 1
    scale = np.abs(omega).max()
    assert np.abs(src - omega).max() < 1e-6 * scale
    assert not np.array_equal(src, omega)
    # without the exit the one substep covers the step: the finite-rate source to the bit
    src0, n0 = substep(sim, oc, hb, 200, 0.0, 1e-6)
    assert n0 == 1 and np.array_equal(src0, omega)


def test_a_right_hand_side_over_no_time_is_the_finite_rate_source(h2):
    """Before a step sets it the step is zero: no exit is judged over no time, one substep."""
    sim, oc = h2
    hb = cell(sim, oc, 10e5, 400.0, stoichiometric(sim))
    src, n = substep(sim, oc, hb, 200, 1e-7, 0.0)
    assert n == 1 and np.array_equal(src, explicit(sim, oc, hb))


def test_an_unsettled_cell_substeps_as_the_one_pass_form(h2):
    """Products at 2500 K never settle: with and without the exit, a step ends alike to the bit."""
    sim, oc = h2
    p0, T0, Y0 = 10e5, 2500.0, stoichiometric(sim)
    out = []
    for tol in (1e-7, 0.0):
        hb = sp.HostBlock((1, 1, 1), 1, sim.ne, 3)
        q = np.zeros(hb.cshape + (sim.ne,))
        q[..., 0], q[..., 4], q[..., 5:] = p0, T0, Y0[:-1]
        O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state),
                            O.box((0, 0, 0), hb.cshape))
        b = sp.oracle_block(hb, "tvd-rk3", 2, walls=(0, 1, 2, 3, 4, 5))
        counts = np.zeros(1, dtype=np.int32)
        b.subcap, b.subeqtol, b.substeps = 200, tol, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        for _ in range(3):
            O.lib.pgo_step(oc.ref, ctypes.byref(b), 1e-6)
        out.append((hb.regs[0].copy(), int(counts[0])))
    assert out[0][1] == out[1][1] and out[0][1] > 1
    assert np.array_equal(out[0][0], out[1][0])
