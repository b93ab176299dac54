"""The oracle pinned against what the mathematics fixes, before any device result is compared with it.

Each check is chosen so a plausible slip in the restatement (a dropped term, a wrong sign or
index, a transposed operand) fails it:
- table log/exp against the library within the reference's stated accuracy;
- production rates against textbook mass action (tests/chem_exact.py), element and mass
  conservation, exact zeros for species no reaction can reach;
- the EOS round trip (p, T) -> Q -> (p, T) and the calorically perfect closed form;
- every flux function on equal sides equals the Euler flux of that state;
- freestream preservation on a jostled periodic grid and telescoping conservation;
- MUSCL sides exact for linear data;
- the RK combines' amplification polynomials.
"""

import math

import numpy as np
import pytest

import support as sp
from chem_exact import density, production
from oracle import oracle as O
from peregrine_amd.mixture.constants import Ru, Tref
from peregrine_amd.simulator import get_simulator


def test_table_log_and_exp_are_within_their_stated_accuracy():
    rng = np.random.default_rng(1)
    x = np.exp(rng.uniform(-700.0, 700.0, 20000))
    # backend/base/kernels/mathfns.mako:5: within 3e-14 absolute, two ulps past |log x| = 100
    err = np.array([abs(O.tlog(v) - math.log(v)) for v in x])
    bound = np.maximum(3e-14, 2 * np.spacing(np.abs(np.log(x))))
    assert (err <= bound).all()
    y = rng.uniform(-708.0, 709.0, 20000)
    rel = np.array([abs(O.texp(v) / math.exp(v) - 1.0) for v in y])
    assert rel.max() <= 2.5 * np.finfo(float).eps
    # e^0 is exact; log 1 has no relative accuracy, only the absolute bound
    assert O.texp(0.0) == 1.0
    assert abs(O.tlog(1.0)) < 3e-14


def test_the_eos_log_is_within_an_ulp():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(50.0, 20000.0, 20000), np.exp(rng.uniform(-700, 700, 5000)),
                        1.0 + rng.uniform(-0.05, 0.05, 5000)])
    got = np.array([O.elog(v) for v in x])
    want = np.log(x)
    assert (np.abs(got - want) <= np.spacing(np.abs(want))).all()


@pytest.fixture(scope="module")
def h2():
    sim = get_simulator(sp.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True))
    return sim.mixture, O.OracleCase(sim.mixture, "muscl", "koren", "hllc")


def _random_state(rng, ns):
    T = rng.uniform(900.0, 2600.0)
    p = 101325.0 * rng.uniform(0.5, 20.0)
    Y = rng.random(ns) ** 3
    Y /= Y.sum()
    return p, T, Y


def test_rates_are_mass_action(h2):
    mix, oc = h2
    rng = np.random.default_rng(7)
    for _ in range(40):
        p, T, Y = _random_state(rng, mix.ns)
        rho = density(mix, Y, T, p)
        got = oc.production(rho, Y, T)
        # with ln T taken as the kernels take it (the table log), only rounding separates the forms:
        # the Gibbs polynomials' terms cancel by six orders, and the oracle's fused Horner steps
        # round apart from numpy's plain ones: measured 8.1e-11 of the largest rate (1.0e-14 when
        # both take plain steps), the polynomials' own rounding floor carried into K
        want = production(mix, rho, Y, T, u=O.tlog(T))
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 2e-10 * scale
        # with the library log the table log's error shows through the Gibbs polynomials (measured 5.7e-11)
        want = production(mix, rho, Y, T)
        assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()


def test_troe_centre_fit_is_within_its_budget(h2):
    mix, _ = h2
    records, sc = mix.chemistry.records(mix.species)

    T = np.geomspace(300.0, 3500.0, 400)
    for r in records:
        if r.troe:
            t = (np.log(T) - sc["troeuc"]) * sc["troeuhinv"]
            fit = np.polyval(np.flip(r.lfcpoly), t)
            assert np.abs(fit - r.fcent.log10(T)).max() < 1e-3


def test_rates_conserve_mass_and_elements(h2):
    mix, oc = h2
    rng = np.random.default_rng(5)
    comp = np.array([[sp_["comp"].get(e, 0.0) for e in ("H", "O", "N")] for sp_ in mix.species.values()])
    MW = np.array([s["MW"] for s in mix.species.values()])
    for _ in range(20):
        p, T, Y = _random_state(rng, mix.ns)
        w = oc.production(density(mix, Y, T, p), Y, T)
        scale = np.abs(w).max()
        assert abs(w.sum()) < 1e-12 * scale
        # atoms per kmol of each element produced: sum_n omega_n / MW_n * a_{e,n}
        atoms = (w / MW) @ comp
        assert np.abs(atoms).max() < 1e-12 * np.abs(w / MW).max()


def test_an_unreachable_species_rate_is_exactly_nothing(h2):
    mix, oc = h2
    names = mix.spnames
    Y = np.zeros(mix.ns)
    Y[names.index("H2")], Y[names.index("O2")] = 0.028, 0.226
    Y[names.index("N2")] = 1.0 - Y.sum()
    T = 1200.0
    w = oc.production(density(mix, Y, T, 101325.0), Y, T)
    assert np.isfinite(w).all()
    # with H2, O2 and N2 alone only O2 + H2 -> H + HO2 and O2's dissociation run
    assert w[names.index("H2")] < 0 and w[names.index("O2")] < 0
    assert w[names.index("H")] > 0 and w[names.index("HO2")] > 0 and w[names.index("O")] > 0
    still = [names.index(n) for n in ("OH", "H2O", "H2O2", "N2")]
    assert np.array_equal(w[still], np.zeros(len(still)))


def test_a_cell_of_one_species_is_still(h2):
    mix, oc = h2
    Y = np.zeros(mix.ns)
    Y[mix.spnames.index("N2")] = 1.0
    w = oc.production(density(mix, Y, 300.0, 1e5), Y, 300.0)
    # the bath only ever takes part as a third body
    assert np.isfinite(w).all() and np.array_equal(w, np.zeros(mix.ns))
    # cold O2 alone dissociates, at 1e-75 kg/m^3/s, and stays finite
    Y[:] = 0.0
    Y[mix.spnames.index("O2")] = 1.0
    w = oc.production(density(mix, Y, 300.0, 1e5), Y, 300.0)
    assert np.isfinite(w).all() and np.abs(w).max() < 1e-60


def _eos_block(gas, ncell=(6, 5, 4)):
    sim = get_simulator(sp.config(gas))
    oc = O.OracleCase(sim.mixture)
    hb = sp.HostBlock(ncell, 1, sim.ne, 2)
    return sim, oc, hb


@pytest.mark.parametrize("gas", ["air-cpg", "air-tpg", "h2o2", "binary"])
def test_eos_round_trip(gas):
    sim, oc, hb = _eos_block(gas)
    rng = np.random.default_rng(2)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = 101325.0 * rng.uniform(0.5, 3.0, hb.cshape)
    q[..., 1:4] = rng.uniform(-100, 100, hb.cshape + (3,))
    q[..., 4] = rng.uniform(350.0, 2500.0, hb.cshape)
    if sim.ns > 1:
        Y = rng.random(hb.cshape + (sim.ns,))
        Y /= Y.sum(axis=-1, keepdims=True)
        q[..., 5:] = Y[..., :-1]
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    # a warm start (the state's own T) and a cold one (T = 0 starts at 300 K)
    for cold in (False, True):
        st = hb.state.copy()
        if cold:
            st[1] = 0.0
        O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(st), full)
        tol = 1e-10 if not cold else 1e-9
        assert np.abs(st[1] / sp.to_dev(q)[4] - 1).max() < tol
        assert np.abs(st[0] / sp.to_dev(q)[0] - 1).max() < tol


def test_cpg_state_is_the_closed_form():
    sim, oc, hb = _eos_block("air-cpg")
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0], q[..., 4] = 2e5, 450.0
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    MW, cp = sp.AIR_CPG["Air"]["MW"], sp.AIR_CPG["Air"]["cp0"]
    R = Ru / MW
    cv = cp - R
    # h = cp (T - Tref) with no formation enthalpy, so e = cv T - cp Tref (tests/simulator/oracle.py:19-28)
    want = dict(rho=2e5 / (R * 450.0), e=cv * 450.0 - cp * Tref, gamma=cp / cv, c=math.sqrt(cp / cv * R * 450.0))
    st = hb.state[:, 1, 1, 1]
    assert abs(st[3] / want["rho"] - 1) < 1e-15
    assert abs(st[4] / want["e"] - 1) < 1e-13
    assert abs(st[6] / want["gamma"] - 1) < 1e-15
    assert abs(st[7] / want["c"] - 1) < 1e-15


def _euler_flux(rho, u, p, e, Y, S):
    U = u @ S
    E = e + 0.5 * (u @ u)
    F = [rho * U, *(rho * u * U + p * S), rho * E * U + p * U]
    return np.array(F + [rho * y * U for y in Y])


@pytest.mark.parametrize("primary,recon", [("kepaec", "constant"), ("hllc", "constant"), ("rusanov", "constant"),
                                           ("hllc", "muscl"), ("rusanov", "muscl")])
@pytest.mark.parametrize("gas", ["air-tpg", "binary"])
def test_a_uniform_state_gives_the_euler_flux_and_no_change(primary, recon, gas):
    cfg = sp.config(gas, recon=recon, primary=primary, limiter="koren" if recon == "muscl" else None)
    sim = get_simulator(cfg)
    oc = O.OracleCase(sim.mixture, recon, "koren" if recon == "muscl" else None, primary)
    hb = sp.HostBlock((6, 5, 4), sim.ng, sim.ne, 3, jostle=0.3, seed=4)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[...] = [1.3e5, 40.0, -25.0, 15.0, 500.0] + ([0.3] if sim.ns == 2 else [])
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    st = hb.state[:, 2, 2, 2]
    Y = [0.3] if sim.ns == 2 else []
    ng = hb.ng
    for a in range(3):
        ext = list(hb.ncell)
        ext[a] += 1
        O.lib.pgo_flux(oc.ref, a, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.S[a]), O.arr(hb.F[a]),
                       O.box((ng,) * 3, ext))
        S = hb.S[a][:, ng, ng, ng]
        want = _euler_flux(st[3], np.array([40.0, -25.0, 15.0]), st[0], st[4], Y, S)
        got = hb.F[a][:, ng, ng, ng]
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    dQ = hb.regs[-1]
    O.lib.pgo_applyflux(sim.ne, O.arr(dQ), *(O.arr(f) for f in hb.F), O.arr(hb.detJinv), 0,
                        O.box((ng,) * 3, hb.ncell))
    # freestream: the face areas of a closed cell sum to nothing (tests/minimal/test_freestream.py:12-28)
    Fscale = max(np.abs(f[:, ng:-ng, ng:-ng, ng:-ng]).max() for f in hb.F) * hb.detJinv[0].max()
    assert np.abs(dQ[hb.interior()]).max() < 1e-13 * Fscale


@pytest.mark.parametrize("primary,recon", [("kepaec", "constant"), ("hllc", "muscl"), ("rusanov", "muscl")])
def test_the_divergence_telescopes(primary, recon):
    limiter = "koren" if recon == "muscl" else None
    sim = get_simulator(sp.config("air-tpg", recon=recon, primary=primary, limiter=limiter))
    oc = O.OracleCase(sim.mixture, recon, limiter, primary)
    hb = sp.HostBlock((8, 6, 5), sim.ng, sim.ne, 3, jostle=0.3, seed=9)
    q = sp.smooth_prims(hb.xcc, sim.ne)
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    regs = (hb.regs[0], hb.regs[2])
    # the halo takes the partner's Q and its state as they are: a halo EOS started elsewhere lands
    # within the tpg Newton's rounding floor (1e-11 of T, the ln T fit's cancellation), which
    # would show here as a boundary face taken twice
    ng = hb.ng
    for a in range(3):
        O.lib.pgo_fill_periodic(O.arr(hb.regs[0]), sim.ne, O.ints(hb.ncell), ng, ng, a)
        O.lib.pgo_fill_periodic(O.arr(hb.state), 8, O.ints(hb.ncell), ng, ng, a)
    for a in range(3):
        ext = list(hb.ncell)
        ext[a] += 1
        O.lib.pgo_flux(oc.ref, a, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.S[a]), O.arr(hb.F[a]),
                       O.box((ng,) * 3, ext))
    dQ = regs[1]
    O.lib.pgo_applyflux(sim.ne, O.arr(dQ), *(O.arr(f) for f in hb.F), O.arr(hb.detJinv), 0,
                        O.box((ng,) * 3, hb.ncell))
    vol = 1.0 / hb.detJinv[0][hb.interior()[1:]]
    total = (dQ[hb.interior()] * vol).reshape(sim.ne, -1).sum(axis=1)
    scale = np.array([np.abs(f[:, ng:-ng, ng:-ng, ng:-ng]).max() for f in hb.F]).max()
    # periodic: every face flux leaves one cell and enters another (tests/minimal/test_conservation.py:55)
    assert np.abs(total).max() < 1e-12 * scale * np.prod(hb.ncell) ** (2 / 3)


def test_muscl_sides_are_exact_for_linear_data():
    sim = get_simulator(sp.config("air-tpg", recon="muscl", primary="hllc", limiter="koren"))
    oc = O.OracleCase(sim.mixture, "muscl", "koren", "hllc")
    hb = sp.HostBlock((8, 4, 4), 2, sim.ne, 3)
    x = hb.xcc[..., 0]
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = 1e5 * (1.0 + 0.5 * x)
    q[..., 1] = 30.0 + 20.0 * x
    q[..., 4] = 400.0 * (1.0 + 0.5 * x)
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)
    # an interior face along i: the four cells' p, T and u lie on lines, so both sides take the face value
    f = (4, 2, 2)
    O.lib.pgo_flux(oc.ref, 0, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.S[0]), O.arr(hb.F[0]), O.box(f, (1, 1, 1)))
    xf = 0.5 * (x[3, 2, 2] + x[4, 2, 2])
    oc1 = O.OracleCase(sim.mixture)
    face = np.zeros((3, 3, 3, sim.ne))
    face[..., 0], face[..., 1], face[..., 4] = 1e5 * (1 + 0.5 * xf), 30.0 + 20.0 * xf, 400.0 * (1 + 0.5 * xf)
    fQ, fs = np.zeros((sim.ne, 3, 3, 3)), np.zeros((8, 3, 3, 3))
    O.lib.pgo_fromprims(oc1.ref, O.arr(sp.to_dev(face)), O.arr(fQ), O.arr(fs), O.box((0, 0, 0), (3, 3, 3)))
    st = fs[:, 1, 1, 1]
    S = hb.S[0][:, 2, 2, 4]
    want = _euler_flux(st[3], np.array([30.0 + 20.0 * xf, 0.0, 0.0]), st[0], st[4], [], S)
    got = hb.F[0][:, 2, 2, 4]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


RK = {
    "rk1": [1, 1],
    "rk2": [1, 1, 1 / 2],
    "tvd-rk3": [1, 1, 1 / 2, 1 / 6],
    "rk4": [1, 1, 1 / 2, 1 / 6, 1 / 24],
}


@pytest.mark.parametrize("integrator", sorted(RK))
def test_the_oracle_rk_amplification(integrator):
    # dQ = z Q / dt at every stage: one step multiplies Q by the amplification polynomial R(z)
    # (tests/integrators/test_order.py:145-162)
    nregs = 4 if integrator == "rk4" else (2 if integrator == "rk1" else 3)
    nst = {"rk1": 1, "rk2": 2, "tvd-rk3": 3, "rk4": 4}[integrator]
    hb = sp.HostBlock((1, 1, 1), 1, 5, nregs)
    b = sp.oracle_block(hb, integrator, False)
    oc = O.OracleCase(get_simulator(sp.config("air-cpg")).mixture)
    dt = 0.5
    for z in (0.1, -0.37, 0.8):
        for r in hb.regs:
            r[...] = 0.0
        hb.regs[0][...] = 1.0
        for n in range(nst):
            Qin = hb.regs[0] if n == 0 else hb.regs[1]
            hb.regs[-1][...] = z / dt * Qin
            O.lib.pgo_advance(oc.ref, O.C.byref(b), n, dt)
        R = np.polyval(np.flip(RK[integrator]), z)
        assert abs(hb.regs[0][0, 1, 1, 1] - R) < 1e-15
