"""The oracle's binary (mixture-averaged) diffusion (diffusion/kernels/binary.mako:11-62) against
what fixes it, after the reference's own tests (tests/transport/test_diffusion.py:9-112): a lone
species takes its self-diffusion coefficient and every other species its pair with the lone one;
traces of either sign below 64 eps read as the pure cell; a bath sums over its own species by the
rule written out; a bath species with none of its bath present is alone; and each pair's fit is
within the refit tolerance of Chapman-Enskog's coefficient evaluated directly (Cantera's part in
the reference's test, absent here)."""

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd.mixture.constants import kB
from peregrine_amd.mixture.fits import evaluate
from peregrine_amd.mixture.integrals import CollisionIntegrals
from peregrine_amd.mixture.transport import pair_parameters
from peregrine_amd.simulator import get_simulator


def _case(bath=None):
    cfg = sp.config("h2o2", trans="kinetic-theory")
    cfg["mixture"]["diffusion"] = "binary"
    cfg["mixture"]["bath"] = bath
    sim = get_simulator(cfg)
    return sim, O.OracleCase(sim.mixture)


def _D(sim, oc, p, T, Y):
    """trstate's D_n of one cell at (p, T, Y) (Y may hold traces of either sign)."""
    ns = sim.ns
    hb = sp.HostBlock((1, 1, 1), 0, sim.ne, 1)
    q = np.zeros((1, 1, 1, sim.ne))
    q[..., 0], q[..., 4] = p, T
    q[..., 5:] = np.abs(Y[:-1])
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), (1, 1, 1)))
    # the mass fractions as given, signs included: rho Y into Q, the state's p and T kept
    hb.regs[0][5:, 0, 0, 0] = hb.regs[0][0, 0, 0, 0] * Y[:-1]
    tr, hs = np.zeros((2 + ns, 1, 1, 1)), np.zeros((ns, 1, 1, 1))
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(tr), O.arr(hs), O.box((0, 0, 0), (1, 1, 1)))
    return tr[2:, 0, 0, 0]


def _pairs(sim, T):
    """D_nm at unit pressure from the case's own pair fits at T, by species pair."""
    k = sim.mixture.constants()
    ns = sim.ns
    slot = np.zeros((ns, ns), dtype=int)
    slot[np.triu_indices(ns)] = np.arange(ns * (ns + 1) // 2)
    slot = np.maximum(slot, slot.T)
    lnT = np.log(T)
    fits = np.array([evaluate(row, lnT) for row in k["dijinv"]])
    return T**1.5 / fits[slot]


@pytest.mark.parametrize("lone", [0, 4, 8])
def test_a_pure_cell(lone):
    sim, oc = _case()
    p, T = 3e5, 1300.0
    Y = np.zeros(sim.ns)
    Y[lone] = 1.0
    D = _D(sim, oc, p, T, Y)
    pairs = _pairs(sim, T)
    assert D[lone] == pytest.approx(pairs[lone, lone] / p, rel=1e-12)
    others = np.arange(sim.ns) != lone
    # the others' mixture-averaged formula collapses to the pair with the lone species
    assert D[others] == pytest.approx(pairs[others, lone] / p, rel=1e-12)


def test_traces_of_either_sign_read_as_the_pure_cell():
    sim, oc = _case()
    names = sim.mixture.spnames
    p, T = 2e5, 900.0
    pairs = _pairs(sim, T)
    lone, up = sim.ns - 1, names.index("OH")
    Y = np.zeros(sim.ns)
    trace = 1e-44
    Y[names.index("O2")], Y[up], Y[names.index("H")] = -3 * trace, 5 * trace, -0.8 * trace
    Y[lone] = 1.0 - Y[:-1].sum()
    D = _D(sim, oc, p, T, Y)
    assert np.isfinite(D).all()
    others = np.arange(sim.ns) != lone
    assert D[others] == pytest.approx(pairs[others, lone] / p, rel=1e-12)
    assert D[lone] == pytest.approx(pairs[lone, lone] / p, rel=1e-12)
    # a trace that counts sets the direction: the dominant species' pair with it
    Y[up] = 1e-9
    Y[lone] = 1.0 - Y[:-1].sum()
    D = _D(sim, oc, p, T, Y)
    assert D[lone] == pytest.approx(pairs[lone, up] / p, rel=1e-12)


def _mixed_state(sim, seed):
    rng = np.random.default_rng(seed)
    Y = rng.random(sim.ns) + 0.05
    return 1e5 * rng.uniform(0.5, 20.0), rng.uniform(400.0, 2500.0), Y / Y.sum()


@pytest.mark.parametrize("bath", [None, ["H2", "O2", "H2O", "N2"]])
def test_the_rule_written_out(bath):
    """D_n = 1 / (p (first + second)): first = (1 - X_n) / sum_b X_b * sum_b X_b / D_nb, second =
    X_n sum_b X_b MW_b / D_nb / sum_b X_b MW_b, over the bath's partners b of n."""
    sim, oc = _case(bath)
    k = sim.mixture.constants()
    names = sim.mixture.spnames
    if bath is not None:
        assert list(k["bath"]) == [names.index(n) for n in bath]
    MW = k["MW"]
    for seed in range(6):
        p, T, Y = _mixed_state(sim, seed)
        D = _D(sim, oc, p, T, Y)
        pairs = _pairs(sim, T)
        X = Y / MW
        X /= X.sum()
        for n in range(sim.ns):
            partners = [m for m in k["bath"] if m != n]
            Xb, Db, MWb = X[partners], pairs[n, partners], MW[partners]
            first = (1.0 - X[n]) / Xb.sum() * np.sum(Xb / Db)
            second = X[n] * np.sum(Xb * MWb / Db) / np.sum(Xb * MWb)
            assert D[n] == pytest.approx(1.0 / (p * (first + second)), rel=1e-10), (seed, n)


def test_a_bath_species_with_none_of_its_bath_present_is_alone():
    sim, oc = _case(["H2", "O2", "H2O", "N2"])
    names = sim.mixture.spnames
    p, T = 4e5, 1100.0
    pairs = _pairs(sim, T)
    o2 = names.index("O2")
    others = np.arange(sim.ns) != o2
    for trace in (0.0, 1e-44):
        Y = np.zeros(sim.ns)
        Y[o2], Y[names.index("H")], Y[names.index("H2O")] = 0.5, 0.5, trace
        Y[-1] = 1.0 - Y[:-1].sum()
        D = _D(sim, oc, p, T, Y)
        assert D[o2] == pytest.approx(pairs[o2, o2] / p, rel=1e-12)
        assert D[others] == pytest.approx(pairs[others, o2] / p, rel=1e-12)


def test_the_pair_fits_follow_chapman_enskog():
    """Each pair's fitted D_nm against the formula evaluated directly, off the fit's nodes."""
    sim, _ = _case()
    T = np.linspace(330.0, 3400.0, 17)
    pp = pair_parameters(sim.mixture.species)
    ci = CollisionIntegrals()
    got = np.array([_pairs(sim, t) for t in T])
    for n in range(sim.ns):
        for m in range(sim.ns):
            Tstar = T * kB / pp["rwell"][n, m]
            delta = np.full_like(T, pp["rdelta"][n, m])
            omega11 = ci.omega22(Tstar, delta) / ci.astar(Tstar, delta)
            D = 3.0 / 16.0 * np.sqrt(2.0 * np.pi / pp["rmass"][n, m]) * (kB * T) ** 1.5 / (np.pi * pp["rdiam"][n, m] ** 2 * omega11)
            assert np.abs(got[:, n, m] / D - 1.0).max() < 2e-3, (n, m)


def test_the_pairs_give_the_measured_coefficients():
    """At 1 atm: N2-O2 0.202 cm^2/s at 293 K and H2-N2 0.784 cm^2/s at 298 K (Marrero and Mason,
    J. Phys. Chem. Ref. Data 1, 3 (1972)), to the few percent a Lennard-Jones estimate reaches."""
    sim, _ = _case()
    names = sim.mixture.spnames
    n2, o2, h2 = names.index("N2"), names.index("O2"), names.index("H2")
    cm2 = 1e4 / 101325.0
    assert _pairs(sim, 293.0)[n2, o2] * cm2 == pytest.approx(0.202, rel=0.03)
    assert _pairs(sim, 298.0)[h2, n2] * cm2 == pytest.approx(0.784, rel=0.04)
