"""The oracle's cubic real gas (Soave-Redlich-Kwong as the reference has it, and Peng-Robinson on
the same code) and Chung's dense transport, against what the mathematics fixes: the textbook
cubic in Z and its least-Gibbs root (tests/eos/support.py:75-122 states the same), the round
trip (p, T) -> Q -> (p, T), cp and the sound speed as derivatives of the EOS itself, the ideal-gas
limit, Euler's identity for the partial enthalpies, and Chung's low-density limit."""

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd.config import Config
from peregrine_amd.mixture import get_mixture


def _mixture(eos, trans=None):
    c = Config()
    m = c["mixture"]
    m["species"], m["eos"], m["T-range"], m["trans"] = ["O2", "N2"], eos, (100.0, 1000.0), trans
    return get_mixture(c)


def _states(oc, mix, p, T, Y):
    """(p, T, Y) arrays -> Q and state by the oracle's fromprims, on a 1-D block."""
    n = len(p)
    hb = sp.HostBlock((n, 1, 1), 0, 4 + mix.ns, 1, nstate=9)
    prims = np.zeros((n, 1, 1, 4 + mix.ns))
    prims[:, 0, 0, 0], prims[:, 0, 0, 4], prims[:, 0, 0, 5] = p, T, Y
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(prims)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), (n, 1, 1)))
    return hb, sp.to_host(hb.state)[:, 0, 0], sp.to_host(hb.regs[0])[:, 0, 0]


def _grid():
    p = np.repeat(np.array([1e5, 3e6, 6e6, 1e7]), 12)
    T = np.tile(np.linspace(100.0, 600.0, 12), 4)
    Y = np.tile(np.linspace(0.05, 0.95, 12), 4)[::-1]
    return p, T, Y


@pytest.mark.parametrize("eos", ["real-gas", "peng-robinson"])
def test_density_is_the_least_gibbs_root_of_the_textbook_cubic(eos):
    mix = _mixture(eos)
    k = mix.constants()
    oc = O.OracleCase(mix)
    p, T, Y = _grid()
    hb, st, Q = _states(oc, mix, p, T, Y)
    u, w, sig, eps = k["cubic"]
    MW = np.array([s["MW"] for s in mix.species.values()])
    Ru = k["Ru"]
    for n in range(len(p)):
        Yv = np.array([Y[n], 1 - Y[n]])
        X = Yv / MW / (Yv / MW).sum()
        Tc = np.array([s["Tcrit"] for s in mix.species.values()])
        pc = np.array([s["pcrit"] for s in mix.species.values()])
        om = k["acentric"]
        mcoef = mix.eos.cubic["m"]
        m = mcoef[0] + mcoef[1] * om + mcoef[2] * om**2
        a = mix.eos.cubic["omega_a"] * (Ru * Tc) ** 2 / pc * (1 + m * (1 - np.sqrt(T[n] / Tc))) ** 2
        b = mix.eos.cubic["omega_b"] * Ru * Tc / pc
        am = (X @ np.sqrt(a)) ** 2
        bm = X @ b
        A, B = am * p[n] / (Ru * T[n]) ** 2, bm * p[n] / (Ru * T[n])
        # Z^3 - (1 - (u - 1) B) Z^2 + (A + w B^2 - u B - u B^2) Z - (A B + w B^2 + w B^3) = 0
        roots = np.roots([1.0, -(1 - (u - 1) * B), A + w * B**2 - u * B - u * B**2, -(A * B + w * B**2 + w * B**3)])
        real = [r.real for r in roots if abs(r.imag) < 1e-9 * abs(r) and r.real > B]
        g = [z - 1 - np.log(z - B) + A / (B * (sig - eps)) * np.log((z + eps * B) / (z + sig * B)) for z in real]
        Z = real[int(np.argmin(g))]
        rho = p[n] / (Z * Ru * T[n]) * (X @ MW)
        assert abs(st[n, 8] / Z - 1) < 1e-10, (n, st[n, 8], Z)
        assert abs(st[n, 3] / rho - 1) < 1e-10


@pytest.mark.parametrize("eos", ["real-gas", "peng-robinson"])
def test_round_trip_and_thermodynamic_derivatives(eos):
    mix = _mixture(eos)
    oc = O.OracleCase(mix)
    p, T, Y = _grid()
    hb, st, Q = _states(oc, mix, p, T, Y)
    # fromQ from T off by 5%: the Newton returns (p, T)
    guess = hb.state.copy()
    guess[1] *= 1.05
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(guess), O.box((0, 0, 0), (len(p), 1, 1)))
    back = sp.to_host(guess)[:, 0, 0]
    assert np.abs(back[:, 1] / T - 1).max() < 1e-10
    assert np.abs(back[:, 0] / p - 1).max() < 1e-9
    # cp = dh/dT at constant p, by central differences of the EOS's own h = e + p / rho
    dT = 1e-5 * T
    _, sTp, _ = _states(oc, mix, p, T + dT, Y)
    _, sTm, _ = _states(oc, mix, p, T - dT, Y)
    h = lambda s: s[:, 4] + s[:, 0] / s[:, 3]  # noqa: E731
    cp_fd = (h(sTp) - h(sTm)) / (2 * dT)
    assert np.abs(cp_fd / st[:, 5] - 1).max() < 1e-6
    # 1/c^2 = (drho/dp)_s = (drho/dp)_T - T / (rho^2 cp) (drho/dT)_p^2, both derivatives by central
    # differences of the EOS's own density
    dp = 1e-6 * p
    _, spp, _ = _states(oc, mix, p + dp, T, Y)
    _, spm, _ = _states(oc, mix, p - dp, T, Y)
    rp = (spp[:, 3] - spm[:, 3]) / (2 * dp)
    rT = (sTp[:, 3] - sTm[:, 3]) / (2 * dT)
    c2 = 1.0 / (rp - T / (st[:, 3] ** 2 * st[:, 5]) * rT**2)
    assert np.abs(c2 / st[:, 7] ** 2 - 1).max() < 1e-6
    # and gamma = c^2 (drho/dp)_T, as the state defines it
    assert np.abs(st[:, 6] / (st[:, 7] ** 2 * rp) - 1).max() < 1e-6


@pytest.mark.parametrize("eos", ["real-gas", "peng-robinson"])
def test_the_dilute_limit_is_the_thermally_perfect_gas(eos):
    p = np.full(6, 10.0)
    T = np.linspace(150.0, 900.0, 6)
    Y = np.linspace(0.1, 0.9, 6)
    real = _mixture(eos)
    ideal = _mixture("tpg")
    _, sr, _ = _states(O.OracleCase(real), real, p, T, Y)
    _, si, _ = _states(O.OracleCase(ideal), ideal, p, T, Y)
    for f in (3, 4, 5, 6, 7):
        assert np.abs(sr[:, f] / si[:, f] - 1).max() < 1e-5, f


@pytest.mark.parametrize("eos", ["real-gas", "peng-robinson"])
def test_partial_enthalpies_make_the_mixture_enthalpy(eos):
    """Euler's theorem for a homogeneous function of degree one: sum_n Y_n h_n = h."""
    mix = _mixture(eos, trans="chung")
    oc = O.OracleCase(mix)
    p, T, Y = _grid()
    hb, st, Q = _states(oc, mix, p, T, Y)
    n = len(p)
    tr, hs = np.zeros((4, 1, 1, n)), np.zeros((2, 1, 1, n))
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(tr), O.arr(hs), O.box((0, 0, 0), (n, 1, 1)))
    h = hs[:, 0, 0, :].T
    hmix = Y * h[:, 0] + (1 - Y) * h[:, 1]
    want = st[:, 4] + st[:, 0] / st[:, 3]
    assert np.abs(hmix - want).max() / np.abs(want).max() < 1e-10


def test_chung_tends_to_the_dilute_values():
    """At vanishing density Chung's factors are one: its transport is the kinetic theory's."""
    dense = _mixture("real-gas", trans="chung")
    dilute = _mixture("real-gas", trans="kinetic-theory")
    p, T, Y = np.full(5, 100.0), np.linspace(150.0, 900.0, 5), np.linspace(0.1, 0.9, 5)
    out = []
    for mix in (dense, dilute):
        oc = O.OracleCase(mix)
        hb, st, Q = _states(oc, mix, p, T, Y)
        tr, hs = np.zeros((4, 1, 1, 5)), np.zeros((2, 1, 1, 5))
        O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(tr), O.arr(hs), O.box((0, 0, 0), (5, 1, 1)))
        out.append(tr[:2, 0, 0, :])
    assert np.abs(out[0] / out[1] - 1).max() < 1e-4
    # and at 60 bar, 130 K the dense factors are well above one (liquid-like density)
    mix = dense
    oc = O.OracleCase(mix)
    hb, st, Q = _states(oc, mix, np.array([6e6]), np.array([130.0]), np.array([0.5]))
    tr, hs = np.zeros((4, 1, 1, 1)), np.zeros((2, 1, 1, 1))
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(tr), O.arr(hs), O.box((0, 0, 0), (1, 1, 1)))
    assert st[0, 3] > 300.0 and tr[0, 0, 0, 0] > 3e-5


def test_the_deterministic_cube_root_is_within_an_ulp():
    rng = np.random.default_rng(9)
    x = np.concatenate([np.exp(rng.uniform(-600, 600, 20000)), rng.uniform(1e-3, 10.0, 20000), -rng.uniform(0.1, 5, 2000)])
    got = np.array([O.dcbrt(v) for v in x])
    want = np.cbrt(x)
    assert (np.abs(got - want) <= np.spacing(np.abs(want))).all()
    assert O.dcbrt(0.0) == 0.0 and O.dcbrt(8.0) == 2.0 and O.dcbrt(-27.0) == -3.0


def _poling_chung(X, T, rho, cvR):
    """Chung's dense-gas factors for one nonpolar fluid as Poling, Prausnitz and O'Connell (5th ed.)
    state them, restated from the book and the species' critical constants alone (after
    tests/dense/test_chung.py:10-62): one-fluid mixing 9-5.24 to 9-5.44, the dense viscosity
    9-6.16 to 9-6.18 and conductivity 10-5.5, each as a ratio to its dilute limit. numpy's exp,
    power and cube root throughout; cvR the ideal cv/R of the mixture."""
    import yaml

    from peregrine_amd.mixture.species import DATA

    db = yaml.safe_load(open(DATA / "species.yaml"))
    tabs = yaml.safe_load(open(DATA / "chung.yaml"))
    names = ("O2", "N2")
    MW = np.array([sum(n * {"O": 15.999, "N": 14.007}[e] for e, n in db[s]["comp"].items()) for s in names])
    Vc = np.array([db[s]["Vcrit"] for s in names]) * MW * 1e3  # cm^3/mol
    Tcr = np.array([db[s]["Tcrit"] for s in names])
    w = np.array([db[s]["acentric"] for s in names])
    sigma = 0.809 * np.cbrt(Vc)  # 9-4.4: angstrom
    epsk = Tcr / 1.2593  # 9-4.5
    xx = np.outer(X, X)
    sij = np.sqrt(np.outer(sigma, sigma))  # 9-5.33
    eij = np.sqrt(np.outer(epsk, epsk))  # 9-5.35
    Mij = 2.0 * np.outer(MW, MW) / np.add.outer(MW, MW)  # 9-5.40
    s3 = (xx * sij**3).sum()  # 9-5.25
    em = (xx * eij * sij**3).sum() / s3  # 9-5.26
    Mm = ((xx * eij * sij**2 * np.sqrt(Mij)).sum() / (em * s3 ** (2.0 / 3.0))) ** 2  # 9-5.28
    wm = (xx * 0.5 * np.add.outer(w, w) * sij**3).sum() / s3  # 9-5.29
    Tc, Vcm = 1.2593 * em, s3 / 0.809**3  # 9-5.42, 9-5.43
    Fc = 1.0 - 0.2756 * wm  # 9-5.44 without the dipole and association terms
    E = np.array(tabs["viscosity"])[:, 0] + np.array(tabs["viscosity"])[:, 1] * wm  # Table 9-6
    B = np.array(tabs["conductivity"])[:, 0] + np.array(tabs["conductivity"])[:, 1] * wm  # Table 10-4
    Ts = T / em
    Om = 1.16145 * Ts**-0.14874 + 0.52487 * np.exp(-0.77320 * Ts) + 2.16178 * np.exp(-2.43787 * Ts)  # 9-4.3
    y = rho / (Mm * 1e3) * Vcm / 6.0  # 9-6.18: rho in mol/cm^3
    G1 = (1.0 - 0.5 * y) / (1.0 - y) ** 3

    def G2(c):
        return (c[0] * (1.0 - np.exp(-c[3] * y)) / y + c[1] * G1 * np.exp(c[4] * y) + c[2] * G1) / (
            c[0] * c[3] + c[1] + c[2])

    g = G2(E)
    eta2 = E[6] * y**2 * g * np.exp(E[7] + E[8] / Ts + E[9] / Ts**2)
    # eta = eta* 36.344 sqrt(M Tc) / Vc^2/3 with eta* = sqrt(T*)/Omega Fc (1/G2 + E6 y) + eta**; over
    # the dilute eta0 = 40.785 Fc sqrt(M T) / (Vc^2/3 Omega), which is its y -> 0 limit
    muratio = (1.0 / g + E[5] * y) + eta2 * Om / (np.sqrt(Ts) * Fc)
    eta0 = 4.0785e-6 * np.sqrt(Mm * T) / (Vcm ** (2.0 / 3.0) * Om) * Fc  # 9-4.10, Pa s
    alpha = cvR - 1.5
    beta = 0.7862 - 0.7109 * wm + 1.3168 * wm**2
    Zr = 2.0 + 10.5 * (T / Tc) ** 2
    Psi = 1.0 + alpha * ((0.215 + 0.28288 * alpha - 1.061 * beta + 0.26665 * Zr) / (0.6366 + beta * Zr
                                                                                     + 1.061 * alpha * beta))
    lam0 = 31.2 * eta0 * Psi / (Mm * 1e-3)  # 10-3.14
    h = G2(B)
    kratio = (1.0 / h + B[5] * y) + 3.586e-3 * np.sqrt(Tc / (Mm * 1e-3)) / Vcm ** (2.0 / 3.0) * B[6] * y**2 * h * \
        np.sqrt(T / Tc) / lam0  # 10-5.5
    return muratio, kratio


@pytest.mark.parametrize("eos", ["real-gas", "peng-robinson"])
def test_chungs_dense_factors_are_polings(eos):
    """The oracle's Chung viscosity and conductivity over its kinetic-theory values at the same
    dense states (60 bar, 120 to 300 K, O2/N2 from pure to pure, liquid-like to gas-like density)
    equal Poling's factors restated above. A wrong E or B row, a G2 of the wrong form or a swapped
    T*/Tr fails by far more than the tolerance."""
    dense = _mixture(eos, trans="chung")
    dilute = _mixture(eos, trans="kinetic-theory")
    T = np.tile(np.linspace(120.0, 300.0, 7), 5)
    Y = np.repeat(np.array([1.0, 0.8, 0.5, 0.2, 0.0]), 7)
    p = np.full(T.size, 6e6)
    n = T.size
    trs = []
    for mix in (dense, dilute):
        oc = O.OracleCase(mix)
        hb, st, Q = _states(oc, mix, p, T, Y)
        tr, hs = np.zeros((4, 1, 1, n)), np.zeros((2, 1, 1, n))
        O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(tr), O.arr(hs), O.box((0, 0, 0), (n, 1, 1)))
        trs.append(tr[:2, 0, 0, :])
    rho = st[:, 3]
    assert rho.max() > 900.0 and rho.min() < 100.0  # liquid-like and gas-like states both
    k = dense.constants()
    MW = k["MW"]
    got_mu, got_k = trs[0][0] / trs[1][0], trs[0][1] / trs[1][1]
    for m in range(n):
        Yv = np.array([Y[m], 1.0 - Y[m]])
        X = Yv / MW / (Yv / MW).sum()
        cvR = sum(X[s] * (np.polyval(k["cppoly"][s][::-1], np.log(T[m])) - 1.0) for s in range(2))
        mur, kr = _poling_chung(X, T[m], rho[m], cvR)
        assert mur > 1.2 or rho[m] < 200.0
        assert got_mu[m] == pytest.approx(mur, rel=1e-12), (m, T[m], Y[m])
        assert got_k[m] == pytest.approx(kr, rel=1e-12), (m, T[m], Y[m])
