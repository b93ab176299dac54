"""Toro's shock tubes on a walled line against the exact solution, after the reference's own
tests (tests/simulator/test_riemann.py:6-33, tests/simulator/test_flux.py:83-126,
tests/simulator/support.py:144-195): the exact solver reproduces Toro's star states (his Table
4.3), and each scheme's mean density error over the wave fan is the one the reference measured for
the same scheme on the same tube (to the four digits it printed; it locks a quarter above), with the
walls passing no mass or energy and the plateau's p and u at the star state. The oracle runs the tubes; test_gpu_parity.py steps
one on the device against it."""

import ctypes

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd.config import Config
from peregrine_amd.simulator import get_simulator
from riemann_exact import Riemann

Ru = 8314.46261815324
# the reference tubes' gas: constants only, one species
AIR = {"Air": {"MW": 28.97, "cp0": 1000.0}}
R_AIR = Ru / 28.97
GAMMA = 1000.0 / (1000.0 - R_AIR)

SCHEMES = {
    "rusanov": dict(primary="rusanov", recon="constant", limiter=None),
    "hllc": dict(primary="hllc", recon="constant", limiter=None),
    "muscl-koren-rusanov": dict(primary="rusanov", recon="muscl", limiter="koren"),
    "muscl-koren-hllc": dict(primary="hllc", recon="muscl", limiter="koren"),
}

# per tube the length, the cells per unit length and the reference's measured mean density error
# over the fan for each scheme (tests/simulator/test_flux.py:86-93)
TUBES = {
    0: dict(length=1.0, cells=200, rho={"rusanov": 5.242e-2, "hllc": 3.822e-2, "muscl-koren-rusanov": 9.576e-3,
                                        "muscl-koren-hllc": 7.295e-3}),
    1: dict(length=2.0, cells=200, rho={"rusanov": 6.451e-2, "hllc": 3.410e-2, "muscl-koren-rusanov": 1.280e-2,
                                        "muscl-koren-hllc": 8.921e-3}),
    2: dict(length=2.5, cells=200, rho={"rusanov": 6.984e-2, "hllc": 7.255e-2, "muscl-koren-rusanov": 7.062e-3,
                                        "muscl-koren-hllc": 7.066e-3}),
    3: dict(length=1.0, cells=400, rho={"rusanov": 1.935e-1, "hllc": 1.606e-1, "muscl-koren-rusanov": 7.468e-2,
                                        "muscl-koren-hllc": 8.740e-2}),
    4: dict(length=3.0, cells=200, rho={"rusanov": 1.238e-1, "hllc": 9.191e-2, "muscl-koren-rusanov": 3.490e-2,
                                        "muscl-koren-hllc": 2.980e-2}),
    5: dict(length=3.5, cells=200, rho={"rusanov": 2.116e-1, "hllc": 6.203e-2}),
}
# the plateau's p and u against the star state (tests/simulator/test_flux.py:95)
PLATEAU = {0: dict(constant=1e-2, muscl=2e-3), 1: dict(constant=5e-3, muscl=2e-3), 3: dict(constant=2e-2, muscl=2e-2),
           4: dict(constant=2e-2, muscl=2e-2), 5: dict(constant=2e-2, muscl=2e-2)}


def test_the_star_states_are_toros():
    """Toro's Table 4.3: p*, u*, and the densities either side of the contact."""
    table = {0: (0.30313, 0.92745, 0.42632, 0.26557), 2: (0.00189, 0.0, 0.02185, 0.02185),
             3: (460.894, 19.5975, 0.57506, 5.99924), 4: (1691.64, 8.68975, 14.2823, 31.0426),
             5: (460.894, 0.0, 0.57506, 5.99924)}
    for case, (pM, uM, rhoL, rhoR) in table.items():
        r = Riemann.toro(case)
        p, u = r.star()
        assert p == pytest.approx(pM, rel=3e-3 if case == 2 else 1e-4)
        assert u == pytest.approx(uM, rel=1e-4, abs=1e-4)
        assert r.sample(p, u, u - 1e-9)[2] == pytest.approx(rhoL, rel=1e-4)
        assert r.sample(p, u, u + 1e-9)[2] == pytest.approx(rhoR, rel=1e-4)


def test_sods_waves_and_rarefaction():
    r = Riemann.toro(0)
    x = np.linspace(0.0, 1.0, 2001)
    sol = r.solve(x)
    pM, uM = r.star()
    contact = r.x0 + uM * r.t
    jumps = x[np.where(np.abs(np.diff(sol["rho"])) > 0.02)[0]]
    assert len(jumps) == 2 and abs(contact - 0.68549) < 1e-5
    # isentropic from the undisturbed left state through the fan to the contact
    entropy = sol["p"] / sol["rho"] ** r.gamma
    assert np.allclose(entropy[x < contact - 1e-3], r.pL / r.rhoL**r.gamma, rtol=1e-12, atol=0.0)


def tube(case, scheme, walls=True):
    """The tube on the oracle: a line of cells along i, one cell thick in j and k, slip walls all
    round; returns the line's x, rho, u, p, the problem, and the mass and energy moved."""
    spec = TUBES[case]
    r = Riemann.toro(case, gamma=GAMMA, R=R_AIR)
    c = Config()
    c["simulation"]["simulator"] = "euler"
    c["mixture"]["species"], c["mixture"]["eos"] = AIR, "cpg"
    a = c["adv-flux"]
    a["primary"], a["recon"], a["limiter"] = SCHEMES[scheme]["primary"], SCHEMES[scheme]["recon"], SCHEMES[scheme]["limiter"]
    c.validate()
    sim = get_simulator(c)
    oc = O.OracleCase(sim.mixture, a["recon"], a["limiter"], a["primary"])
    n = int(round(spec["length"] * spec["cells"]))
    hb = sp.HostBlock((n, 1, 1), sim.ng, sim.ne, 3, lengths=(spec["length"], 0.1, 0.1))
    x = hb.xcc[..., 0] + (r.x0 - 0.5 * spec["length"])
    left = x <= r.x0
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = np.where(left, r.pL, r.pR)
    q[..., 1] = np.where(left, r.uL, r.uR)
    q[..., 4] = np.where(left, r.TL, r.TR)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))
    # as the solver's set_prims ends (multiblock/solver.py:298-305): the EOS over the interior
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box((hb.ng,) * 3, hb.ncell))
    b = sp.oracle_block(hb, "tvd-rk3", False, walls=(0, 1, 2, 3, 4, 5))
    inner = hb.interior()
    started = hb.regs[0][inner].reshape(sim.ne, -1).sum(axis=1)
    for _ in range(int(round(r.t / r.dt))):
        O.lib.pgo_step(oc.ref, ctypes.byref(b), r.dt)
    Q = hb.regs[0][inner]
    g = hb.ng
    moved = Q.reshape(sim.ne, -1).sum(axis=1)[[0, 4]] / started[[0, 4]] - 1.0
    line = np.s_[0, 0, :]
    rho, u = Q[0][line], Q[1][line] / Q[0][line]
    p = hb.state[0][g:-g, g:-g, g:-g][line]
    return x[g:-g, g, g], rho, u, p, r, np.abs(moved)


def errors(x, rho, u, p, r):
    """The reference's measures (tests/simulator/support.py:164-194): the mean density error over the
    fan widened by a tenth, and the plateau's p and u against the star state."""
    g = r.gamma
    pM, uM = r.star()
    sL, sR = r.heads()
    head = (r.x0 + sL * r.t, r.x0 + sR * r.t)
    assert x[0] < head[0] and head[1] < x[-1]
    lo = max(head[0] - 0.1 * (head[1] - head[0]), x[0])
    hi = min(head[1] + 0.1 * (head[1] - head[0]), x[-1])
    fan = (x > lo) & (x < hi)
    exact = r.solve(x)
    rho_err = np.abs(rho - exact["rho"])[fan].mean() / np.abs(exact["rho"][fan]).mean()
    g1 = (g - 1.0) / (2.0 * g)
    edge = sR if pM > r.pR else uM + r.cR * (pM / r.pR) ** g1
    a, b = r.x0 + uM * r.t, r.x0 + edge * r.t
    plateau = (x > a + 0.3 * (b - a)) & (x < b - 0.3 * (b - a))
    p_err = np.abs(p[plateau] / pM - 1.0).max()
    u_err = np.abs(u[plateau] - uM).max() / max(abs(uM), abs(r.uL), abs(r.uR))
    return rho_err, p_err, u_err


def runs():
    for case, spec in TUBES.items():
        for scheme in spec["rho"]:
            yield pytest.param(case, scheme, id=f"toro{case}-{scheme}")


@pytest.mark.parametrize("case,scheme", list(runs()))
def test_the_tube_against_the_exact_solution(case, scheme):
    x, rho, u, p, r, moved = tube(case, scheme)
    assert np.isfinite(rho).all() and np.isfinite(p).all()
    # the walls pass nothing: the interior mass and energy are kept
    assert moved.max() < 1e-12
    rho_err, p_err, u_err = errors(x, rho, u, p, r)
    print(f"toro{case} {scheme}: rho {rho_err:.3e} (reference {TUBES[case]['rho'][scheme]:.3e}), p {p_err:.2e}, u {u_err:.2e}")
    # the reference locks its measure a quarter above what it measured; this restatement of the
    # same schemes lands on the measured value itself, to the four digits it was printed with
    assert rho_err <= 1.25 * TUBES[case]["rho"][scheme]
    assert rho_err == pytest.approx(TUBES[case]["rho"][scheme], rel=1e-3)
    if case == 2:
        assert u_err < 5e-2
        return
    within = PLATEAU[case]["muscl" if "muscl" in scheme else "constant"]
    assert p_err < within and u_err < within


# --- the density wave once around a periodic line (tests/simulator/test_flux.py:26-48,
# tests/simulator/support.py:118-138): measured (max order, mean order, e80) for each scheme

WAVE_SCHEMES = dict(SCHEMES, kepaec=dict(primary="kepaec", recon="constant", limiter=None))
ORDERS = [
    ("kepaec", 1.95, 1.95, 3.872e-3),
    ("muscl-koren-rusanov", 1.6, 2.0, 1.298e-2),
    ("muscl-koren-hllc", 1.5, 2.0, 6.922e-3),
    ("rusanov", 0.3, 0.3, 4.069e-1),
    ("hllc", 0.7, 0.7, 1.312e-1),
]


def wave_errors(scheme, nx, u=100.0):
    """An entropy wave rho = 1.2 + 0.6 sin(2 pi x) at 1 atm advected once around a periodic line of
    nx - 1 cells (one cell thick, slip walls across), by tvd-rk3 at a tenth of the acoustic CFL;
    the max and mean |rho - rho_0| along the line."""
    c = Config()
    c["simulation"]["simulator"] = "euler"
    c["mixture"]["species"], c["mixture"]["eos"] = AIR, "cpg"
    a = c["adv-flux"]
    for key in ("primary", "recon", "limiter"):
        a[key] = WAVE_SCHEMES[scheme][key]
    c.validate()
    sim = get_simulator(c)
    oc = O.OracleCase(sim.mixture, a["recon"], a["limiter"], a["primary"])
    hb = sp.HostBlock((nx - 1, 1, 1), sim.ng, sim.ne, 3, lengths=(1.0, 0.1, 0.1))
    xc = hb.xcc[..., 0]
    rho = 1.2 + 0.6 * np.sin(2 * np.pi * xc)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0], q[..., 1] = 101325.0, u
    q[..., 4] = q[..., 0] / (rho * R_AIR)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box((hb.ng,) * 3, hb.ncell))
    b = sp.oracle_block(hb, "tvd-rk3", False, walls=(2, 3, 4, 5))
    cmax = np.sqrt(GAMMA * q[..., 0].max() / rho.min())
    dx = 1.0 / (nx - 1)
    n = int(round(1.0 / (abs(u) * 0.1 * dx / (abs(u) + cmax))))
    dt = 1.0 / (abs(u) * n)
    for _ in range(n):
        O.lib.pgo_step(oc.ref, ctypes.byref(b), dt)
    g = hb.ng
    error = np.abs(hb.regs[0][0][g, g, g:-g] - rho[g:-g, g, g])
    return error.max(), error.mean()


@pytest.mark.parametrize("scheme,atleast,mean_atleast,e80", ORDERS, ids=[o[0] for o in ORDERS])
def test_the_order_on_the_wave(scheme, atleast, mean_atleast, e80):
    e40, e80got = wave_errors(scheme, 41), wave_errors(scheme, 81)
    print(f"wave {scheme}: orders {np.log2(e40[0] / e80got[0]):.2f} {np.log2(e40[1] / e80got[1]):.2f}, "
          f"e80 {e80got[0]:.4e} (reference {e80:.3e})")
    assert np.log2(e40[0] / e80got[0]) >= atleast
    assert np.log2(e40[1] / e80got[1]) >= mean_atleast
    # the reference locks e80 a fifth above its measurement; the same schemes land on it
    assert e80got[0] == pytest.approx(e80, rel=1e-3)
