"""The oracle's viscous terms against what the mathematics fixes: gradients and viscous fluxes
of linear fields in closed form (tests/navstokes/test_gradients.py:32,
tests/navstokes/test_diffflux.py:60-140), Wilke's rule and the conductivity mean in their
textbook forms, the species enthalpies against the mixture EOS, and conservation of a periodic
viscous right-hand side."""

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd.mixture.constants import Ru
from peregrine_amd.simulator import get_simulator


def _block(gas, ncell=(6, 5, 4), lengths=(0.6, 0.45, 0.3), trans="constant", jostle=0.0):
    cfg = sp.config(gas, trans=trans)
    sim = get_simulator(cfg)
    hb = sp.HostBlock(ncell, 1, sim.ne, 2, lengths=lengths, jostle=jostle, seed=7)
    return sim, O.OracleCase(sim.mixture), hb


def _linear_prims(hb, ne, A, gT, g0, gY=None, rho_wave=True):
    """p, u = A x, T = T0 + gT . x, Y = Y0 + gY . x at the cell centres (halos included)."""
    x = hb.xcc
    q = np.zeros(hb.cshape + (ne,))
    # a varying density, so the velocity gradient is a ratio of two varying conserved values
    q[..., 0] = 101325.0 * (1.0 + (0.2 * np.sin(2.0 * x[..., 0] + x[..., 1]) if rho_wave else 0.0))
    for r in range(3):
        q[..., 1 + r] = g0[r] + x @ A[r]
    q[..., 4] = 400.0 + x @ gT
    if gY is not None:
        q[..., 5] = 0.4 + x @ gY
    return q


def _fill(oc, hb, q):
    full = O.box((0, 0, 0), hb.cshape)
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), full)


def test_gradients_are_exact_for_linear_fields():
    sim, oc, hb = _block("binary")
    A = np.array([[3.0, -1.0, 0.5], [2.0, 0.25, -4.0], [-1.5, 6.0, 1.0]])
    gT, gY = np.array([40.0, -25.0, 10.0]), np.array([0.05, -0.02, 0.08])
    q = _linear_prims(hb, sim.ne, A, gT, (10.0, -5.0, 2.0), gY)
    _fill(oc, hb, q)
    O.lib.pgo_grads(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.Jinv), O.arr(hb.grads),
                    O.box((1, 1, 1), hb.ncell))
    g = hb.grads[hb.interior()].reshape(sim.ne - 1, 3, -1)
    want = [A[0], A[1], A[2], gT, gY]
    for row, w in enumerate(want):
        err = np.abs(g[row] - w[:, None]).max() / np.abs(w).max()
        assert err < 1e-12, (row, err)


def _face_flux(sim, oc, hb, axis, alpha=1.0):
    ng = hb.ng
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp),
                        O.box((0, 0, 0), hb.cshape))
    O.lib.pgo_grads(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.Jinv), O.arr(hb.grads),
                    O.box((ng,) * 3, hb.ncell))
    ext = list(hb.ncell)
    ext[axis] += 1
    F = np.zeros_like(hb.F[axis])
    O.lib.pgo_diffflux(oc.ref, axis, alpha, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp),
                       O.arr(hb.grads), O.arr(hb.xccd), O.arr(hb.xf[axis]), O.arr(hb.S[axis]), O.arr(F),
                       O.box((ng,) * 3, ext))
    return F


def _wilke(X, mu, MW):
    """Wilke's rule as textbooks state it: phi_ij = [1 + (mu_i/mu_j)^1/2 (M_j/M_i)^1/4]^2 / [8 (1 + M_i/M_j)]^1/2."""
    ns = len(mu)
    phi = np.array([[(1 + np.sqrt(mu[i] / mu[j]) * (MW[j] / MW[i]) ** 0.25) ** 2 / np.sqrt(8 * (1 + MW[i] / MW[j]))
                     for j in range(ns)] for i in range(ns)])
    return sum(X[i] * mu[i] / sum(X[j] * phi[i, j] for j in range(ns)) for i in range(ns))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_viscous_flux_is_newtonian_fourier_fick_for_linear_fields(axis):
    """Two species of equal Lewis number: with linear u, T and Y both cells extrapolate to one face
    value, so the damping term vanishes and the face gradients are the fields' own; the flux is
    then Newton's stress, Fourier's heat flux and Fick's law (no correction velocity with equal
    diffusivities) with the stress work and the species enthalpy flux, the transport
    coefficients being the means of the two cells' (Wilke and the conductivity mean, textbook
    forms, at the cell centres)."""
    species = {n: dict(v, lewis=1.2) for n, v in sp.BINARY_CPG.items()}
    cfg = sp.config("binary", trans="constant")
    cfg["mixture"]["species"] = species
    sim = get_simulator(cfg)
    oc = O.OracleCase(sim.mixture)
    hb = sp.HostBlock((7, 6, 5), 1, sim.ne, 2, lengths=(0.7, 0.48, 0.3))
    A = np.array([[3.0, -1.0, 0.5], [2.0, 0.25, -4.0], [-1.5, 6.0, 1.0]])
    gT, gY, u0 = np.array([40.0, -25.0, 10.0]), np.array([0.05, -0.02, 0.08]), np.array([10.0, -5.0, 2.0])
    q = _linear_prims(hb, sim.ne, A, gT, u0, gY, rho_wave=False)
    _fill(oc, hb, q)
    F = _face_flux(sim, oc, hb, axis)
    mix = sim.mixture.constants()
    mu, kap, MW = mix["mu0"], mix["kappa0"], mix["MW"]
    st = sp.to_host(hb.state)
    x = hb.xcc
    Y0 = q[..., 5]
    Y = np.stack([Y0, 1.0 - Y0])
    X = Y / MW[:, None, None, None]
    X = X / X.sum(axis=0)
    muc = _wilke(X, mu, MW)
    kapc = 0.5 * (sum(X[i] * kap[i] for i in range(2)) + 1.0 / sum(X[i] / kap[i] for i in range(2)))
    Dc = kapc / (st[..., 3] * st[..., 5] * 1.2)
    h = sp.to_host(hb.hsp)
    ng, n = hb.ng, hb.ncell
    # faces one in from each end along the axis: both neighbours hold gradients
    idx = [np.arange(ng, ng + m) for m in n]
    idx[axis] = np.arange(ng + 1, ng + n[axis])
    I, J, K = np.meshgrid(*idx, indexing="ij")
    d = [int(axis == 0), int(axis == 1), int(axis == 2)]
    L = (I - d[0], J - d[1], K - d[2])
    R = (I, J, K)
    Sf = sp.to_host(hb.S[axis])[R]
    xf = sp.to_host(hb.xf[axis])[R]
    got = sp.to_host(F)[R]

    def mean(a):
        return 0.5 * (a[L] + a[R])

    muf, kapf, Df = mean(muc), mean(kapc), mean(Dc)
    rhof = mean(st[..., 3])
    uf = xf @ A.T + u0
    tau = -muf[..., None, None] * (A + A.T) + (2.0 / 3.0) * (muf * np.trace(A))[..., None, None] * np.eye(3)
    mom = np.einsum("...re,...e->...r", tau, Sf)
    err = np.abs(got[..., 1:4] - mom).max() / np.abs(mom).max()
    assert err < 1e-12, err
    J0 = -rhof * Df * (Sf @ gY)
    err = np.abs(got[..., 5] - J0).max() / np.abs(J0).max()
    assert err < 1e-12, err
    energy = np.einsum("...r,...r->...", uf, mom) - kapf * (Sf @ gT) + J0 * (mean(h[..., 0]) - mean(h[..., 1]))
    err = np.abs(got[..., 4] - energy).max() / np.abs(energy).max()
    assert err < 1e-12, err
    assert np.abs(got[..., 0]).max() == 0.0


def test_wilke_and_conductivity_mean_are_the_textbook_forms():
    sim, oc, hb = _block("binary", ncell=(4, 3, 3))
    rng = np.random.default_rng(3)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = 101325.0
    q[..., 4] = rng.uniform(300, 900, hb.cshape)
    q[..., 5] = rng.uniform(0.0, 1.0, hb.cshape)
    _fill(oc, hb, q)
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp),
                        O.box((0, 0, 0), hb.cshape))
    mix = sim.mixture.constants()
    mu, kap, MW, Le = mix["mu0"], mix["kappa0"], mix["MW"], mix["lewis"]
    Y = np.stack([q[..., 5], 1.0 - q[..., 5]])
    X = Y / MW[:, None, None, None]
    X = X / X.sum(axis=0)
    phi = np.array([[(1 + np.sqrt(mu[i] / mu[j]) * (MW[j] / MW[i]) ** 0.25) ** 2 / np.sqrt(8 * (1 + MW[i] / MW[j]))
                     for j in range(2)] for i in range(2)])
    mum = sum(X[i] * mu[i] / sum(X[j] * phi[i, j] for j in range(2)) for i in range(2))
    kapm = 0.5 * (sum(X[i] * kap[i] for i in range(2)) + 1.0 / sum(X[i] / kap[i] for i in range(2)))
    tr = sp.to_host(hb.trstate)
    st = sp.to_host(hb.state)
    assert np.abs(tr[..., 0] / mum - 1).max() < 1e-13
    assert np.abs(tr[..., 1] / kapm - 1).max() < 1e-13
    # Lewis: D_n = kappa / (rho cp Le_n)
    for n in range(2):
        D = kapm / (st[..., 3] * st[..., 5] * Le[n])
        assert np.abs(tr[..., 2 + n] / D - 1).max() < 1e-13


@pytest.mark.parametrize("gas", ["h2o2", "binary", "air-tpg"])
def test_species_enthalpies_make_the_mixture_enthalpy(gas):
    """sum_n Y_n h_n(T) = e + p / rho: the enthalpy table the viscous flux reads is the EOS's."""
    # the mechanism's species carry no constant transport data: H2/O2 borrows the enthalpy table
    # alone, with placeholder transport constants the enthalpies do not read
    sim = get_simulator(sp.config(gas, trans=None if gas == "h2o2" else "constant"))
    mixture = sim.mixture
    oc = O.OracleCase(mixture)
    if gas == "h2o2":
        # the hsp table needs only the caloric constants: borrow them without a transport model
        k = mixture.constants()
        keep = [np.ascontiguousarray(k["hpoly"]), np.ascontiguousarray(np.ones(mixture.ns))]
        oc.keep += keep
        oc.c.hpoly = keep[0].ctypes.data_as(O._dp)
        oc.c.Ru = Ru
        oc.c.trans = 0
        ones = np.ones(mixture.ns)
        for name in ("mu0", "kappa0", "lewis", "MWqinv"):
            a = np.ascontiguousarray(ones)
            oc.keep.append(a)
            setattr(oc.c, name, a.ctypes.data_as(O._dp))
        w = np.ascontiguousarray(np.full((mixture.ns, mixture.ns), 0.25))
        oc.keep.append(w)
        oc.c.wilkepair = w.ctypes.data_as(O._dp)
    hb = sp.HostBlock((5, 4, 3), 1, sim.ne, 2)
    rng = np.random.default_rng(11)
    q = np.zeros(hb.cshape + (sim.ne,))
    q[..., 0] = 101325.0 * rng.uniform(0.5, 2.0, hb.cshape)
    q[..., 4] = rng.uniform(350.0, 2500.0, hb.cshape)
    if sim.ns > 1:
        Yr = rng.random(hb.cshape + (sim.ns,))
        Yr /= Yr.sum(axis=-1, keepdims=True)
        q[..., 5:] = Yr[..., :-1]
    _fill(oc, hb, q)
    O.lib.pgo_transport(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.arr(hb.trstate), O.arr(hb.hsp),
                        O.box((0, 0, 0), hb.cshape))
    st = sp.to_host(hb.state)
    h = sp.to_host(hb.hsp)
    Y = np.concatenate([q[..., 5:], 1.0 - q[..., 5:].sum(axis=-1, keepdims=True)], axis=-1)
    hmix = (Y * h).sum(axis=-1)
    want = st[..., 4] + st[..., 0] / st[..., 3]
    # the enthalpy polynomial and the EOS's recurrence on the mixture's cp coefficients are one
    # integral written two ways; their terms cancel by six orders (measured 1.2e-11 for air)
    assert np.abs(hmix - want).max() / np.abs(want).max() < 5e-11


def test_a_periodic_viscous_rhs_conserves_mass_momentum_and_energy():
    """The divergence of face fluxes telescopes: on a periodic block the volume-weighted sum of
    dQ over the cells is zero for every conserved row, viscous terms included."""
    sim, oc, hb = _block("binary", ncell=(6, 5, 4), lengths=(1.0, 0.8, 0.6), jostle=0.2)
    q = sp.smooth_prims(hb.xcc, sim.ne, ns=2)
    _fill(oc, hb, q)
    b = sp.oracle_block(hb, "rk1", False, viscous=True, alpha=1.0)
    # one stage: prerhs (exchange, halo EOS, transport), rhs (gradients and their exchange,
    # advective and viscous fluxes, divergence)
    dQ = hb.regs[-1]
    Q0 = hb.regs[0].copy()
    O.lib.pgo_stage(oc.ref, O.C.byref(b), 0, 1e-7)
    dQsum = ((hb.regs[0] - Q0)[hb.interior()] / hb.detJinv[hb.interior()]).reshape(sim.ne, -1).sum(axis=1)
    scale = (np.abs(dQ[hb.interior()]) / hb.detJinv[hb.interior()]).reshape(sim.ne, -1).sum(axis=1) * 1e-7
    assert (np.abs(dQsum) <= 1e-12 * scale + 1e-300).all(), dQsum / scale


def test_collision_integrals_reproduce_their_table():
    """The spline in ln T* passes through every tabulated point of each delta* column, and the
    quadratic in delta* returns a column's value at its own delta*."""
    from peregrine_amd.mixture.integrals import CollisionIntegrals
    from peregrine_amd.mixture.species import database

    ci = CollisionIntegrals()
    t = database("integrals")
    ts = np.array(t["tstar22"], float)
    om = np.array(t["omega22"], float)
    for col in (0, 3, 7):
        got = ci.omega22(ts, np.full_like(ts, t["delta"][col]))
        assert np.abs(got / om[:, col] - 1).max() < 1e-12


def test_kinetic_theory_gives_the_measured_gas_properties():
    """Chapman-Enskog viscosity and the Eucken-corrected conductivity of the mechanism's gases at
    300 K against measured values (viscosity: N2 17.9, O2 20.7, H2 8.96 uPa s; conductivity:
    N2 25.9, O2 26.3, H2 186 mW/m/K; CRC Handbook), within 3%: the fits' agreement with the
    reference's own is not pinned beyond re-running the same algorithm."""
    cfg = sp.config("h2o2", trans="kinetic-theory")
    sim = get_simulator(cfg)
    mix = sim.mixture
    k = mix.constants()
    lnT = np.log(300.0)
    names = mix.spnames
    from peregrine_amd.mixture.fits import evaluate

    for name, mu_meas, kap_meas in (("N2", 17.9e-6, 25.9e-3), ("O2", 20.7e-6, 26.3e-3), ("H2", 8.96e-6, 186e-3)):
        n = names.index(name)
        mu = (evaluate(k["mupoly"][n], lnT) * 300.0**0.25) ** 2
        kap = evaluate(k["kappapoly"][n], lnT) * np.sqrt(300.0)
        assert abs(mu / mu_meas - 1) < 0.03, (name, mu)
        assert abs(kap / kap_meas - 1) < 0.03, (name, kap)


def test_the_viscous_sponge_ramps_the_viscosity_along_its_line():
    """plugins/simulator/kernels/viscsponge.mako: before the line's origin the viscosity is as the
    transport made it, past its end multiplier times it, linear in between (halfway: 1 + (m - 1)/2),
    whatever the distance across the line; the other transport fields untouched."""
    from oracle import oracle as O

    n = 21
    x = np.zeros((n, 3, 2, 3))
    x[..., 0] = np.linspace(-1.0, 3.0, n)[:, None, None]  # along the line, 0.2 apart
    x[..., 1] = np.linspace(-5.0, 5.0, 3)[None, :, None]  # across it
    x[..., 2] = np.array([0.0, 7.0])[None, None, :]
    rng = np.random.default_rng(3)
    # the same transport state across the line, so the cells across it can be compared bit for bit
    tr = np.broadcast_to(rng.uniform(1e-5, 2e-5, (n, 1, 1, 4)), (n, 3, 2, 4)).copy()
    got = sp.to_dev(tr)
    # from x = 0 to x = 2 along +x, eightfold
    O.lib.pgo_viscsponge(O.arr(sp.to_dev(x)), O.arr(got), O.dbls([1.0, 0.0, 0.0, 0.0, 2.0, 8.0]),
                         O.box((0, 0, 0), (n, 3, 2)))
    got = sp.to_host(got)
    factor = got[..., 0] / tr[..., 0]
    s = x[..., 0] / 2.0
    assert np.array_equal(got[s <= 0.0], tr[s <= 0.0])
    assert np.array_equal(got[..., 0][s >= 1.0], tr[..., 0][s >= 1.0] * 8.0)
    mid = np.isclose(s, 0.5)
    assert mid.any() and np.array_equal(got[..., 0][mid], tr[..., 0][mid] * 4.5)
    assert np.allclose(factor, 1.0 + np.clip(s * 7.0, 0.0, 7.0), rtol=1e-15, atol=0)
    # across the line it is the same
    assert np.array_equal(got[..., 0], np.broadcast_to(got[:, :1, :1, 0], got.shape[:3]))
    assert np.array_equal(got[..., 1:], tr[..., 1:])
