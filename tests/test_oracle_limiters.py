"""The oracle's five MUSCL limiters, pinned by what fixes them independently of any code here:
Sweby's phi(r) for each (the textbook forms the reference's own test oracle writes out,
tests/simulator/oracle.py:39-60), the symmetry phi(r)/r = phi(1/r) four of them have and Koren's
does not, the TVD region, and on a line of cells (after tests/simulator/test_recon.py:78-120)
linear exactness, two-cell boundedness and the Sweby form of each reconstructed side.

A swapped operand (phi(1/r) for phi(r)), a dropped factor of two or a mislabelled limiter fails
one of these: the numpy phi's below are the textbook's, not a retyping of pgo.c."""

import numpy as np
import pytest

import support as sp
from oracle import oracle as O
from peregrine_amd.simulator import get_simulator

O.lib.pgo_limit.restype = O.C.c_double
O.lib.pgo_limit.argtypes = [O.C.c_int, O.C.c_double, O.C.c_double]
O.lib.pgo_sides.argtypes = [O.C.POINTER(O.Case), O.Arr, O.Arr, O.C.c_int, O.C.c_int, O.C.c_int, O.C.c_int,
                            O.C.POINTER(O.C.c_double)]


# Sweby's phi of the slope ratio r (Sweby 1984; Koren 1993; van Leer 1974; Roe 1986; LeVeque,
# Finite Volume Methods, eq. 6.39-6.40)
def koren(r):
    return np.maximum(0.0, np.minimum(np.minimum(2.0 * r, (2.0 + r) / 3.0), 2.0))


def minmod(r):
    return np.maximum(0.0, np.minimum(1.0, r))


def mc(r):
    return np.maximum(0.0, np.minimum(np.minimum(2.0 * r, 0.5 * (1.0 + r)), 2.0))


def superbee(r):
    return np.maximum(0.0, np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0)))


def vanleer(r):
    return (r + np.abs(r)) / (1.0 + np.abs(r))


PHI = {"koren": koren, "minmod": minmod, "mc": mc, "superbee": superbee, "vanleer": vanleer}


def limit(name, a, b):
    return np.array([O.lib.pgo_limit(O.LIMITER[name], x, y) for x, y in zip(a, b)])


def _ratios():
    rng = np.random.default_rng(11)
    r = np.concatenate([rng.uniform(-4, 4, 3000), np.exp(rng.uniform(-8, 8, 2000)), [0.0, 0.5, 1.0, 2.0, 3.0, -1.0]])
    b = np.exp(rng.uniform(-20, 20, r.size))
    return r, b


@pytest.mark.parametrize("name", sorted(PHI))
def test_the_limiter_is_b_times_swebys_phi_of_a_over_b(name):
    r, b = _ratios()
    a = r * b
    got = limit(name, a, b)
    want = b * PHI[name](a / b)
    # a/b differs from r by an ulp; each form rounds a few times
    assert np.allclose(got, want, rtol=4e-16 * 8, atol=0.0)


@pytest.mark.parametrize("name", sorted(PHI))
def test_the_limiter_lies_in_swebys_tvd_region(name):
    r, b = _ratios()
    phi = limit(name, r * b, b) / b
    # 0 <= phi <= min(2r, 2) for r > 0, phi = 0 for r <= 0 (Sweby 1984)
    pos = r > 0
    assert (phi[~pos] == 0.0).all()
    assert (phi[pos] >= 0.0).all() and (phi[pos] <= np.minimum(2.0 * r[pos], 2.0) * (1 + 1e-15)).all()
    # second-order: phi(1) = 1
    assert O.lib.pgo_limit(O.LIMITER[name], 1.0, 1.0) == pytest.approx(1.0, rel=1e-15)


@pytest.mark.parametrize("name", sorted(PHI))
def test_the_symmetric_limiters_treat_both_slopes_alike(name):
    # phi(r)/r = phi(1/r) holds for minmod, MC, superbee and van Leer; Koren's is the one-sided
    # third-order form and breaks it, which tells it from the others (a swap of a and b shows)
    rng = np.random.default_rng(4)
    r = np.exp(rng.uniform(-3, 3, 500))
    one = limit(name, r, np.ones_like(r)) / r
    inv = limit(name, 1.0 / r, np.ones_like(r))
    if name == "koren":
        assert np.abs(one - inv).max() > 0.1
    else:
        assert np.allclose(one, inv, rtol=1e-14, atol=1e-15)


def _line(limiter, n=41):
    """A 1-D line of a calorically perfect binary gas, MUSCL with this limiter."""
    sim = get_simulator(sp.config("binary", recon="muscl", primary="rusanov", limiter=limiter))
    oc = O.OracleCase(sim.mixture, "muscl", limiter, "rusanov")
    hb = sp.HostBlock((n, 1, 1), 2, sim.ne, 1, lengths=(1.0, 1.0, 1.0))
    return sim, oc, hb


def _set(oc, hb, q):
    O.lib.pgo_fromprims(oc.ref, O.arr(sp.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state), O.box((0, 0, 0), hb.cshape))


def _sides(oc, hb, ns):
    """The two sides of every i face of the line's interior (its stencils reach the halo cells)."""
    w = 8 + ns - 1
    n = hb.ncell[0]
    out = np.zeros((n + 1, 2, w))
    buf = (O.C.c_double * (2 * w))()
    for f in range(n + 1):
        O.lib.pgo_sides(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), 0, hb.ng + f, hb.ng, hb.ng, buf)
        out[f] = np.array(buf[:]).reshape(2, w)
    return out


NAMES = ("p", "T", "rho", "e", "c", "u", "v", "w")


@pytest.mark.parametrize("name", sorted(PHI))
def test_a_linear_field_is_reconstructed_exactly(name):
    sim, oc, hb = _line(name)
    x = hb.xcc[..., 0]
    coef = {0: (101325.0, 10132.5), 1: (10.0, 50.0), 2: (-5.0, 20.0), 3: (3.0, -8.0), 4: (300.0, 100.0),
            5: (0.2, 0.5)}
    q = np.zeros(hb.cshape + (sim.ne,))
    for c, (a, b) in coef.items():
        q[..., c] = a + b * x
    _set(oc, hb, q)
    s = _sides(oc, hb, sim.ns)
    xf = np.arange(hb.ncell[0] + 1) / hb.ncell[0]
    # the stencil of the faces 2 .. n-2 stays within the linear field (the halo's extrapolation
    # of x is linear too, but keep to the interior)
    ok = slice(2, hb.ncell[0] - 1)
    want = {"p": coef[0], "u": coef[1], "v": coef[2], "w": coef[3], "T": coef[4]}
    for side in (0, 1):
        for k, (a, b) in want.items():
            got = s[ok, side, NAMES.index(k)]
            exact = a + b * xf[ok]
            assert np.abs(got - exact).max() <= 1e-13 * np.abs(exact).max(), (name, side, k)
        exact = 0.2 + 0.5 * xf[ok]
        assert np.abs(s[ok, side, 8] - exact).max() <= 1e-14, (name, side, "Y")


def _rough(sim, hb, seed=5):
    rng = np.random.default_rng(seed)
    q = np.zeros(hb.cshape + (sim.ne,))
    shape = q.shape[:3]
    q[..., 0] = rng.uniform(0.5e5, 2e5, shape)
    q[..., 1:4] = rng.uniform(-100, 100, shape + (3,))
    q[..., 4] = rng.uniform(250, 900, shape)
    q[..., 5] = rng.uniform(0.0, 1.0, shape)
    return q


def _cells(hb, ns):
    """p, T, u, v, w, Y of every cell along the line (i over the whole row, halos included)."""
    Q, st = sp.to_host(hb.regs[0])[:, hb.ng, hb.ng], sp.to_host(hb.state)[:, hb.ng, hb.ng]
    rho = Q[:, 0]
    return {"p": st[:, 0], "T": st[:, 1], "u": Q[:, 1] / rho, "v": Q[:, 2] / rho, "w": Q[:, 3] / rho,
            "Y": Q[:, 5] / rho}


@pytest.mark.parametrize("name", sorted(PHI))
def test_each_side_stays_within_its_two_cells(name):
    sim, oc, hb = _line(name)
    _set(oc, hb, _rough(sim, hb))
    s = _sides(oc, hb, sim.ns)
    c = _cells(hb, sim.ns)
    n, g = hb.ncell[0], hb.ng
    for side in (0, 1):
        for k in ("p", "T", "u", "v", "w", "Y"):
            col = 8 if k == "Y" else NAMES.index(k)
            L, R = c[k][g - 1:g + n], c[k][g:g + n + 1]
            lo, hi = np.minimum(L, R), np.maximum(L, R)
            got = s[:, side, col]
            tol = 1e-14 * (hi - lo) + 1e-300
            assert (got >= lo - tol).all() and (got <= hi + tol).all(), (name, side, k)


@pytest.mark.parametrize("name", sorted(PHI))
def test_the_sides_follow_sweby_with_the_textbook_phi(name):
    # vL + phi(r)/2 (vR - vL) left of the face, vR - phi(r')/2 (vR - vL) right of it, r the cell's
    # upwind difference over the difference across the face
    sim, oc, hb = _line(name)
    _set(oc, hb, _rough(sim, hb))
    s = _sides(oc, hb, sim.ns)
    c = _cells(hb, sim.ns)
    n, g = hb.ncell[0], hb.ng
    phi = PHI[name]
    for k in ("p", "T", "u", "v", "w", "Y"):
        col = 8 if k == "Y" else NAMES.index(k)
        v = c[k]
        LL, L, R, RR = v[g - 2:g + n - 1], v[g - 1:g + n], v[g:g + n + 1], v[g + 1:g + n + 2]
        d = R - L
        with np.errstate(divide="ignore", invalid="ignore"):
            left = L + 0.5 * np.where(d == 0, 0.0, phi((L - LL) / d)) * d
            right = R - 0.5 * np.where(d == 0, 0.0, phi((RR - R) / d)) * d
        scale = np.abs(v).max()
        assert np.abs(s[:, 0, col] - left).max() <= 1e-13 * scale, (name, k, "left")
        assert np.abs(s[:, 1, col] - right).max() <= 1e-13 * scale, (name, k, "right")


@pytest.mark.parametrize("name", sorted(PHI))
def test_each_side_is_a_state_of_the_eos(name):
    # rho, e and c of each side are the calorically perfect gas's of its p, T and Y in closed form:
    # h = cp (T - Tref) with no heat of formation, so e = cv T - cp Tref (tests/simulator/oracle.py:19-28)
    sim, oc, hb = _line(name)
    _set(oc, hb, _rough(sim, hb))
    s = _sides(oc, hb, sim.ns)
    MW = np.array([28.0, 44.0])
    cp = np.array([1040.0, 846.0])
    Ru = 8314.46261815324
    for side in (0, 1):
        p, T, Y0 = s[:, side, 0], s[:, side, 1], s[:, side, 8]
        Y = np.stack([Y0, 1 - Y0], axis=-1)
        R = Ru * (Y / MW).sum(axis=-1)
        cpm = (Y * cp).sum(axis=-1)
        rho = p / (R * T)
        gamma = cpm / (cpm - R)
        assert np.allclose(s[:, side, 2], rho, rtol=1e-14, atol=0.0)
        assert np.allclose(s[:, side, 4], np.sqrt(gamma * R * T), rtol=1e-14, atol=0.0)
        e = (cpm - R) * T - cpm * 298.15
        # e crosses zero along the line: held at the row's scale
        assert np.abs(s[:, side, 3] - e).max() <= 1e-14 * np.abs(e).max()
