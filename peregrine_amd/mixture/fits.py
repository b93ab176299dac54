"""Species thermodynamic data and the ln T polynomial refits every kernel reads.

Restates:
- the NASA-7 and JANAF sources, ``peregrine/simulator/mixture/thermo.py:38-71``
  (cp/R sampled at 200 points across the NASA-7 range);
- the fitting helpers, ``peregrine/simulator/mixture/polyfit.py:4-34`` (the
  lowest degree whose weighted least-squares fit is within ``refit-tol``
  relative, coefficients rising in ln T);
- the caloric models, ``caloric/tdepcp/model.py:12-70`` (temperature-dependent
  cp) and ``caloric/constcp.py`` (constant cp).

Conventions (the kernels rely on them): with u = ln T,
  cp/R   = sum_m cppoly[m] u^m
  h/R    = T * sum_m hpoly[m] u^m + h298        (hpoly = the e^u integral of cppoly)
  s/R    = sum_m spoly[m] u^m                    (spoly = the u integral of cppoly)
  g/(RT) = sum_m gpoly[m] u^m + h298 / T          (gpoly = hpoly - spoly)
"""

import numpy as np

from peregrine_amd.mixture.constants import Ru, Tref
from peregrine_amd.mixture.species import HALF_ROTATIONAL_DOF, OneOf


class ThermoSource:
    key = None

    @classmethod
    def find(cls, sp):
        for sub in cls.__subclasses__():
            if sub.key in sp:
                return sub
        raise KeyError(f"species {sp.name} has no thermo data")


class Nasa7(ThermoSource):
    key = "NASA7"

    @staticmethod
    def coeffs(n7, T):
        return n7["low"] if T <= n7["T-range"][1] else n7["high"]

    @classmethod
    def cpR(cls, n7, T):
        a = cls.coeffs(n7, T)
        return a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])))

    @classmethod
    def hRT(cls, n7, T):
        a = cls.coeffs(n7, T)
        return a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T

    @classmethod
    def sR(cls, n7, T):
        a = cls.coeffs(n7, T)
        return a[0] * np.log(T) + T * (a[1] + T * (a[2] / 2 + T * (a[3] / 3 + T * a[4] / 4))) + a[6]

    @classmethod
    def cp_r(cls, sp):
        n7 = sp["NASA7"]
        T = np.linspace(n7["T-range"][0], n7["T-range"][-1], 200)
        return T, np.array([cls.cpR(n7, t) for t in T])

    @classmethod
    def hsref_r(cls, sp):
        n7 = sp["NASA7"]
        return cls.hRT(n7, Tref) * Tref, cls.sR(n7, Tref)


class Janaf(ThermoSource):
    key = "janaf"

    @staticmethod
    def cp_r(sp):
        j = sp["janaf"]
        return np.array(j["T"], float), np.array(j["cp"], float) / Ru

    @staticmethod
    def hsref_r(sp):
        j = sp["janaf"]
        return j["dHf298"] / Ru, j["s298"] / Ru


# ---------------------------------------------------------------------------
# polynomial helpers: coefficient arrays rise in the power
# ---------------------------------------------------------------------------
def fit(x, y, tol, maxdeg):
    """Lowest degree whose 1/y^2-weighted fit is within ``tol`` relative."""
    weights = 1.0 / y**2
    for degree in range(maxdeg + 1):
        c = np.polyfit(x, y, degree, w=weights)
        error = np.abs(np.polyval(c, x) / y - 1).max()
        if error <= tol:
            break
    return np.flip(c), error


def evaluate(poly, x):
    return np.polyval(np.flip(np.asarray(poly, float)), x)


def expintegral(P):
    """Q with d/du (e^u Q(u)) = e^u P(u): Q = P - P' + P'' - ..."""
    P = np.asarray(P, float)
    Q = np.zeros(len(P))
    D = np.flip(P)
    for k in range(len(P)):
        Q[: len(D)] += (-1) ** k * np.flip(D)
        D = np.polyder(D)
    return Q


def integral(P):
    return np.flip(np.polyint(np.flip(np.asarray(P, float))))


def padded(rows, dtype=np.float64, width=None):
    terms = max((len(r) for r in rows), default=1)
    terms = max(terms, 1) if width is None else width
    out = np.zeros((len(rows), terms), dtype=dtype)
    for k, r in enumerate(rows):
        out[k, : len(r)] = r
    return out


# ---------------------------------------------------------------------------
# caloric models
# ---------------------------------------------------------------------------
class TempDepCp:
    """cp(T) refit in ln T over the case's T-range."""

    needs = (OneOf(("NASA7", "janaf")),)

    def __init__(self, cfgsect):
        self.cfgsect = cfgsect

    def cppoly(self, sp):
        T, cpR = self._sample(sp)
        return fit(np.log(T), cpR, self.cfgsect["refit-tol"], self.cfgsect["refit-max-deg"])

    def hpoly(self, sp):
        return expintegral(sp["cppoly"])

    def h298(self, sp):
        h298R, _ = ThermoSource.find(sp).hsref_r(sp)
        return h298R - Tref * evaluate(sp["hpoly"], np.log(Tref))

    def spoly(self, sp):
        _, srefR = ThermoSource.find(sp).hsref_r(sp)
        S = integral(sp["cppoly"])
        S[0] = srefR - evaluate(S, np.log(Tref))
        return S

    def _sample(self, sp):
        """cp/R samples over the case's T-range: the thermo source's own points, carried past its
        ends where the range reaches further (caloric/tdepcp/model.py:35-64 of the reference's
        mixture) and then cut to the range. Below the first point the samples continue at its
        spacing down to Tlow on the straight line from the frozen 0 K value (translation, plus
        rotation where the shape is known) to the first sample; above the last, at its spacing up to
        Thigh, holding the last sample."""
        Tlow, Thigh = self.cfgsect["T-range"]
        T, cpR = ThermoSource.find(sp).cp_r(sp)
        if Tlow < T[0]:
            below = _outward(T[0], T[1] - T[0], Tlow)[::-1]
            frozen = 2.5 + HALF_ROTATIONAL_DOF[sp["geometry"]] if "geometry" in sp else cpR[0]
            T, cpR = np.concatenate([below, T]), np.concatenate([frozen + (cpR[0] - frozen) * below / T[0], cpR])
        if Thigh > T[-1]:
            above = _outward(T[-1], T[-1] - T[-2], Thigh)
            T, cpR = np.concatenate([T, above]), np.concatenate([cpR, np.full_like(above, cpR[-1])])
        keep = (T >= Tlow) & (T <= Thigh)
        return T[keep], cpR[keep]


def _outward(edge, spacing, bound):
    """Points from edge outward (away from the samples) at the given spacing, short of bound, then
    bound itself."""
    step = abs(spacing) if bound > edge else -abs(spacing)
    return np.append(np.arange(edge + step, bound, step), bound)


class ConstantCp:
    needs = ("cp0",)

    def __init__(self, cfgsect):
        self.cfgsect = cfgsect

    def cppoly(self, sp):
        return [sp["cp0"] * sp["MW"] / Ru], 0.0

    def hpoly(self, sp):
        return sp["cppoly"]

    def h298(self, sp):
        return sp.get("dHf298", 0.0) / Ru - Tref * sp["cppoly"][0]

    def spoly(self, sp):
        c = sp["cppoly"][0]
        return [sp.get("s298", 0.0) / Ru - c * np.log(Tref), c]
