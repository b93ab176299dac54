"""Finite-rate mass-action chemistry: the per-reaction records and tables.

Restates ``peregrine/simulator/mixture/chemistry/finiterate.py:21-313`` and
``peregrine/simulator/mixture/reaction.py:1-43``:

- rates in log space: ln k = ln A + b ln T - (Ea/R)/T; falloff pressure ratios
  as their own Arrhenius rate relative to the high-pressure one;
- Troe's log10 Fcent as a degree-12 Chebyshev least-squares fit in
  t = (ln T - uc)/uh over the case's T-range, converted to the power basis;
- per reaction, the *product form* (one table exponential, then products of
  c e^(g/RuT)) when every exponential argument stays inside the table exp's
  range over the T-range, else the *log form*;
- the log-concentration floor ``logcfloor`` and the clamp ``expclamp``;
- third-body sets shared between reactions.

The records feed both the mechanism code generator (the product kernels) and
the flat tables the oracle reads.
"""

import math
from types import SimpleNamespace

import numpy as np
from numpy.polynomial import chebyshev

from peregrine_amd.mixture.constants import Ru, pref
from peregrine_amd.mixture.eos import BaseEos, BaseModel
from peregrine_amd.mixture.fits import padded
from peregrine_amd.mixture.parser import KINDS

LOG10E = 0.4342944819032518
LN10 = 2.302585092994046
# the table exp's arguments stay inside this over the case's temperatures for the product form
EXPBOUND = 680.0
# the boost a kind adds to ln k beyond its Arrhenius rate, every concentration at most e^logcmax
BOOST = {"elementary": lambda lc: 0.0, "three-body": lambda lc: lc + 2.0, "lindemann": lambda lc: 1.0, "troe": lambda lc: 1.0}
LFCDEG = 12


class TroeCentre:
    """Troe's broadening centre F_cent(T) = (1 - a) exp(-T/T3) + a exp(-T/T1) + exp(-T2/T), held as
    its exponential terms sign * exp(c0 + cT T + cTinv / T). The mechanism's sentinels (T3 or T1 zero,
    T2 absent) switch a term off; over the temperatures a gas takes (100 to 20000 K) a term whose
    exponent is below -745 throughout is dropped, and a coefficient whose part of the exponent stays
    below 1e-14 is zeroed (peregrine/simulator/mixture/chemistry/finiterate.py:240-277). The kernels
    read log10 F_cent from a polynomial in t = (ln T - uc) / uh fitted here."""

    TMIN, TMAX = 100.0, 20000.0

    def __init__(self, troe):
        self.terms = []
        if troe is None:
            return
        a, T3, T1, T2 = troe
        parts = []
        if a != 1.0 and T3 != 0:
            parts.append((1.0 - a, -1.0 / T3, 0.0))
        if a != 0.0 and T1 != 0:
            parts.append((a, -1.0 / T1, 0.0))
        if T2:
            parts.append((1.0, 0.0, -T2))
        for coeff, cT, cTinv in parts:
            if cT * self.TMIN < -745 or cTinv / self.TMAX < -745:
                continue
            if abs(cT) * self.TMAX < 1e-14:
                cT = 0.0
            if abs(cTinv) / self.TMIN < 1e-14:
                cTinv = 0.0
            self.terms.append((float(np.sign(coeff)), float(np.log(abs(coeff))), cT, cTinv))

    def __len__(self):
        return len(self.terms)

    def log10(self, T):
        """log10 F_cent at T by a signed log-sum-exp about the largest exponent; with no terms,
        log10 of the smallest normal double (Cantera's floor)."""
        T = np.asarray(T, dtype=np.float64)
        tiny = np.finfo(np.float64).tiny
        if not self.terms:
            return np.full(T.shape, np.log(tiny) * LOG10E)
        exps = [c0 + cT * T + cTinv / T for _, c0, cT, cTinv in self.terms]
        top = np.max(exps, axis=0)
        acc = 0
        for (sign, *_), x in zip(self.terms, exps):
            acc = acc + sign * np.exp(x - top)
        return (top + np.log(np.maximum(tiny, acc))) * LOG10E

    def poly(self, uc, uh, degree=LFCDEG):
        """log10 F_cent's power-basis coefficients in t (rising), from the least-squares Chebyshev
        series of the given degree through 8 degree + 64 Chebyshev points of the first kind on
        [-1, 1]: degree 12 is within 1e-3 of the centre, the precision Troe's parameters carry."""
        if not self.terms:
            return []
        nodes = 8 * degree + 64
        t = np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes)
        series = chebyshev.chebfit(t, self.log10(np.exp(uc + uh * t)), degree)
        return [float(c) for c in chebyshev.cheb2poly(series)]


class FiniteRate(BaseModel):
    name = "finite-rate"

    def __init__(self, cfgsect, reactions):
        super().__init__(cfgsect)
        self.reactions = reactions

    # -- records ------------------------------------------------------------
    def records(self, species):
        """(records, scalars): one SimpleNamespace per reaction, and the case's scalars."""
        names = list(species)
        MW = np.array([species[n]["MW"] for n in names])
        index = {n: k for k, n in enumerate(names)}
        reactions = [r for r in self.reactions if r["rate"][0] != 0.0]
        for r in reactions:
            if r["rate"][0] < 0 or r.get("lowrate", (1.0,))[0] <= 0:
                raise ValueError(f"{r['equation']}: rates are evaluated in log space, A must be positive")
        Tlow, Thigh = self.trange
        uc = 0.5 * (np.log(Thigh) + np.log(Tlow))
        uh = 0.5 * (np.log(Thigh) - np.log(Tlow))
        records, nets, fwds = [], [], []
        ntroe = 0
        for r in reactions:
            A, b, EaR = r["rate"]
            A0, b0, EaR0 = r.get("lowrate", (1.0, 0.0, 0.0))
            logA = np.log(A)
            nu = {}
            for sp, v in r["reactants"].items():
                nu[sp] = nu.get(sp, 0.0) - v
            for sp, v in r["products"].items():
                nu[sp] = nu.get(sp, 0.0) + v
            net = [(index[sp], v) for sp, v in nu.items() if v != 0]
            kind = KINDS[r["type"]]
            fcent = TroeCentre(r.get("troe"))
            rec = SimpleNamespace(
                name=r["type"],
                equation=r["equation"],
                **kind,
                reversible=bool(r["reversible"]),
                troeidx=ntroe if kind["troe"] else None,
                logA=float(logA),
                b=float(b),
                EaR=float(EaR),
                logprA=float(np.log(A0) - logA),
                prb=float(b0 - b),
                prEaR=float(EaR0 - EaR),
                fcent=fcent,
                lfcpoly=fcent.poly(uc, uh),
                fwd=[(index[sp], v) for sp, v in r["fwd"].items() if v != 0],
                react=[(index[sp], v) for sp, v in r["reactants"].items() if v != 0],
                prod=[(index[sp], v) for sp, v in r["products"].items() if v != 0],
                net=[(sp, v, float(v * MW[sp])) for sp, v in net],
                eff=[(index[sp], e - r["defeff"]) for sp, e in r["efficiencies"].items() if e != r["defeff"]],
                nusum=float(sum(v for _, v in net)),
                defeff=float(r["defeff"]),
            )
            ntroe += kind["troe"]
            records.append(rec)
            nets.append(net)
            fwds.append(rec.fwd)
        grid = np.geomspace(Tlow, Thigh, 2048)
        logT, Tinv, gRT = np.log(grid), 1.0 / grid, BaseEos.gibbs(species, grid)
        for rec in records:
            rec.product = self._productform(rec, grid, logT, Tinv, gRT, len(species))
        gain = np.zeros(len(species))
        for row in nets:
            for sp, v in row:
                gain[sp] += abs(v)
        maxgain = max(1.0, (gain * MW).max()) if records else 1.0
        logcfloor = self._logcfloor(species, reactions, records, nets, fwds, index)
        self._thirdbodies(records)
        scalars = dict(
            logprefRu=float(np.log(pref / Ru)),
            prefRu=float(pref / Ru),
            Rupref=float(Ru / pref),
            logcfloor=float(logcfloor),
            cfloor=float(math.exp(logcfloor)),
            maxgain=float(maxgain),
            troeuc=float(uc),
            troeuhinv=float(1.0 / uh),
            expclamp=float(math.log(np.finfo(np.float64).max) - math.log(2.0 * maxgain)),
        )
        return records, scalars

    @staticmethod
    def _thirdbodies(records):
        sets, keys = [], []
        for r in records:
            r.tb = None
            if r.thirdbody:
                key = (r.defeff, tuple(r.eff))
                if key not in keys:
                    keys.append(key)
                    sets.append(SimpleNamespace(defeff=r.defeff, eff=list(r.eff)))
                r.tb = keys.index(key)
        return sets

    def thirdbodies(self, species):
        records, _ = self.records(species)
        return self._thirdbodies(records)

    @staticmethod
    def _productform(r, T, logT, Tinv, gRT, ns):
        whole = all(float(v).is_integer() and v > 0 for _, v in r.react + r.prod)
        if not whole or dict(r.fwd) != dict(r.react):
            return False
        sps = [sp for sp, _ in r.react + r.prod]
        if np.abs(gRT[sps]).max() >= EXPBOUND:
            return False
        xk = r.logA + r.b * logT - r.EaR * Tinv - sum(nu * gRT[sp] for sp, nu in r.react)
        low, high = xk.min(), xk.max()
        if r.troe:
            l10 = r.fcent.log10(T)
            low, high = low + LN10 * min(l10.min(), 0.0), high + LN10 * max(l10.max(), 0.0)
        if r.falloff:
            xpr = r.logprA + r.prb * logT - r.prEaR * Tinv
            if np.abs(xpr).max() >= EXPBOUND - 8.0:
                return False
        if not (low > -EXPBOUND and high < EXPBOUND):
            return False
        logc = 12.0
        extra = logc + np.log(r.defeff * ns + sum(max(d, 0.0) for _, d in r.eff)) if r.thirdbody and not r.falloff else 0.0
        chain = xk + extra
        for sp, nu in r.react:
            for _ in range(int(nu)):
                chain = chain + gRT[sp] + logc
                if chain.max() >= EXPBOUND:
                    return False
        chain = xk + abs(r.nusum) * 6.0
        for sp, nu in r.prod:
            for _ in range(int(nu)):
                chain = chain + gRT[sp] + logc
                if chain.max() >= EXPBOUND:
                    return False
        return True

    def _logcfloor(self, species, reactions, records, nets, fwds, index):
        logcmax, margin = 12.0, 45.0
        Tlow, Thigh = self.trange
        T = np.geomspace(max(Tlow, 100.0), Thigh, 256)
        logT, gRT = np.log(T), BaseEos.gibbs(species, T)
        required = 90.0
        for r, rec, netrow, fwdrow in zip(reactions, records, nets, fwds):
            A, b, EaR = r["rate"]
            logkf = np.log(A) + b * logT - EaR / T
            aux = BOOST[rec.name](logcmax)
            top = logkf.max() + aux
            for k, vk in fwdrow:
                if vk <= 0:
                    continue
                other = sum(v * logcmax for n, v in fwdrow if n != k and v > 0)
                required = max(required, (top + other + margin) / vk)
            if r["reversible"]:
                affinity = sum(v * gRT[n] for n, v in netrow)
                nusum = sum(v for _, v in netrow)
                logkr = logkf + affinity + nusum * np.log(Ru * T / pref)
                rev = [(index[sp], float(v)) for sp, v in r["products"].items() if v > 0]
                top = logkr.max() + aux
                for k, vk in rev:
                    other = sum(v * logcmax for n, v in rev if n != k)
                    required = max(required, (top + other + margin) / vk)
        return -float(required)

    # -- flat tables, for the oracle -----------------------------------------
    def tables(self, species):
        """Every per-reaction column as a fixed-width array, plus the scalars."""
        records, scalars = self.records(species)
        sets = self._thirdbodies(records)
        W = 4  # widest reaction side / net list / fwd list held
        for r in records:
            if max(len(r.fwd), len(r.react), len(r.prod), len(r.net)) > W:
                raise ValueError(f"{r.equation}: more than {W} species on a side")
        kindcode = {"elementary": 0, "three-body": 1, "lindemann": 2, "troe": 3}
        cols = dict(
            kind=np.array([kindcode[r.name] for r in records], np.int32),
            reversible=np.array([int(r.reversible) for r in records], np.int32),
            product=np.array([int(r.product) for r in records], np.int32),
            tb=np.array([-1 if r.tb is None else r.tb for r in records], np.int32),
            troeidx=np.array([-1 if r.troeidx is None else r.troeidx for r in records], np.int32),
            logA=np.array([r.logA for r in records]),
            b=np.array([r.b for r in records]),
            EaR=np.array([r.EaR for r in records]),
            logprA=np.array([r.logprA for r in records]),
            prb=np.array([r.prb for r in records]),
            prEaR=np.array([r.prEaR for r in records]),
            nusum=np.array([r.nusum for r in records]),
            nfwd=np.array([len(r.fwd) for r in records], np.int32),
            fwdsp=padded([[s for s, _ in r.fwd] for r in records], np.int32, W),
            fwdorder=padded([[v for _, v in r.fwd] for r in records], width=W),
            nreact=np.array([len(r.react) for r in records], np.int32),
            reactsp=padded([[s for s, _ in r.react] for r in records], np.int32, W),
            reactnu=padded([[v for _, v in r.react] for r in records], width=W),
            nprod=np.array([len(r.prod) for r in records], np.int32),
            prodsp=padded([[s for s, _ in r.prod] for r in records], np.int32, W),
            prodnu=padded([[v for _, v in r.prod] for r in records], width=W),
            nnet=np.array([len(r.net) for r in records], np.int32),
            netsp=padded([[s for s, _, _ in r.net] for r in records], np.int32, W),
            netnu=padded([[v for _, v, _ in r.net] for r in records], width=W),
            netnuMW=padded([[m for _, _, m in r.net] for r in records], width=W),
            lfcpoly=padded([r.lfcpoly for r in records if r.troe], width=LFCDEG + 1)
            if any(r.troe for r in records)
            else np.zeros((1, LFCDEG + 1)),
            ntbeff=np.array([len(t.eff) for t in sets] or [0], np.int32),
            tbdefeff=np.array([t.defeff for t in sets] or [0.0]),
            tbeffsp=padded([[s for s, _ in t.eff] for t in sets] or [[]], np.int32, len(species)),
            tbeffdev=padded([[d for _, d in t.eff] for t in sets] or [[]], width=len(species)),
        )
        return records, sets, cols, scalars
