"""Mass-action production rates by the textbook formulas, in numpy with the library exp and log.

Independent of the kernels' forms (no table exp/log, no product/log-form split, no clamps):
k = A T^b exp(-Ea/RT); a third body's concentration sum(eff_n c_n); falloff
k = k_inf Pr/(1 + Pr) F with Troe's log10 F = log10 Fc / (1 + ((log10 Pr + C)/(N - 0.14 (log10 Pr + C)))^2);
the reverse rate k / Kc with Kc = exp(-sum nu g/RT) (pref / Ru T)^(sum nu).
log10 Fc is taken from the case's own fit (lfcpoly) unless exact_fcent, then from its terms.
The Gibbs polynomials are evaluated at u = ln T, or at the u given (the kernels' table log of T:
its 3e-14 absolute error, through polynomials in ln T of degree 7, moves a rate by up to 1e-10).
"""

import numpy as np

from peregrine_amd.mixture.constants import Ru, pref
from peregrine_amd.mixture.eos import BaseEos


def production(mixture, rho, Y, T, exact_fcent=False, u=None):
    records, sc = mixture.chemistry.records(mixture.species)
    sets = mixture.chemistry.thirdbodies(mixture.species)
    MW = np.array([s["MW"] for s in mixture.species.values()])
    c = rho * np.asarray(Y) / MW
    if u is None:
        g = BaseEos.gibbs(mixture.species, np.array([T]))[:, 0]
    else:
        gpoly, h298 = BaseEos.gpoly(mixture.species), BaseEos.collect("h298", mixture.species)
        g = np.array([np.polyval(np.flip(gp), u) + h / T for gp, h in zip(gpoly, h298)])
    ctot = c.sum()
    omega = np.zeros(len(c))
    for r in records:
        k = np.exp(r.logA) * T**r.b * np.exp(-r.EaR / T)
        if r.thirdbody:
            t = sets[r.tb]
            M = t.defeff * ctot + sum(d * c[s] for s, d in t.eff)
        if r.thirdbody and not r.falloff:
            k *= M
        if r.falloff:
            k0_over_kinf = np.exp(r.logprA) * T**r.prb * np.exp(-r.prEaR / T)
            Pr = k0_over_kinf * M
            k *= Pr / (1.0 + Pr)
            if r.troe:
                if exact_fcent:
                    l10 = float(r.fcent.log10(np.array([T]))[0])
                else:
                    t = (np.log(T) - sc["troeuc"]) * sc["troeuhinv"]
                    l10 = np.polyval(np.flip(r.lfcpoly), t)
                C = -0.4 - 0.67 * l10
                N = 0.75 - 1.27 * l10
                x = np.log10(Pr) + C
                k *= 10.0 ** (l10 / (1.0 + (x / (N - 0.14 * x)) ** 2))
        fwd = k * np.prod([c[s] ** v for s, v in r.fwd])
        rev = 0.0
        if r.reversible:
            dG = sum(v * g[s] for s, v, _ in r.net)
            Kc = np.exp(-dG) * (pref / (Ru * T)) ** r.nusum
            rev = k / Kc * np.prod([c[s] ** v for s, v in r.prod])
        for s, v, _ in r.net:
            omega[s] += v * MW[s] * (fwd - rev)
    return omega


def density(mixture, Y, T, p):
    MW = np.array([s["MW"] for s in mixture.species.values()])
    return p / (Ru * T * (np.asarray(Y) / MW).sum())
