"""Equations of state: the per-species tables the state kernels read.

Follows ``peregrine/simulator/mixture/eos/base.py:7-74`` (the fits each eos
stores on the species, the constants it emits, gpoly and the Gibbs function)
and the cpg/tpg models (``eos/cpg/model.py``, ``eos/tpg/model.py``).

The device kernels read fixed-width rows: cppoly padded with zeros to
``NT = 7`` terms (refit degree at most 6) and gpoly to ``NG = 8``. Zero
padding is exact in Horner's scheme and in the h-coefficient recurrence, so a
lower fitted degree gives the same bits.
"""

import numpy as np

from peregrine_amd.mixture.constants import Ru
from peregrine_amd.mixture.fits import ConstantCp, TempDepCp, evaluate, padded
from peregrine_amd.mixture.species import AnyOf

NT = 7
NG = 8


class BaseModel:
    name = None

    def __init__(self, cfgsect):
        self.cfgsect = cfgsect
        self.needs = ("MW",)
        self.provides = {}

    @property
    def trange(self):
        return tuple(self.cfgsect["T-range"])

    def check(self, species):
        missing = {}
        for name, sp in species.items():
            lack = sp.missing(self.needs)
            if lack:
                missing[name] = lack
        if missing:
            text = "; ".join(
                f"{', '.join('|'.join(k) if isinstance(k, AnyOf) else k for k in keys)} for {name}"
                for name, keys in missing.items()
            )
            raise ValueError(f"{self.name} needs {text}")

    def populate(self, species):
        for method, quantities in self.provides.items():
            values = getattr(self, method)(species)
            if isinstance(quantities, str):
                quantities, values = (quantities,), (values,)
            for quantity, column in zip(quantities, values):
                for sp, v in zip(species.values(), column):
                    sp[quantity] = v

    def constants(self, species):
        return {}

    @staticmethod
    def collect(key, species):
        return np.array([sp[key] for sp in species.values()], float)

    @staticmethod
    def collect_poly(key, species, width=None):
        return padded([sp[key] for sp in species.values()], width=width)


class BaseEos(BaseModel):
    # the state fields a cell carries, in storage order
    fields = ("p", "T", "rhoinv", "rho", "e", "cp", "gamma", "c")

    def __init__(self, cfgsect, caloric):
        super().__init__(cfgsect)
        self.caloric = caloric(cfgsect)
        self.needs += self.caloric.needs
        # ordered: each reads what the one before stored
        self.provides["cppoly"] = ("cppoly", "cpfiterr")
        self.provides.update(hpoly="hpoly", h298="h298", spoly="spoly")

    def cppoly(self, species):
        fits = [self.caloric.cppoly(sp) for sp in species.values()]
        return [f for f, _ in fits], [e for _, e in fits]

    def hpoly(self, species):
        return [self.caloric.hpoly(sp) for sp in species.values()]

    def h298(self, species):
        return [self.caloric.h298(sp) for sp in species.values()]

    def spoly(self, species):
        return [self.caloric.spoly(sp) for sp in species.values()]

    def constants(self, species):
        terms = max(len(sp["cppoly"]) for sp in species.values())
        if terms > NT:
            raise ValueError(f"cp refits of {terms} terms; the kernels hold at most {NT} (refit-max-deg <= {NT - 1})")
        cppoly = self.collect_poly("cppoly", species, NT)
        hpoly = self.collect_poly("hpoly", species, NT)
        h298 = self.collect("h298", species)
        spoly = self.collect_poly("spoly", species, NG)
        Rsp = Ru / self.collect("MW", species)
        return dict(
            cppoly=cppoly,
            hpoly=hpoly,
            h298=h298,
            spoly=spoly,
            gpoly=self.gpoly(species),
            cppolyR=cppoly * Rsp[:, None],
            hpolyR=hpoly * Rsp[:, None],
            h298R=h298 * Rsp,
        )

    @classmethod
    def gpoly(cls, species):
        hpoly = cls.collect_poly("hpoly", species, NG)
        spoly = cls.collect_poly("spoly", species, NG)
        return hpoly - spoly

    @classmethod
    def gibbs(cls, species, T):
        """g / (Ru T) per species at T, as the reverse rates read it."""
        T = np.asarray(T, dtype=np.float64)
        u, Tinv = np.log(T), 1.0 / T
        gpoly, h298 = cls.gpoly(species), cls.collect("h298", species)
        return np.array([evaluate(g, u) + h * Tinv for g, h in zip(gpoly, h298)])


class Cpg(BaseEos):
    name = "cpg"
    kind = 0

    def __init__(self, cfgsect):
        super().__init__(cfgsect, ConstantCp)

    def constants(self, species):
        return dict(super().constants(species), cp0=self.collect("cp0", species))


class Tpg(BaseEos):
    name = "tpg"
    kind = 1

    def __init__(self, cfgsect):
        super().__init__(cfgsect, TempDepCp)


EOS_MODELS = {m.name: m for m in (Cpg, Tpg)}
