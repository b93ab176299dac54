"""Species records: what a mixture knows about each species before any fit.

Follows ``peregrine/simulator/mixture/species.py:19-136``: a species is a bag
of named quantities (``MW``, ``NASA7`` or ``janaf``, ``cp0``, critical and
transport data); ``MW`` can be derived from ``comp``; the case's own data win
over the library, and a species that states one form of a quantity (a NASA7
fit, say) does not take the library's other form of it (JANAF tables).
"""

from functools import lru_cache
from pathlib import Path

import yaml

DATA = Path(__file__).resolve().parent.parent / "data"


class AnyOf(tuple):
    """A need met by any one of several keys."""


@lru_cache(maxsize=None)
def database(name):
    path = DATA / f"{name}.yaml"
    if not path.is_file():
        return {}
    with open(path) as f:
        return yaml.safe_load(f) or {}


class Species:
    # kg/kmol (IUPAC standard atomic weights as Cantera tabulates them)
    atomwts = {
        "H": 1.008, "He": 4.002602, "Li": 6.94, "Be": 9.0121831, "B": 10.81,
        "C": 12.011, "N": 14.007, "O": 15.999, "F": 18.998403163, "Ne": 20.1797,
        "Na": 22.98976928, "Mg": 24.305, "Al": 26.9815384, "Si": 28.085,
        "P": 30.973761998, "S": 32.06, "Cl": 35.45, "Ar": 39.95, "K": 39.0983,
        "Ca": 40.078, "Kr": 83.798,
    }
    # rotational degrees of freedom over two, by molecular shape
    rotdof = {"atom": 0.0, "linear": 1.0, "nonlinear": 1.5}

    def __init__(self, name, data=None):
        self.name = name
        self.data = {}
        self.supplement(data)

    def __repr__(self):
        return f"<Species {self.name}: {', '.join(sorted(self.data))}>"

    def has(self, key):
        return key in self.data or (key == "MW" and "comp" in self.data)

    __contains__ = has

    def __getitem__(self, key):
        if key in self.data:
            return self.data[key]
        if key == "MW" and "comp" in self.data:
            return self.molwt(self.data["comp"])
        raise KeyError(f"species {self.name} has no {key}")

    def __setitem__(self, key, value):
        self.data[key] = value

    def get(self, key, default=None):
        try:
            return self[key]
        except KeyError:
            return default

    def supplement(self, data):
        for key, value in (data or {}).items():
            if value is not None:
                self.data.setdefault(key, value)
        return self

    def missing(self, keys):
        out = []
        for key in keys:
            ok = any(map(self.has, key)) if isinstance(key, AnyOf) else self.has(key)
            if not ok:
                out.append(key)
        return out

    @classmethod
    def molwt(cls, comp, weights=None):
        weights = weights or cls.atomwts
        return sum(weights[el] * n for el, n in comp.items())

    @classmethod
    def build(cls, usersp, models):
        library = database("species")
        # keys that are one quantity in different forms: having one is having it
        sameas = {}
        for model in models:
            for group in (k for k in model.needs if isinstance(k, AnyOf)):
                for key in group:
                    sameas[key] = group
        built = {}
        for name, data in usersp.items():
            sp = cls(name, data)
            known = library.get(name) or {}
            had = {k for k in sameas if not sp.missing([sameas[k]])}
            sp.supplement({k: v for k, v in known.items() if k not in had})
            built[name] = sp
        return built
