"""Species records: what a case knows about each species before any fit.

A record is a mapping of named quantities (``MW``, ``comp``, ``NASA7`` or ``janaf``, ``cp0``,
critical and Lennard-Jones data). It is assembled from two layers, the case's own entry for the
species over the species library (``data/species.yaml``), with two rules the case files of the
reference rely on (``peregrine/simulator/mixture/species.py:19-136`` states them): the case's value
of a quantity is the one used, and a quantity that comes in interchangeable forms (a ``NASA7`` fit
or ``janaf`` tables) is taken whole from whichever layer states any form of it. ``MW`` follows from
``comp`` when no layer gives it.
"""

from collections.abc import Mapping
from functools import lru_cache
from pathlib import Path

import yaml

DATA = Path(__file__).resolve().parent.parent / "data"

# kg/kmol (IUPAC standard atomic weights as Cantera tabulates them)
ATOMIC_WEIGHTS = {
    "H": 1.008, "He": 4.002602, "Li": 6.94, "Be": 9.0121831, "B": 10.81,
    "C": 12.011, "N": 14.007, "O": 15.999, "F": 18.998403163, "Ne": 20.1797,
    "Na": 22.98976928, "Mg": 24.305, "Al": 26.9815384, "Si": 28.085,
    "P": 30.973761998, "S": 32.06, "Cl": 35.45, "Ar": 39.95, "K": 39.0983,
    "Ca": 40.078, "Kr": 83.798,
}

# half the rotational degrees of freedom of a molecule of each shape
HALF_ROTATIONAL_DOF = {"atom": 0.0, "linear": 1.0, "nonlinear": 1.5}


class OneOf(tuple):
    """A requirement any one of several keys meets: the interchangeable forms of one quantity."""

    def __str__(self):
        return "|".join(self)


def requirement_text(req):
    return str(req) if isinstance(req, OneOf) else req


@lru_cache(maxsize=None)
def database(name):
    """A YAML table of the package's data directory by name, empty when there is none."""
    path = DATA / f"{name}.yaml"
    if not path.is_file():
        return {}
    with open(path) as f:
        return yaml.safe_load(f) or {}


def molecular_weight(comp, weights=None):
    """kg/kmol of an elemental composition {element: atoms}."""
    table = ATOMIC_WEIGHTS if weights is None else weights
    return float(sum(table[element] * count for element, count in comp.items()))


class Species(Mapping):
    """One species' quantities. Models write what they derive into it (``record[key] = value``); a
    read of a quantity no layer holds raises KeyError naming the species."""

    def __init__(self, name, values):
        self.name = name
        self._values = {k: v for k, v in values.items() if v is not None}

    def __getitem__(self, key):
        try:
            return self._values[key]
        except KeyError:
            if key == "MW" and "comp" in self._values:
                return molecular_weight(self._values["comp"])
            raise KeyError(f"species {self.name} has no {key}") from None

    def __setitem__(self, key, value):
        self._values[key] = value

    def __contains__(self, key):
        return key in self._values or (key == "MW" and "comp" in self._values)

    def __iter__(self):
        keys = list(self._values)
        if "MW" not in self._values and "comp" in self._values:
            keys.append("MW")
        return iter(keys)

    def __len__(self):
        return sum(1 for _ in self)

    def __repr__(self):
        return f"Species({self.name!r}, {sorted(self)})"

    def meets(self, req):
        return any(k in self for k in req) if isinstance(req, OneOf) else req in self

    def unmet(self, requirements):
        """The requirements (keys, or OneOf groups of keys) this record does not meet."""
        return [r for r in requirements if not self.meets(r)]


def layered(name, own, library, forms):
    """The record of one species: the library's quantities under the case's own, except that a
    quantity in interchangeable forms (each group in forms) the case states in any form keeps none
    of the library's forms of it."""
    own = {k: v for k, v in (own or {}).items() if v is not None}
    shadowed = set()
    for group in forms:
        if any(k in own for k in group):
            shadowed.update(group)
    values = {k: v for k, v in (library or {}).items() if k not in shadowed}
    values.update(own)
    return Species(name, values)


def build_species(casesp, models):
    """{name: Species} for the case's species, in the case's order, each layered over the library
    with the interchangeable forms every model's requirements name."""
    forms = [req for model in models for req in model.needs if isinstance(req, OneOf)]
    library = database("species")
    return {name: layered(name, own, library.get(name), forms) for name, own in casesp.items()}
