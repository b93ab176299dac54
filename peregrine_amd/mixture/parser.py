"""Reader for Cantera YAML mechanisms: species thermo and the reaction list.

Restates ``peregrine/simulator/mixture/parser/base.py:9-263``,
``parser/cantera.py:17-27`` and ``parser/units.py:1-67``: the file is loaded
with YAML booleans switched off (so a species ``NO`` stays a name), reaction
sides are split into stoichiometric dicts, third bodies and falloff colliders
are classified, explicit orders override stoichiometry, and Arrhenius
parameters are converted to SI (kmol, m, s, activation energy as T in K).
"""

import re

import yaml

from peregrine_amd.mixture.constants import Ru, Tref, avogadro, debye, kB
from peregrine_amd.mixture.species import Species

# J/kmol per unit of energy
ENERGY = {"cal/mol": 4184.0, "kcal/mol": 4.184e6, "J/mol": 1.0e3, "kJ/mol": 1.0e6, "J/kmol": 1.0, "K": Ru}
LENGTH = {"cm": 1.0e-2, "m": 1.0, "angstrom": 1.0e-10}
QUANTITY = {"mol": 1.0e-3, "kmol": 1.0, "molec": 1.0 / avogadro}
TRANSPORT = {
    "well-depth": ("well", kB),
    "diameter": ("diam", 1.0e-10),
    "dipole": ("dipole", debye),
    "polarizability": ("polarize", 1.0e-30),
}
# which reaction kinds carry a third body, a falloff blend, Troe's blend, and the body's share of the order
KINDS = {
    "elementary": dict(thirdbody=False, falloff=False, troe=False, bodyorder=0),
    "three-body": dict(thirdbody=True, falloff=False, troe=False, bodyorder=1),
    "lindemann": dict(thirdbody=True, falloff=True, troe=False, bodyorder=0),
    "troe": dict(thirdbody=True, falloff=True, troe=True, bodyorder=0),
}


class _Loader(yaml.SafeLoader):
    pass


_Loader.yaml_implicit_resolvers = {
    k: [(tag, rx) for tag, rx in v if tag != "tag:yaml.org,2002:bool"]
    for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()
}


def _factor(table, unit):
    try:
        return table[unit]
    except KeyError:
        raise KeyError(f"unit {unit!r}; one of {', '.join(table)}") from None


def _molar(text):
    value, unit = str(text).split()
    return float(value) / _factor(QUANTITY, unit.split("/")[1])


def _truthy(sect, key):
    v = sect.get(key, False)
    if isinstance(v, str):
        return v.strip().lower() in ("true", "yes", "on")
    return bool(v)


def _side(text):
    species, hasm = {}, False
    text = re.sub(r"\(\s*\+[^)]+\)", "", text).strip()
    for term in (t.strip() for t in text.split("+")):
        if not term:
            continue
        if term == "M":
            if hasm:
                raise ValueError("more than one generic third body M on a side")
            hasm = True
            continue
        m = re.match(r"^(\d+\.?\d*)\s+(.+)$", term)
        coeff, name = (float(m.group(1)), m.group(2).strip()) if m else (1.0, term)
        species[name] = species.get(name, 0.0) + coeff
    return species, hasm


def _equation(equation):
    equation = equation.split("#")[0].strip()
    colliders = {c.strip() for c in re.findall(r"\(\s*\+([^)]+)\)", equation)}
    if len(colliders) > 1:
        raise ValueError(f"inconsistent falloff colliders in {equation!r}")
    collider = colliders.pop() if colliders else None
    if "<=>" in equation:
        lhs, rhs = equation.split("<=>")
        reversible = True
    elif "=>" in equation:
        lhs, rhs = equation.split("=>")
        reversible = False
    else:
        raise ValueError(f"cannot parse equation {equation!r}")
    reactants, lm = _side(lhs)
    products, rm = _side(rhs)
    if lm != rm:
        raise ValueError(f"third body M must be on both sides: {equation!r}")
    return reactants, products, reversible, lm, collider


class CanteraMechanism:
    suffixes = (".yaml", ".yml")
    defunits = {"length": "m", "quantity": "kmol", "activation-energy": "J/kmol"}

    def __init__(self, path, species=None):
        with open(path) as f:
            data = yaml.load(f, Loader=_Loader)
        if species is None:
            if "phases" not in data:
                raise ValueError(f"{path} has no phases section; give the species list")
            species = list(data["phases"][0]["species"])
        lookup = {sp["name"]: sp for sp in data.get("species", [])}
        weights = dict(Species.atomwts)
        for el in data.get("elements", []):
            weights[el["symbol"]] = float(el["atomic-weight"])
        self.species = {}
        for name in species:
            if name not in lookup:
                raise KeyError(f"species {name!r} not in {path}")
            self.species[name] = self._species(lookup[name], weights)
        self.reactions = []
        units = {**self.defunits, **data.get("units", {})}
        self.energy = _factor(ENERGY, units["activation-energy"])
        self.length = _factor(LENGTH, units["length"])
        self.quantity = _factor(QUANTITY, units["quantity"])
        for raw in data.get("reactions") or []:
            rxn = self._reaction(raw)
            unknown = (set(rxn["reactants"]) | set(rxn["products"]) | set(rxn["efficiencies"])) - set(self.species)
            if unknown:
                raise KeyError(f"reaction {rxn['equation']!r} names species not in the phase: {sorted(unknown)}")
            self.reactions.append(rxn)

    def arrhenius(self, sect, order):
        A = float(sect["A"]) * (self.length**3 / self.quantity) ** (order - 1)
        return A, float(sect["b"]), float(sect["Ea"]) * self.energy / Ru

    def _reaction(self, raw):
        equation = raw["equation"]
        reactants, products, reversible, hasm, collider = _equation(equation)
        kind = raw.get("type")
        effs = {k: float(v) for k, v in raw.get("efficiencies", {}).items()}
        defeff = float(raw.get("default-efficiency", 1.0))
        kind = {"Arrhenius": "elementary", "three-body-Arrhenius": "three-body", "Lindemann": "falloff", "Troe": "falloff"}.get(kind, kind)
        hastroe = "Troe" in raw
        if kind == "falloff":
            rtype = "troe" if hastroe else "lindemann"
            if collider is None:
                raise ValueError(f"falloff reaction without a (+M) collider: {equation!r}")
            if collider != "M":
                if "default-efficiency" in raw and defeff != 0.0:
                    raise ValueError(f"default efficiency with an explicit collider: {equation!r}")
                if effs and (len(effs) != 1 or collider not in effs):
                    raise ValueError(f"incompatible collider definitions: {equation!r}")
                defeff = 0.0
                effs = effs or {collider: 1.0}
        elif kind == "three-body" or (kind is None and hasm):
            rtype = "three-body"
            if collider is not None:
                raise ValueError(f"(+{collider}) needs type falloff: {equation!r}")
            if not hasm:
                if len(effs) != 1:
                    raise ValueError(f"an explicit third body needs one efficiency: {equation!r}")
                (sp,) = effs
                for side in (reactants, products):
                    if side.get(sp, 0) < 1:
                        raise ValueError(f"collider {sp!r} must be on both sides: {equation!r}")
                    if side[sp] == 1:
                        del side[sp]
                    else:
                        side[sp] -= 1
                defeff = 0.0
        elif kind in (None, "elementary"):
            rtype = "elementary"
            if hasm or collider is not None:
                raise ValueError(f"elementary reaction with a third body: {equation!r}")
            if "efficiencies" in raw or "default-efficiency" in raw:
                raise ValueError(f"efficiencies without a third body: {equation!r}")
        else:
            raise ValueError(f"unsupported reaction type {kind!r}: {equation!r}")
        orders = {sp: float(v) for sp, v in raw.get("orders", {}).items()}
        if orders:
            if reversible:
                raise ValueError(f"orders given for a reversible reaction: {equation!r}")
            if any(sp not in reactants for sp in orders) and not _truthy(raw, "nonreactant-orders"):
                raise ValueError(f"order for a non-reactant: {equation!r}")
            if any(v < 0 for v in orders.values()) and not _truthy(raw, "negative-orders"):
                raise ValueError(f"negative order: {equation!r}")
        fwd = dict(reactants)
        fwd.update(orders)
        order = sum(fwd.values()) + KINDS[rtype]["bodyorder"]
        rxn = dict(
            equation=equation,
            type=rtype,
            reversible=reversible,
            duplicate=_truthy(raw, "duplicate"),
            reactants=reactants,
            products=products,
            fwd=fwd,
            efficiencies=effs,
            defeff=defeff,
        )
        if KINDS[rtype]["falloff"]:
            rxn["rate"] = self.arrhenius(raw["high-P-rate-constant"], order)
            rxn["lowrate"] = self.arrhenius(raw["low-P-rate-constant"], order + 1)
        else:
            rxn["rate"] = self.arrhenius(raw["rate-constant"], order)
        if hastroe:
            troe = {"T2": 0.0, **raw["Troe"]}
            rxn["troe"] = [float(troe[k]) for k in ("A", "T3", "T1", "T2")]
        return rxn

    @staticmethod
    def _species(raw, weights):
        sp = {}
        if "composition" in raw:
            sp["comp"] = {k: float(v) for k, v in raw["composition"].items()}
            sp["MW"] = Species.molwt(sp["comp"], weights)
        thermo = raw.get("thermo")
        if thermo and thermo["model"] == "constant-cp":
            cp0 = _molar(thermo["cp0"])
            sp["cp0"] = cp0 / sp["MW"]
            T0 = float(str(thermo.get("T0", "298.15 K")).split()[0])
            h0 = _molar(thermo["h0"]) if "h0" in thermo else 0.0
            sp["dHf298"] = h0 + cp0 * (Tref - T0)
        elif thermo:
            if thermo["model"].upper() != "NASA7":
                raise ValueError(f"unsupported thermo model {thermo['model']} for {raw['name']}")
            ranges = [float(r) for r in thermo["temperature-ranges"]]
            data = [list(map(float, c)) for c in thermo["data"]]
            low, high = (data[0], data[0]) if len(data) == 1 else data[:2]
            Tmid = ranges[1] if len(data) > 1 else ranges[-1]
            sp["NASA7"] = {"T-range": [ranges[0], Tmid, ranges[-1]], "high": high, "low": low}
        if "transport" in raw:
            tr = raw["transport"]
            sp["geometry"] = tr["geometry"]
            for key, (ours, scale) in TRANSPORT.items():
                sp[ours] = float(tr.get(key, 0.0)) * scale
            sp["zrot"] = float(tr.get("rotational-relaxation", 0.0))
            if "acentric-factor" in tr:
                sp["acentric"] = float(tr["acentric-factor"])
        return sp


def read_mechanism(path, species=None):
    return CanteraMechanism(path, species)
