"""The mixture: species, models, and the flat constant tables of a case.

Follows ``peregrine/simulator/mixture/mixture.py:18-115``: the species come
from the case (a dict, a list of library names, or a mechanism file), the eos
is populated first and every other model reads what it stored; the last
species is the one the solution does not carry.
"""

from pathlib import Path

import numpy as np

from peregrine_amd.mixture.constants import Ru
from peregrine_amd.mixture.eos import EOS_MODELS
from peregrine_amd.mixture.finiterate import FiniteRate
from peregrine_amd.mixture.parser import read_mechanism
from peregrine_amd.mixture.species import DATA, build_species
from peregrine_amd.mixture.transport import TRANSPORT_MODELS, DIFFUSION_MODELS, MIXRULE_MODELS


class Mixture:
    def __init__(self, cfgsect, root=None):
        self.cfgsect = cfgsect
        if cfgsect["eos"] not in EOS_MODELS:
            raise ValueError(f"eos {cfgsect['eos']!r}: this build has {sorted(EOS_MODELS)}")
        self.eos = EOS_MODELS[cfgsect["eos"]](cfgsect)
        self.trans = self.diffusion = self.mixrule = None
        if cfgsect["trans"]:
            for key, models in (("trans", TRANSPORT_MODELS), ("diffusion", DIFFUSION_MODELS),
                                ("mixing-rule", MIXRULE_MODELS)):
                if cfgsect[key] not in models:
                    raise ValueError(f"mixture.{key} {cfgsect[key]!r}: this build has {sorted(models)}")
            self.trans = TRANSPORT_MODELS[cfgsect["trans"]](cfgsect)
            self.diffusion = DIFFUSION_MODELS[cfgsect["diffusion"]](cfgsect)
            self.mixrule = MIXRULE_MODELS[cfgsect["mixing-rule"]](cfgsect)
        usersp, self.reactions = self.read(cfgsect, root)
        chem = cfgsect["chemistry"]
        if chem and chem != "finite-rate":
            raise ValueError(f"chemistry {chem!r}: this build has finite-rate")
        self.chemistry = FiniteRate(cfgsect, self.reactions) if chem else None
        if self.chemistry and cfgsect["eos"] != "tpg":
            raise ValueError("finite-rate needs the thermally perfect caloric model (eos tpg)")
        if cfgsect["trans"] == "chung" and self.eos.kind != 2:
            raise ValueError("chung's dense-gas transport reads the critical point of a real-gas eos")
        self.species = build_species(usersp, self.models)
        self.spnames = list(self.species)
        self.ns = len(self.spnames)
        for model in self.models:
            model.check(self.species)
        for model in self.models:
            model.populate(self.species)

    @property
    def models(self):
        return tuple(m for m in (self.eos, self.trans, self.diffusion, self.mixrule, self.chemistry) if m is not None)

    @property
    def carried(self):
        return self.spnames[:-1]

    def constants(self):
        MW = np.array([sp["MW"] for sp in self.species.values()], float)
        out = dict(Ru=Ru, MW=MW, MWinv=1.0 / MW, Rsp=Ru / MW)
        for m in self.models:
            if m is self.chemistry:
                continue
            out.update(m.constants(self.species))
        return out

    @staticmethod
    def read(cfgsect, root):
        mixture = cfgsect["species"]
        if isinstance(mixture, dict):
            return dict(mixture), []
        if isinstance(mixture, (list, tuple)):
            return {sp: {} for sp in mixture}, []
        locs = [] if root is None else [Path(root) / mixture]
        locs += [Path(mixture), DATA / "mechanisms" / mixture]
        for loc in locs:
            if loc.is_file():
                mech = read_mechanism(str(loc))
                return mech.species, mech.reactions
        raise FileNotFoundError(f"cannot find the mixture {mixture!r}; tried {', '.join(map(str, locs))}")

    def described(self):
        parts = [self.eos.name]
        if self.trans is not None:
            parts.append(f"{self.trans.name}/{self.diffusion.name}/{self.mixrule.name}")
        if self.chemistry is not None:
            parts.append(f"finite-rate ({len(self.reactions)} reactions)")
        return f"{self.ns} species, " + " + ".join(parts)


def get_mixture(cfg, root=None):
    return Mixture(cfg["mixture"], root)
