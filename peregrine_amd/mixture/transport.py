"""Transport models: viscosity, conductivity and species diffusion constants.

Follows ``peregrine/simulator/mixture/transport/constant.py`` (the case states
mu0 and kappa0 per species), ``diffusion/lewis.py`` (a Lewis number per
species, 1 unless stated) and ``mixrule/wilke.py`` (Wilke's pair factors
phi_nm = (1 + s_n/s_m)^2 / sqrt(8 (1 + MW_n/MW_m)), s = sqrt(mu) MW^-1/4).
"""

import numpy as np

from peregrine_amd.mixture.eos import BaseModel


class ConstantProps(BaseModel):
    name = "constant"
    kind = 0

    def __init__(self, cfgsect):
        super().__init__(cfgsect)
        self.needs += ("mu0", "kappa0")

    def constants(self, species):
        return dict(mu0=self.collect("mu0", species), kappa0=self.collect("kappa0", species))


class Lewis(BaseModel):
    name = "lewis"
    kind = 0

    def constants(self, species):
        return dict(lewis=np.array([sp.get("lewis", 1.0) for sp in species.values()], float))


class Wilke(BaseModel):
    name = "wilke"
    kind = 0

    def constants(self, species):
        MW = self.collect("MW", species)
        return dict(MWqinv=MW**-0.25, wilkepair=1.0 / np.sqrt(8.0 * (1.0 + MW[:, None] / MW[None, :])))


TRANSPORT_MODELS = {m.name: m for m in (ConstantProps,)}
DIFFUSION_MODELS = {m.name: m for m in (Lewis,)}
MIXRULE_MODELS = {m.name: m for m in (Wilke,)}
