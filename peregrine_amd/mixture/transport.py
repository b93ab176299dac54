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


class KineticTheory(BaseModel):
    """Dilute-gas viscosity and conductivity from the Lennard-Jones (Stockmayer) parameters, fitted
    in ln T over the case's T-range (``peregrine/simulator/mixture/transport/kinetic.py:12-141``):

    - mu = 5/16 sqrt(pi m kB T) / (pi sigma^2 Omega(2,2)*), fitted as sqrt(mu) / T^1/4;
    - kappa by the Warnatz form of the Eucken correction: self-diffusion
      D = 3/16 sqrt(2 pi / m_r) (kB T)^3/2 / (pi sigma^2 Omega(1,1)*), Parker's rotational
      relaxation Z_rot(T) = Z_rot(298) F(298) / F(T), and translational, rotational and internal
      shares of cv; fitted as kappa / sqrt(T);

    a polar/non-polar pair's well and diameter corrected for induction, and delta* from the
    dipole. The fits take the lowest degree within refit-tol (at most refit-max-deg), as the
    caloric refits do."""

    name = "kinetic-theory"
    kind = 1

    def __init__(self, cfgsect):
        super().__init__(cfgsect)
        self.needs += ("well", "diam", "dipole", "polarize", "zrot", "geometry")
        # the conductivity reads cppoly, which the eos stores first; mupoly before kappapoly
        self.provides.update(mupoly="mupoly", kappapoly="kappapoly")

    def constants(self, species):
        from peregrine_amd.mixture.eos import NT

        return dict(mupoly=self.collect_poly("mupoly", species, NT), kappapoly=self.collect_poly("kappapoly", species, NT))

    def _params(self, species):
        from peregrine_amd.mixture.constants import avogadro, epsilon0
        from peregrine_amd.mixture.integrals import CollisionIntegrals
        from peregrine_amd.mixture.species import HALF_ROTATIONAL_DOF

        MW = self.collect("MW", species)
        well, diam = self.collect("well", species), self.collect("diam", species)
        dipole = self.collect("dipole", species)
        rotdof = np.array([HALF_ROTATIONAL_DOF[sp["geometry"]] for sp in species.values()])
        mass = MW / avogadro
        # the reduced dipole of each species with itself (the pair terms a mixture rule would read
        # are the diagonal here: a species' own viscosity and conductivity)
        delta = 0.5 * dipole**2 / (4.0 * np.pi * epsilon0 * well * diam**3)
        return dict(Ts=np.linspace(*self.trange, 50), ci=CollisionIntegrals(), MW=MW, well=well, diam=diam,
                    mass=mass, delta=delta, zrot=self.collect("zrot", species), rotdof=rotdof)

    def _fit(self, lnT, values):
        from peregrine_amd.mixture.fits import fit

        tol, deg = self.cfgsect["refit-tol"], self.cfgsect["refit-max-deg"]
        return [fit(lnT, v, tol, deg)[0] for v in values.T]

    def mupoly(self, species):
        from peregrine_amd.mixture.constants import kB

        p = self._params(species)
        Ts = p["Ts"]
        Tstar = np.outer(Ts, kB / p["well"])
        omega22 = p["ci"].omega22(Tstar, np.broadcast_to(p["delta"], Tstar.shape))
        mu = (5.0 / 16.0) * np.sqrt(np.pi * p["mass"] * kB * Ts[:, None]) / (np.pi * p["diam"] ** 2 * omega22)
        return self._fit(np.log(Ts), np.sqrt(mu / np.sqrt(Ts)[:, None]))

    def kappapoly(self, species):
        from peregrine_amd.mixture.constants import Ru, Tref, kB
        from peregrine_amd.mixture.fits import evaluate

        p = self._params(species)
        Ts = p["Ts"]
        lnT = np.log(Ts)
        T = Ts[:, None]
        sps = list(species.values())
        mu = np.array([evaluate(sp["mupoly"], lnT) for sp in sps]).T ** 2 * np.sqrt(T)
        cpR = np.array([evaluate(sp["cppoly"], lnT) for sp in sps]).T
        Tstar = np.outer(Ts, kB / p["well"])
        delta = np.broadcast_to(p["delta"], Tstar.shape)
        omega11 = p["ci"].omega22(Tstar, delta) / p["ci"].astar(Tstar, delta)
        # a species with itself: the reduced mass is half the molecular mass
        mr = 0.5 * p["mass"]
        D = (3.0 / 16.0) * np.sqrt(2.0 * np.pi / mr) * (kB * T) ** 1.5 / (np.pi * p["diam"] ** 2 * omega11)

        def parker(ts):
            return 1.0 + np.pi**1.5 / np.sqrt(ts) * (0.5 + 1.0 / ts) + (0.25 * np.pi**2 + 2.0) / ts

        MW, rotdof = p["MW"], p["rotdof"]
        fint = MW / (Ru * T) * D / mu
        A = 2.5 - fint
        zrot = p["zrot"] * parker(kB * Tref / p["well"]) / parker(Tstar)
        B = zrot + 2.0 / np.pi * (5.0 / 3.0 * rotdof + fint)
        c1 = 2.0 / np.pi * A / B
        cvint = cpR - 2.5 - rotdof
        frot = fint * (1.0 + c1)
        ftrans = 2.5 * (1.0 - c1 * rotdof / 1.5)
        kappa = mu / MW * Ru * (ftrans * 1.5 + frot * rotdof + fint * cvint)
        return self._fit(lnT, kappa / np.sqrt(T))


class ChungDenseGas(KineticTheory):
    """Chung's dense-gas corrections on the kinetic-theory dilute values
    (``peregrine/simulator/mixture/transport/chung.py:9-33``, ``kernels/chung.mako``): Chung,
    Ajlan, Lee and Starling, Ind. Eng. Chem. Res. 27, 671 (1988) as Poling et al. (5th ed.,
    9-5, 9-6, 10-5) state it, without the dipole and association terms, one-fluid mixing of
    sigma, epsilon, M and the acentric factor. Its sigma (angstrom) and epsilon/k (K) come from
    the real gas's critical point."""

    name = "chung"

    def constants(self, species):
        from peregrine_amd.mixture.species import database

        chung = database("chung")
        MW = self.collect("MW", species)
        Vc = self.collect("Vcrit", species) * MW * 1e3
        return dict(
            super().constants(species),
            chungsigma=0.809 * Vc ** (1.0 / 3.0),
            chungepsk=self.collect("Tcrit", species) / 1.2593,
            chungvisc=np.array(chung["viscosity"], dtype=np.float64),
            chungcond=np.array(chung["conductivity"], dtype=np.float64),
        )


def pair_parameters(species):
    """Every species pair's Lennard-Jones parameters by the usual combining rules: the reduced mass,
    the mean diameter, the geometric-mean well, the reduced dipole, and for a polar/non-polar pair
    the induction correction xi that deepens the well by xi^2 and shrinks the diameter by
    xi^-1/6 (Hirschfelder, Curtiss and Bird; as ``transport/kinetic.py:30-64`` combines them)."""
    from peregrine_amd.mixture.constants import avogadro, epsilon0

    def col(key):
        return np.array([sp[key] for sp in species.values()], float)

    MW, well, diam, dipole, polarize = col("MW"), col("well"), col("diam"), col("dipole"), col("polarize")
    mass = MW / avogadro
    rmass = np.outer(mass, mass) / np.add.outer(mass, mass)
    rdiam = 0.5 * np.add.outer(diam, diam)
    rwell = np.sqrt(np.outer(well, well))
    rdipole = np.sqrt(np.outer(dipole, dipole))
    rdelta = 0.5 * rdipole**2 / (4.0 * np.pi * epsilon0 * rwell * rdiam**3)
    polar = dipole > 0.0
    mixed = polar[:, None] != polar[None, :]
    n = np.arange(len(MW))
    ip = np.where(polar[:, None], n[:, None], n[None, :])
    inp = np.where(polar[:, None], n[None, :], n[:, None])
    alphastar = polarize[inp] / diam[inp] ** 3
    mustar = dipole[ip] / np.sqrt(4.0 * np.pi * epsilon0 * diam[ip] ** 3 * well[ip])
    xi = 1.0 + 0.25 * alphastar * mustar**2 * np.sqrt(well[ip] / well[inp])
    rwell = np.where(mixed, rwell * xi**2, rwell)
    rdiam = np.where(mixed, rdiam * xi ** (-1.0 / 6.0), rdiam)
    return dict(rmass=rmass, rdiam=rdiam, rwell=rwell, rdelta=rdelta)


class BinaryDiffusion(BaseModel):
    """Mixture-averaged diffusion from binary pair coefficients (``diffusion/binary.py:9-64``,
    ``diffusion/kernels/binary.mako:11-62``): each pair's Chapman-Enskog coefficient at unit
    pressure, D_nm = 3/16 sqrt(2 pi / m_nm) (kB T)^3/2 / (pi sigma_nm^2 Omega(1,1)*), fitted as
    T^3/2 / D_nm in ln T over the case's T-range; the kernel sums over the bath (the case's
    `bath` species, or all of them) and divides by the pressure.

    Table layout: pairfit[n][k] the fit of species n with its k-th partner, bathof[n][k] that
    partner's index (every other species in order for the whole bath; the bath's species, its
    own slot included, for a partial one), selffit[n] its own pair, diffmeta (width, partial)."""

    name = "binary"
    kind = 1

    def __init__(self, cfgsect):
        super().__init__(cfgsect)
        self.needs += ("well", "diam", "dipole", "polarize")

    def bath(self, species):
        names = list(species)
        chosen = self.cfgsect["bath"]
        unknown = sorted(set(chosen or ()) - set(names))
        if unknown:
            raise ValueError(f"mixture.bath names {unknown}, not species of the case")
        return np.array([names.index(n) for n in (names if chosen is None else chosen)], dtype=np.int32)

    def dijinv(self, species):
        """The fit of T^3/2 / D_nm (at unit pressure) per unordered pair, in triu order."""
        from peregrine_amd.mixture.constants import kB
        from peregrine_amd.mixture.fits import fit
        from peregrine_amd.mixture.integrals import CollisionIntegrals

        pp = pair_parameters(species)
        Ts = np.linspace(*self.trange, 50)
        ns = len(species)
        k, j = np.triu_indices(ns)
        Tstar = np.outer(Ts, kB / pp["rwell"][k, j])
        delta = np.broadcast_to(pp["rdelta"][k, j], Tstar.shape)
        ci = CollisionIntegrals()
        omega11 = ci.omega22(Tstar, delta) / ci.astar(Tstar, delta)
        kT = kB * Ts[:, None]
        D = 3.0 / 16.0 * np.sqrt(2.0 * np.pi / pp["rmass"][k, j]) * kT**1.5 / (np.pi * pp["rdiam"][k, j] ** 2 * omega11)
        tol, deg = self.cfgsect["refit-tol"], self.cfgsect["refit-max-deg"]
        return [fit(np.log(Ts), Dinv, tol, deg)[0] for Dinv in (Ts[:, None] ** 1.5 / D).T]

    def constants(self, species):
        from peregrine_amd.mixture.eos import NT
        from peregrine_amd.mixture.fits import padded

        ns = len(species)
        dijinv = padded(self.dijinv(species), width=NT)
        bath = self.bath(species)
        slot = np.zeros((ns, ns), dtype=int)
        slot[np.triu_indices(ns)] = np.arange(ns * (ns + 1) // 2)
        slot = np.maximum(slot, slot.T)
        whole = len(bath) == ns
        rows = [[m for m in range(ns) if m != n] if whole else list(bath) for n in range(ns)]
        width = len(rows[0])
        pairfit = np.zeros((ns, ns, NT))
        bathof = np.zeros((ns, ns))
        for n, r in enumerate(rows):
            for k, m in enumerate(r):
                bathof[n, k] = m
                if m != n:
                    pairfit[n, k] = dijinv[slot[n, m]]
        return dict(pairfit=pairfit, selffit=dijinv[np.diag(slot)], bathof=bathof, bath=bath,
                    diffmeta=np.array([width, 0.0 if whole else 1.0]), dijinv=dijinv)


TRANSPORT_MODELS = {m.name: m for m in (ConstantProps, KineticTheory, ChungDenseGas)}
DIFFUSION_MODELS = {m.name: m for m in (Lewis, BinaryDiffusion)}
MIXRULE_MODELS = {m.name: m for m in (Wilke,)}
