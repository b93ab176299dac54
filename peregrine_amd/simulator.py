"""Simulators: what a case solves (Euler), its conserved vector and its kernel choices.

Mirrors ``peregrine/simulator/euler/simulator.py:19-161``:
Q = [rho, rho u, rho v, rho w, rho E, rho Y_0 .. rho Y_{ns-2}], the primitives
[p, u, v, w, T, Y_0 .. Y_{ns-2}], the last species what the others leave; the
advective flux is chosen by ``adv-flux`` (recon, limiter, primary); the
chemistry integration by ``chem-integration.kind``.
"""

import hashlib

import numpy as np

from peregrine_amd import mechgen, native
from peregrine_amd.config import ConfigError
from peregrine_amd.misc import find_subclass
from peregrine_amd.mixture import get_mixture

EOS = {"cpg": 0, "tpg": 1, "real-gas": 2, "peng-robinson": 2}
RECON = {"constant": 0, "muscl": 1}
LIMITER = {None: 0, "koren": 1, "minmod": 2, "mc": 3, "superbee": 4, "vanleer": 5}
PRIMARY = {"kepaec": 0, "hllc": 1, "rusanov": 2}
TRANS = {None: -1, "constant": 0, "kinetic-theory": 1, "chung": 2}


class IoVars:
    def __init__(self, sim):
        self.sim = sim
        sp = sim.mixture.spnames
        self.privars = ["p", "u", "v", "w", "T", *sp[:-1]]
        self.exported = ["rho", *self.privars, sp[-1]]

    def data(self, blk, names):
        """Named fields of a block on the host, from Q and the state (as the reference's export)."""
        Q, state = blk.Q.get(), blk.state.get()
        rho = Q[..., 0]
        sp = self.sim.mixture.spnames
        with np.errstate(divide="ignore", invalid="ignore"):
            rhoinv = 1.0 / rho
            values = {"p": state[..., 0], "T": state[..., 1], "rho": rho}
            for n, name in enumerate("uvw"):
                values[name] = Q[..., 1 + n] * rhoinv
            for n, name in enumerate(sp[:-1]):
                values[name] = Q[..., 5 + n] * rhoinv
            values[sp[-1]] = np.ones_like(rho) - sum((values[n] for n in sp[:-1]), np.zeros_like(rho))
        return {name: values[name] for name in names}


class BaseSimulator:
    name = None

    def __init__(self, cfg):
        self.cfg = cfg


class EulerSimulator(BaseSimulator):
    name = "euler"
    viscous = False
    # the CFL reduction's three speeds, in its order (simulator/euler/simulator.py:23)
    cflspeeds = ("acoustic", "convective", "combined")

    def __init__(self, cfg):
        super().__init__(cfg)
        self.mixture = get_mixture(cfg)
        adv = cfg["adv-flux"]
        if adv["formulation"] != "single" or adv["secondary"] or adv["correction"] or adv["switch"]:
            raise ConfigError("adv-flux", "this build has the single formulation with no switch or correction")
        if adv["recon"] not in RECON:
            raise ConfigError("adv-flux.recon", f"one of {sorted(RECON)}")
        if adv["primary"] not in PRIMARY:
            raise ConfigError("adv-flux.primary", f"one of {sorted(PRIMARY)}")
        if adv["recon"] == "muscl" and adv["limiter"] not in LIMITER or (adv["recon"] == "muscl" and not adv["limiter"]):
            raise ConfigError("adv-flux.limiter", f"muscl needs one of {sorted(k for k in LIMITER if k)}")
        if adv["recon"] == "muscl" and adv["primary"] == "kepaec":
            raise ConfigError("adv-flux", "kepaec is the scheme of the two cells as they are; it takes recon constant")
        self.recon, self.limiter, self.primary = adv["recon"], adv["limiter"], adv["primary"]
        ci = cfg["chem-integration"]
        kind = ci["kind"]
        if kind not in (None, "coupled", "sub-stepped", "implicit"):
            raise ConfigError("chem-integration.kind", "this build has coupled, sub-stepped and implicit chemistry")
        if kind and self.mixture.chemistry is None:
            raise ConfigError("chem-integration.kind", f"{kind} needs mixture.chemistry")
        self.chemistry = bool(kind)
        self.chemkind = kind
        self.substep = None
        if kind == "sub-stepped":
            # chemint/substepped.py:22-33
            cap, bulk, bins, tol = (ci[k] for k in ("max-substeps", "bulk-substeps", "stiff-bins",
                                                    "equilibrium-tolerance"))
            if not isinstance(cap, int) or isinstance(cap, bool) or cap < 1:
                raise ConfigError("chem-integration.max-substeps", f"{cap!r} is not a count of one or more")
            if not isinstance(bulk, int) or isinstance(bulk, bool) or not 1 <= bulk <= cap:
                raise ConfigError("chem-integration.bulk-substeps", f"{bulk!r} is not a count from one to max-substeps")
            if not isinstance(bins, int) or isinstance(bins, bool) or not 1 <= bins <= 16:
                raise ConfigError("chem-integration.stiff-bins", f"{bins!r} is not a count from one to sixteen")
            if not isinstance(tol, (int, float)) or isinstance(tol, bool) or not 0.0 <= tol < 1.0:
                raise ConfigError("chem-integration.equilibrium-tolerance",
                                  f"{tol!r} is not a mass fraction below one, zero for no exit")
            self.substep = dict(cap=cap, bulk=bulk, bins=bins, eqtol=float(tol))
        self.implicit = None
        if kind == "implicit":
            # the linearly implicit Euler form (no counterpart in the reference): its parts per step,
            # and the refinement of each FP32-factored solve against the FP64 matrix
            parts, sweeps, rtol = (ci[k] for k in ("implicit-parts", "max-sweeps", "refine-tolerance"))
            if not isinstance(parts, int) or isinstance(parts, bool) or parts < 1:
                raise ConfigError("chem-integration.implicit-parts", f"{parts!r} is not a count of one or more")
            if not isinstance(sweeps, int) or isinstance(sweeps, bool) or sweeps < 1:
                raise ConfigError("chem-integration.max-sweeps", f"{sweeps!r} is not a count of one or more")
            if not isinstance(rtol, (int, float)) or isinstance(rtol, bool) or not 0.0 < rtol < 1.0:
                raise ConfigError("chem-integration.refine-tolerance", f"{rtol!r} is not a share between 0 and 1")
            self.implicit = dict(parts=parts, sweeps=sweeps, tol=float(rtol))
        self.iovars = IoVars(self)

    @property
    def ns(self):
        return self.mixture.ns

    @property
    def ne(self):
        return 4 + self.mixture.ns

    @property
    def nstate(self):
        """The state fields per cell: p, T, 1/rho, rho, e, cp, gamma, c, and Z for a real gas."""
        return len(self.mixture.eos.fields)

    @property
    def ng(self):
        # the stencil's reach: MUSCL reads two cells each side of a face
        return 2 if self.recon == "muscl" else 1

    def ics(self):
        ic = self.cfg["ics"]
        Y, names = ic["Y"], self.mixture.spnames
        unknown = sorted(set(Y) - set(names))
        if unknown:
            raise ConfigError("ics.Y", f"names {unknown}, not species of {names}")
        if sum(Y.values()) > 1.0 + 1e-12:
            raise ConfigError("ics.Y", "sums to more than one")
        return [ic[k] for k in ("p", "u", "v", "w", "T")] + [Y.get(n, 0.0) for n in names[:-1]]

    def report(self):
        return (
            f"  Simulator: {self.name}\n  Species: {self.mixture.spnames}\n"
            f"  Mixture: {self.mixture.described()}\n"
            f"  Advective flux: {self.primary} / {self.recon}{' ' + self.limiter if self.limiter else ''}\n"
        )


class NavierStokesSimulator(EulerSimulator):
    """``peregrine/simulator/navstokes/simulator.py:10-60``: the Euler terms plus the viscous
    flux (``diff-flux.formula``: alpha-damping) from the transport state (mu, kappa, a
    diffusivity per species), the species enthalpies and the gradients of every primitive but
    the pressure (u, v, w, T, the carried Y: ngrad = ne - 1)."""

    name = "navier-stokes"
    viscous = True

    def __init__(self, cfg):
        super().__init__(cfg)
        if self.mixture.trans is None:
            raise ConfigError("mixture.trans", "a viscous case names a transport model")
        diff = cfg["diff-flux"]
        if diff["formula"] != "alpha-damping":
            raise ConfigError("diff-flux.formula", "this build has alpha-damping")
        self.alpha = float(dict(diff["values"] or {}).get("alpha", 1.0))

    @property
    def ngrad(self):
        return self.ne - 1

    def report(self):
        return super().report() + f"  Viscous flux: alpha-damping (alpha {self.alpha})\n"


def get_simulator(cfg):
    name = cfg["simulation"]["simulator"]
    try:
        return find_subclass(BaseSimulator, name=name)(cfg)
    except KeyError:
        raise ConfigError("simulation.simulator", f"{name!r}: this build has euler and navier-stokes") from None


def case_mechanism(mix):
    """A reacting mixture's mechanism as the library takes it: (name, records, third-body sets,
    scalars, signature), the name the one a library built for the case gives it."""
    records, scalars = mix.chemistry.records(mix.species)
    sets = mix.chemistry.thirdbodies(mix.species)
    sig = mechgen.signature(records, sets, mix.ns)
    return "case_" + hashlib.sha256(sig.encode()).hexdigest()[:12], records, sets, scalars, sig


class NativeCase:
    """The constant tables of a case on the device and the library's case handle. The handle is
    made by the main library when it is compiled for the case, else by a library built for the
    case (``build.build_variant``, cached by source hash, as the reference compiles each case's
    kernels on demand: ``peregrine/backend/base/compiler.py:42-64``); ``L`` is the library every
    call on the handle goes to."""

    def __init__(self, sim, backend):
        import ctypes as C

        mix = sim.mixture
        ns = mix.ns
        self.L = native.lib()
        layout, size = native.const_layout(ns, self.L)
        if size < 0:
            raise native.NativeError(f"no constant layout for {ns} species")
        k = mix.constants()
        table = np.zeros(size)
        for name, (off, count) in layout.items():
            if name in k:
                table[off : off + count] = np.asarray(k[name], float).ravel()[:count]
        self.consts = backend.tensor(table)
        self.chem = None
        mech, sig = None, None
        if mix.chemistry is not None:
            name, records, sets, scalars, sig = case_mechanism(mix)
            mech = (name, records, sets)
            self.chem = backend.tensor(mechgen.pack(records, sets, scalars, ns))
        desc = native.CaseDesc()
        desc.ns, desc.eos = ns, EOS[mix.eos.name]
        desc.recon, desc.limiter, desc.primary = RECON[sim.recon], LIMITER[sim.limiter], PRIMARY[sim.primary]
        desc.trans = TRANS[mix.trans.name if mix.trans is not None else None]
        desc.diffusion = mix.diffusion.kind if mix.diffusion is not None else 0
        self.desc = desc
        self.handle = self._create(self.L, sig)
        if self.handle is None:
            if not backend.cfg["backend-mi355x"]["jit"]:
                have = ", ".join(f"{n} ({s})" for n, s, _ in native.mechanisms())
                raise native.NativeError(
                    f"the library is not compiled for this case ({ns} species, eos {mix.eos.name}"
                    + (f", mechanism {sig}; compiled: {have}" if sig else "")
                    + "); set backend-mi355x.jit to build one for it, or add the mechanism to "
                    "peregrine_amd/data/mechanisms/compiled.yaml and rebuild"
                )
            from peregrine_amd import build

            self.L = native.lib_at(build.build_variant(ns, realgas=desc.eos == 2, mechanism=mech))
            self.handle = self._create(self.L, sig)
            if self.handle is None:
                raise native.NativeError(f"the library built for this case does not take it ({ns} species)")

    def _create(self, L, sig):
        """The case's handle from library L, or None when L is not compiled for the case."""
        import ctypes as C

        desc = self.desc
        desc.mechanism = -1
        if sig is not None:
            desc.mechanism = L.pga_mechanism_find(sig.encode())
            if desc.mechanism < 0:
                return None
        handle = C.c_void_p()
        chem_ptr = self.chem.data_ptr() if self.chem is not None else None
        rc = L.pga_case_create(C.byref(desc), self.consts.data_ptr(), chem_ptr, C.byref(handle))
        if rc == native.PGA_EUNSUP:
            return None
        native.check(rc, "pga_case_create", L)
        if not L.pga_stage_supported(handle):
            L.pga_case_destroy(handle)
            return None
        return handle

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.L.pga_case_destroy(h)
            self.handle = None
