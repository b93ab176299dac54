"""Solver plugins: called after steps, with the reference's start/due/tnext/__call__/finalize
protocol and its cadence (niter-out or dt-out).

Mirrors ``peregrine/plugins/solver/base.py:5-26``; ``report`` prints the banner and the
seconds per iteration per cell (``plugins/solver/report.py:9-60``), ``nancheck`` stops a run
whose interior holds a non-finite value (``plugins/solver/nancheck.py:10-44``). A physics'
plugin acts with its physics instead: ``viscsponge`` raises the viscosity along a line after
every transport (``plugins/simulator/viscsponge.py:9-26``), on the device.
"""

import time

import numpy as np


class Cadence:
    """When a plugin acts: every niter-out steps, or each time the clock passes a multiple of
    dt-out; every step when neither is given (``plugins/solver/cadence.py:4-25``)."""

    def __init__(self, cfgsect):
        self.niterout = cfgsect.get("niter-out")
        self.dtout = cfgsect.get("dt-out")
        if self.niterout is None and self.dtout is None:
            self.niterout = 1

    def nreached(self, tme):
        # a step that lands on a multiple of dt-out can fall a few ulps short of it: counted reached
        tol = 8.0 * np.spacing(max(abs(tme), self.dtout))
        return int((tme + tol) / self.dtout)

    def next(self, solver):
        if self.dtout is None:
            return None
        return (self.nreached(solver.tme.get()) + 1) * self.dtout

    def due(self, solver):
        if self.dtout is None:
            return solver.nrt % self.niterout == 0
        tme, dt = solver.tme.get(), solver.dt.get()
        return self.nreached(tme) != self.nreached(tme - dt)


class BasePlugin:
    name = None

    def __init__(self, cfgsect, solver):
        self.cfg = dict(cfgsect or {})
        self.cadence = Cadence(self.cfg)

    def start(self, solver):
        pass

    def due(self, solver):
        return self.cadence.due(solver)

    def tnext(self, solver):
        return self.cadence.next(solver)

    def __call__(self, solver):
        pass

    def finalize(self, solver):
        pass


class Report(BasePlugin):
    """At the start the case; after each step it is due on the step number, the time, the step
    size and the largest CFL numbers (combined, then each speed); at the end the wall time over
    this run's steps and the seconds per step per cell."""

    name = "report"

    def start(self, solver):
        print(" >>> peregrine-mi355x <<<", flush=True)
        print(repr(solver), flush=True)
        self.t0, self.n0 = time.perf_counter(), solver.nrt

    def __call__(self, solver):
        cfl = solver.cflmax()
        dt = solver.dt.get()
        speeds = "".join(f"         {n:10s}: {cfl[n] * dt:.3f}\n" for n in solver.sim.cflspeeds if n != "combined")
        print(f" >>> --------- nrt: {solver.nrt:<6} ---------- <<<\n"
              f"     tme: {solver.tme.get():.6E} s\n"
              f"     dt : {dt:.6E} s\n"
              f"     MAX CFL       : {cfl['combined'] * dt:.3f}\n" + speeds, flush=True)

    def finalize(self, solver):
        solver.synchronize()
        steps = max(solver.nrt - self.n0, 1)
        elapsed = time.perf_counter() - self.t0
        hrs, rem = divmod(int(elapsed), 3600)
        mins, secs = divmod(rem, 60)
        print(f"Simulation time: {hrs}h : {mins}m : {secs}s over {steps} steps\n"
              f"Seconds/Iteration/Cell: {elapsed / steps / max(solver.ncells, 1):.3e}", flush=True)


class NanCheckError(RuntimeError):
    pass


class NanCheck(BasePlugin):
    name = "nancheck"

    def __call__(self, solver):
        torch = solver.backend.torch
        ng = solver.ng
        bad = [blk for blk in solver.blocks if not bool(torch.isfinite(blk.Q.t[:, ng:-ng, ng:-ng, ng:-ng]).all())]
        if not bad:
            return
        # where: the centres of the cells holding a non-finite value, a file per block
        for blk in bad:
            Q = blk.Q.get()[blk.interior]
            where = np.where((~np.isfinite(Q)).any(axis=-1))
            cells = blk.xcc.get()[blk.interior][where]
            with open(f"nans_{blk.nblki}.log", "w") as f:
                f.write(f"Non-finite state, block {blk.nblki}\n")
                for x, y, z in cells:
                    f.write(f"x = {x} y = {y} z = {z}\n")
        raise NanCheckError(f"non-finite state at step {solver.nrt} in blocks {[b.nblki for b in bad]}; "
                            "see the nans_<block>.log files")


class ViscSponge:
    """The viscosity raised along the line from origin to ending, to multiplier times itself at
    the end and beyond, over the cells the transport has just made (interior and first halo
    layer); a Navier-Stokes case only."""

    name = "viscsponge"
    simulators = ("navier-stokes",)

    def __init__(self, cfgsect, solver):
        from peregrine_amd import native

        cfgsect = dict(cfgsect or {})
        missing = [k for k in ("origin", "ending", "multiplier") if k not in cfgsect]
        if missing:
            raise ValueError(f"plugins.viscsponge needs {missing}")
        origin, ending = (np.array(cfgsect[k], float) for k in ("origin", "ending"))
        length = float(np.linalg.norm(ending - origin))
        if not length > 0.0:
            raise ValueError("plugins.viscsponge: origin and ending must differ")
        normal = (ending - origin) / length
        self.line = native.dbls([*normal, float(origin @ normal), length, float(cfgsect["multiplier"])])


PLUGINS = {p.name: p for p in (Report, NanCheck)}
PHYSICS_PLUGINS = {p.name: p for p in (ViscSponge,)}


def claim(cfgsect, solver):
    """The solver's plugins and the physics' plugins the plugins section names."""
    unknown = sorted(set(cfgsect) - set(PLUGINS) - set(PHYSICS_PLUGINS))
    if unknown:
        raise ValueError(f"{unknown}: no plugin of that name in this build (have "
                         f"{sorted(PLUGINS) + sorted(PHYSICS_PLUGINS)})")
    physics = []
    for name, sect in cfgsect.items():
        if name in PHYSICS_PLUGINS:
            cls = PHYSICS_PLUGINS[name]
            if solver.sim.name not in cls.simulators:
                raise ValueError(f"plugins.{name} is for {', '.join(cls.simulators)}, not {solver.sim.name}")
            physics.append(cls(sect, solver))
    return [PLUGINS[name](sect, solver) for name, sect in cfgsect.items() if name in PLUGINS], physics
