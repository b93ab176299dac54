"""Solver plugins: called after steps, with the reference's start/due/tnext/__call__/finalize
protocol and its cadence (niter-out or dt-out).

Mirrors ``peregrine/plugins/solver/base.py:5-26``; ``report`` prints the banner and the
seconds per iteration per cell (``plugins/solver/report.py:9-60``), ``nancheck`` stops a run
whose interior holds a non-finite value (``plugins/solver/nancheck.py:10-44``), ``trace``
samples the primitives at chosen points into csv files (``plugins/solver/trace.py:30-140``). A
physics' plugin acts with its physics instead: ``viscsponge`` raises the viscosity along a line after
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
        """The all-finite reduction over each block's interior on the device (pga_allfinite,
        ``plugins/solver/kernels/allfinite.mako:3-19``), one flag per block read back together."""
        torch = solver.backend.torch
        flags = torch.zeros(len(solver.blocks), dtype=torch.int32, device=solver.backend.device)
        for n, blk in enumerate(solver.blocks):
            solver._call("pga_allfinite", blk.Q.view(), solver.sim.ne, blk.interior_box(), flags.data_ptr() + 4 * n)
        ok = flags.cpu().numpy()
        bad = [blk for blk, f in zip(solver.blocks, ok) if not f]
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
        # the state it stopped on, by the writer among the plugins if there is one
        for plugin in solver.plugins:
            if isinstance(plugin, Writer):
                plugin(solver)
        raise NanCheckError(f"non-finite state at step {solver.nrt} in blocks {[b.nblki for b in bad]}; "
                            "see the nans_<block>.log files")


def _trace_points(spec):
    """A trace's points: a point p0, a line of n points from p0 to p1, or a plane of n01 by n02
    points spanned from p0 by p1 and p2 (``plugins/solver/trace.py:12-27``)."""
    p0 = np.array(spec["p0"], float)
    if spec["type"] == "point":
        return p0[None, :]
    if spec["type"] == "line":
        return p0 + np.linspace(0.0, 1.0, spec["n"])[:, None] * (np.array(spec["p1"], float) - p0)
    a = np.linspace(0.0, 1.0, spec["n01"])[:, None, None]
    b = np.linspace(0.0, 1.0, spec["n02"])[None, :, None]
    p1, p2 = (np.array(spec[k], float) for k in ("p1", "p2"))
    return (p0 + a * (p1 - p0) + b * (p2 - p0)).reshape(-1, 3)


class Trace(BasePlugin):
    """The primitive variables (p, u, v, w, T, the carried Y) at the cells whose centres are
    nearest chosen points, a row per point appended to each trace's csv when it is due, with the
    time and the cell centre (``plugins/solver/trace.py:30-140``). A trace may carry its own
    cadence; the others act on the section's."""

    name = "trace"

    def __init__(self, cfgsect, solver):
        super().__init__(cfgsect, solver)
        self.specs = dict(self.cfg.get("traces") or {})
        for tag, spec in self.specs.items():
            if spec.get("type") not in ("point", "line", "plane"):
                raise ValueError(f"trace {tag}: type is point, line or plane")
            if "file" not in spec:
                raise ValueError(f"trace {tag}: names the file it writes")
        self.cadences = {tag: Cadence(spec) if {"niter-out", "dt-out"} & set(spec) else self.cadence
                         for tag, spec in self.specs.items()}

    def due(self, solver):
        return any(c.due(solver) for c in self.cadences.values())

    def tnext(self, solver):
        times = [c.next(solver) for c in self.cadences.values()]
        return min((t for t in times if t is not None), default=None)

    def start(self, solver):
        from pathlib import Path

        names = ["p", "u", "v", "w", "T", *solver.sim.mixture.spnames[:-1]]
        centres = [(blk, blk.xcc.get()[blk.interior]) for blk in solver.blocks]
        self.cells = {}
        for tag, spec in self.specs.items():
            cells = []
            for p in _trace_points(spec):
                best = None
                for blk, c in centres:
                    d = ((c - p) ** 2).sum(axis=-1)
                    m = np.unravel_index(int(d.argmin()), d.shape)
                    if best is None or d[m] < best[0]:
                        best = (d[m], blk, tuple(int(x) + blk.ng for x in m), c[m])
                cells.append(best[1:])
            self.cells[tag] = cells
            if not Path(spec["file"]).exists():
                Path(spec["file"]).write_text("time, x, y, z, " + ", ".join(names) + "\n")

    def __call__(self, solver):
        torch = solver.backend.torch
        tme = solver.tme.get()
        for tag, spec in self.specs.items():
            if not self.cadences[tag].due(solver):
                continue
            cells = self.cells[tag]
            rows = [None] * len(cells)
            for blk in {id(c[0]): c[0] for c in cells}.values():
                mine = [n for n, c in enumerate(cells) if c[0] is blk]
                i, j, k = (torch.tensor([cells[n][1][a] for n in mine], device=blk.Q.t.device) for a in range(3))
                # one gather per block; the device arrays hold (component, k, j, i)
                Q, st = blk.Q.t[:, k, j, i].cpu().numpy(), blk.state.t[:, k, j, i].cpu().numpy()
                rhoinv = 1.0 / Q[0]
                for m, n in enumerate(mine):
                    r = rhoinv[m]
                    values = [st[0, m], Q[1, m] * r, Q[2, m] * r, Q[3, m] * r, st[1, m], *(Q[5:, m] * r)]
                    rows[n] = ", ".join(f"{v:.8e}" for v in (tme, *cells[n][2], *values))
            with open(spec["file"], "a") as f:
                f.write("\n".join(rows) + "\n")


class Writer(BasePlugin):
    """Results into a directory as numbered files (``plugins/solver/writer.py:12-36``): dir
    (made if need be), precision (single or double), basename formatted from the step n and the
    time t. Each is a complete restart (``peregrine_amd.restart``)."""

    name = "writer"

    def start(self, solver):
        from pathlib import Path

        from peregrine_amd.restart import RestartWriter

        self.dir = Path(self.cfg.get("dir", "."))
        self.dir.mkdir(parents=True, exist_ok=True)
        self.writer = RestartWriter(solver, self.cfg.get("basename", "q.{n:08d}") + ".npz",
                                    self.cfg.get("precision", "single"), cfg=solver.cfg)

    def __call__(self, solver):
        self.writer.write(solver, solver.nrt, solver.tme.get(), self.dir)


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


PLUGINS = {p.name: p for p in (Report, NanCheck, Trace, Writer)}
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
