"""Solver plugins: called after steps, with the reference's start/due/__call__/finalize protocol.

Mirrors ``peregrine/plugins/solver/base.py:5-26``; ``report`` prints the banner and the
seconds per iteration per cell (``plugins/solver/report.py:9-60``), ``nancheck`` stops a run
whose interior holds a non-finite value (``plugins/solver/nancheck.py:10-44``).
"""

import time


class BasePlugin:
    name = None
    defaults = {"every": 1}

    def __init__(self, cfgsect, solver):
        self.cfg = {**self.defaults, **(cfgsect or {})}

    def start(self, solver):
        pass

    def due(self, solver):
        return solver.nrt % int(self.cfg["every"]) == 0

    def __call__(self, solver):
        pass

    def finalize(self, solver):
        pass


class Report(BasePlugin):
    name = "report"

    def start(self, solver):
        print(repr(solver), flush=True)
        self.t0, self.n0 = time.perf_counter(), solver.nrt

    def __call__(self, solver):
        print(f"  step {solver.nrt:8d}  t = {solver.tme.get():.6e}", flush=True)

    def finalize(self, solver):
        solver.synchronize()
        steps = max(solver.nrt - self.n0, 1)
        dt = time.perf_counter() - self.t0
        print(f"  Seconds/Iteration/Cell: {dt / steps / max(solver.ncells, 1):.3e}", flush=True)


class NanCheckError(RuntimeError):
    pass


class NanCheck(BasePlugin):
    name = "nancheck"

    def __call__(self, solver):
        torch = solver.backend.torch
        ng = solver.ng
        for blk in solver.blocks:
            Q = blk.Q.t[:, ng:-ng, ng:-ng, ng:-ng]
            if not bool(torch.isfinite(Q).all()):
                raise NanCheckError(f"block {blk.nblki}: a non-finite value in Q at step {solver.nrt}")


PLUGINS = {p.name: p for p in (Report, NanCheck)}


def claim(cfgsect, solver):
    unknown = sorted(set(cfgsect) - set(PLUGINS))
    if unknown:
        raise ValueError(f"{unknown}: no plugin of that name in this build (have {sorted(PLUGINS)})")
    return [PLUGINS[name](sect, solver) for name, sect in cfgsect.items()]
