"""Runge-Kutta schemes: the registers each stage reads and writes and its combine weights.

Restates ``peregrine/integrators/rungekutta.py:4-95`` and
``peregrine/integrators/base.py:27-42``. A combine is
``out = [wout out] + sum_m w_m X_m``, dt multiplying the last term when the
combine is a step; the weights of a register named twice are summed and the
output register's own weight becomes ``wout``, in the order the reference
collects them (Q, then the start register, then dQ).
"""

from collections import namedtuple

from peregrine_amd.misc import find_subclass

StageRegisters = namedtuple("StageRegisters", "Q dQ Qout")
Combine = namedtuple("Combine", "out own wout X w step")


def combine(out, *pairs, step=True):
    weights = {}
    for reg, w in pairs:
        weights[reg] = weights.get(reg, 0.0) + w
    own = weights.pop(out, 0.0)
    reads = [r for r in weights if weights[r] != 0.0]
    return Combine(out, own != 0.0, own, tuple(reads), tuple(weights[r] for r in reads), step)


class FixedController:
    """``integrators/controllers.py:13-17``: the configured dt."""

    name = "fixed"

    def __init__(self, cfg):
        self.dt = cfg["simulation"]["dt"]

    def stepsize(self, solver):
        return self.dt


class CFLController:
    """``integrators/controllers.py:19-29``: cfl over the largest combined inverse time scale, no
    more than dt-max, shortened to land on the next time a plugin acts at."""

    name = "cfl"

    def __init__(self, cfg):
        self.cfl = float(cfg["simulation"]["cfl"])
        self.dtmax = float(cfg["simulation"]["dt-max"])

    def stepsize(self, solver):
        dt = min(self.cfl / solver.cflmax()["combined"], self.dtmax)
        times = [p.tnext(solver) for p in solver.plugins]
        now = solver.tme.get()
        return min([dt, *(t - now for t in times if t is not None)])


CONTROLLERS = {c.name: c for c in (FixedController, CFLController)}


class BaseRungeKutta:
    name = None
    nregs = 3
    # (fraction of dt, wQ0, wQ, wdQ): Qout = wQ Q + wQ0 Qstart + wdQ dt dQ
    stages = ()

    def __init__(self, controller):
        self.controller = controller

    @property
    def nstages(self):
        return len(self.stages)

    def stageregs(self, regs):
        start, stage, dQ = regs[0], regs[1], regs[-1]
        last = len(self.stages) - 1
        return [
            StageRegisters(Q=start if n == 0 else stage, dQ=dQ, Qout=start if n == last else stage)
            for n in range(len(self.stages))
        ]

    def stagekerns(self, n, regs):
        _, wQ0, wQ, wdQ = self.stages[n]
        Q, dQ, Qout = self.stageregs(regs)[n]
        return [combine(Qout, (Q, wQ), (regs[0], wQ0), (dQ, wdQ))]

    def stagetimes(self, start, dt):
        return [start + f * dt for f, *_ in self.stages[1:]] + [start + dt]


class RK1(BaseRungeKutta):
    name = "rk1"
    nregs = 2
    stages = ((0.0, 0.0, 1.0, 1.0),)


class RK2(BaseRungeKutta):
    name = "rk2"
    stages = ((0.0, 0.0, 1.0, 1.0), (1.0, 0.5, 0.5, 0.5))


# Gottlieb and Shu, Total variation diminishing Runge-Kutta schemes, Math. Comp. 67(221), 1998
class RK3(BaseRungeKutta):
    name = "tvd-rk3"
    stages = ((0.0, 0.0, 1.0, 1.0), (1.0, 0.75, 0.25, 0.25), (0.5, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0))


class RK34(BaseRungeKutta):
    name = "rk34"
    stages = ((0.0, 0.0, 1.0, 0.5), (0.5, 0.0, 1.0, 0.5), (1.0, 2.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0), (0.5, 0.0, 1.0, 0.5))


class RK4(BaseRungeKutta):
    name = "rk4"
    nregs = 4
    # (fraction of dt, wdQ, wsum); the last stage has no wdQ
    stages = ((0.0, 0.5, 1.0 / 6.0), (0.5, 0.5, 1.0 / 3.0), (0.5, 1.0, 1.0 / 3.0), (1.0, None, 1.0 / 6.0))

    def stagekerns(self, n, regs):
        start, stage, total, dQ = regs
        _, wdQ, wsum = self.stages[n]
        share = combine(total, (total, 0.0 if n == 0 else 1.0), (dQ, wsum))
        if wdQ is None:
            return [share, combine(start, (start, 1.0), (total, 1.0), step=False)]
        return [share, combine(stage, (start, 1.0), (dQ, wdQ))]


def get_integrator(cfg):
    name = cfg["time-integration"]["integrator"]
    cls = find_subclass(BaseRungeKutta, name=name)
    name = cfg["simulation"]["controller"]
    if name not in CONTROLLERS:
        raise ValueError(f"controller {name!r}: this build has {sorted(CONTROLLERS)}")
    return cls(CONTROLLERS[name](cfg))
