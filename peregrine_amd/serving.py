"""Serving: a request loop over small solves whose state stays on the device.

Requests of three kinds arrive over time (a seeded Poisson stream in the benchmark):

- ``Ignition``: a 0D H2-air reactor at constant volume (the reference's own ignition case,
  tests/chemint/cases.py:45-58, examples/autoIgnition.py:18-53), T0, p0 and the equivalence
  ratio phi given, integrated to ``tend`` by the implicit chemistry (linearly implicit Euler, the
  per-cell systems factored in FP32 on the matrix cores and refined in FP64). Its result is the
  final mass fractions and T and the time T took to rise by 400 K. Constant pressure is not built:
  the reference has no constant-pressure reactor either.
- ``Tube``: Sod's shock tube on a walled line of n cells (the reference's walled lines,
  tests/simulator/support.py:144-195, tests/riemann.py:7-14), MUSCL-Koren + HLLC, tvd-rk3, to a
  fixed time; its result is the density along the line.
- ``TaylorGreen``: a periodic n^3 Taylor-Green box, Navier-Stokes with KEEP and rk4, a fixed
  number of steps; its result is the box's mean kinetic energy.

The loop keeps every admitted request resident: the reactors as the cells of one device row (a
``ReactorPool``), each tube or box as a solver of its own. A request is admitted when it has arrived,
a slot is free and its HBM footprint fits under the budget left by those resident; every tick
advances every resident request by one step (all the reactors in one launch of the implicit
chemistry, each solver on a stream of its own, none waiting on the host), and a request that has
taken its steps is read back and retired, freeing its slot and memory. Latency is finish minus
arrival (queueing included); throughput is finished requests over the run's wall time.

A reactor's arithmetic is the same whichever slot it takes and whoever shares the launch, and a
solver steps only its own arrays, so each result is the same alone and in the mix (tested to the
bit).
"""

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

import peregrine_amd as pg
from peregrine_amd import native
from peregrine_amd.config import Config
from peregrine_amd.simulator import NativeCase, get_simulator

# Sod's tube, Toro's test 1 (tests/riemann.py:7-8) in air at SI values: left and right states
SOD = {"rhoL": 1.0, "uL": 0.0, "pL": 1.0e5, "rhoR": 0.125, "uR": 0.0, "pR": 1.0e4}
AIR = {"Air": {"MW": 28.96, "cp0": 1005.0, "mu0": 1.8591191080521142e-05, "kappa0": 0.02625394405190068}}
R_AIR = 8314.46261815324 / 28.96
RISE = 400.0  # K: the rise that marks ignition


@dataclass
class Request:
    kind: str
    params: dict
    arrival: float = 0.0
    rid: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    result: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def ignition(T0, p0, phi, tend=2e-4, dt=1e-7, **kw):
    return Request("ignition", dict(T0=T0, p0=p0, phi=phi, tend=tend, dt=dt), **kw)


def tube(n, **kw):
    return Request("tube", dict(n=int(n)), **kw)


def taylor_green(n, steps=20, **kw):
    return Request("tgv", dict(n=int(n), steps=int(steps)), **kw)


def h2_air(phi, names):
    """Mass fractions of H2-air at equivalence ratio phi (air 0.21 O2 + 0.79 N2 by mole)."""
    X = {"H2": 2.0 * phi, "O2": 1.0, "N2": 3.76}
    MW = {"H2": 2.01588, "O2": 31.9988, "N2": 28.0134}
    m = {k: X[k] * MW[k] for k in X}
    total = sum(m.values())
    return np.array([m.get(n, 0.0) / total for n in names])


# -- the reactors -------------------------------------------------------------------------------
class ReactorPool:
    """0D constant-volume reactors as the cells of one device row: Q, state and the implicit form's
    work for `capacity` slots. A tick is one step of every slot: the implicit chemistry's mean rates
    over dt (pga_implicit), rho Y advanced by dt times them (pga_combine), the EOS of the result
    (pga_fromQ), and T recorded. Free slots hold pure N2 at 300 K, whose rates are exactly zero."""

    def __init__(self, backend, capacity, dt, parts=1, history=4096):
        torch = backend.torch
        cfg = Config()
        cfg["simulation"]["simulator"] = "euler"
        m = cfg["mixture"]
        m["species"], m["eos"], m["T-range"], m["chemistry"] = "h2o2.yaml", "tpg", (300.0, 3500.0), "finite-rate"
        ci = cfg["chem-integration"]
        ci["kind"], ci["implicit-parts"] = "implicit", int(parts)
        cfg.validate()
        self.backend, self.cfg = backend, cfg
        self.sim = get_simulator(cfg)
        self.case = NativeCase(self.sim, backend)
        self.capacity, self.history = int(capacity), int(history)
        ne, ns = self.sim.ne, self.sim.ns
        self.Q = backend.alloc((capacity, 1, 1, ne), name="reactor-Q")
        self.state = backend.alloc((capacity, 1, 1, 8), name="reactor-state")
        self.rates = backend.alloc((capacity, 1, 1, ns - 1), name="reactor-rates")
        self.prims = backend.alloc((capacity, 1, 1, ne), name="reactor-prims")
        self.box = native.box((0, 0, 0), (capacity, 1, 1))
        n = int(self.case.L.pga_implicit_worksize(self.case.handle, self.box))
        self.work = torch.zeros(max(n, 1), dtype=torch.float64, device=backend.device)
        self.sweeps = torch.zeros(capacity, dtype=torch.int32, device=backend.device)
        self.dt = backend.scalar("dt", dt)
        self.dtval = float(dt)
        self.opts = native.ImplicitOpts(int(parts), 8, 1e-13)
        self.T = torch.zeros((self.history, capacity), dtype=torch.float64, device=backend.device)
        self.ticks = 0
        self.free = list(range(capacity - 1, -1, -1))
        self.slot = {}
        # every slot pure N2 at 300 K, in one launch
        inert = np.zeros(ne)
        inert[0], inert[4] = 101325.0, 300.0
        self.prims.t.copy_(torch.from_numpy(inert).to(self.prims.t.device)[:, None, None, None].expand_as(self.prims.t))
        native.check(self.case.L.pga_fromprims(self.case.handle, self.prims.view(), self.Q.view(), self.state.view(),
                                                self.box, backend.stream), "pga_fromprims", self.case.L)

    @staticmethod
    def footprint(capacity, ns=9):
        """Device bytes of a pool of `capacity` slots (Q, prims, state, rates, T history, the
        implicit form's work), per slot."""
        ne, nv = ns + 4, ns + 1
        work = nv * nv * 8 + 3 * nv * 8
        return 8 * (2 * ne + 8 + ns - 1 + 4096) + work + 4

    def _set(self, s, T0, p0, Y):
        q = np.zeros(self.sim.ne)
        q[0], q[4], q[5:] = p0, T0, Y[:-1]
        self.prims.t[:, 0, 0, s] = self.backend.torch.from_numpy(q).to(self.prims.t.device)
        b = native.box((s, 0, 0), (1, 1, 1))
        native.check(self.case.L.pga_fromprims(self.case.handle, self.prims.view(), self.Q.view(), self.state.view(), b,
                                                self.backend.stream), "pga_fromprims", self.case.L)

    def admit(self, req):
        s = self.free.pop()
        p = req.params
        self._set(s, p["T0"], p["p0"], h2_air(p["phi"], self.sim.mixture.spnames))
        req.slot, req.first, req.steps = s, self.ticks, int(round(p["tend"] / self.dtval))
        req.done = 0
        self.slot[s] = req

    def tick(self):
        L, stream = self.case.L, self.backend.stream
        ns = self.sim.ns
        native.check(L.pga_implicit(self.case.handle, ctypes.byref(self.opts), self.Q.view(), self.state.view(),
                                    self.rates.view(), 0, self.sweeps.data_ptr(), self.capacity, self.work.data_ptr(),
                                    self.work.numel(), self.dt.ptr, self.box, stream), "pga_implicit", self.case.L)
        qv = self.Q.view()
        species = native.view(self.Q.ptr + 8 * 5 * self.Q.strides[2], self.capacity, 1, 1, ns - 1, strides=self.Q.strides)
        X = (native.View * 1)(self.rates.view())
        native.check(L.pga_combine(ns - 1, species, 1, 1.0, 1, X, native.dbls([1.0]), self.dt.ptr, 1, self.box, stream),
                     "pga_combine", self.case.L)
        native.check(L.pga_fromQ(self.case.handle, qv, self.state.view(), (native.Box * 1)(self.box), 1, stream),
                     "pga_fromQ", self.case.L)
        self.T[self.ticks % self.history].copy_(self.state.t[1, 0, 0, :])
        self.ticks += 1
        for req in self.slot.values():
            req.done += 1

    def finished(self):
        return [r for r in self.slot.values() if r.done >= r.steps]

    def retire(self, req):
        """The request's result (final Y and T, the time to a 400 K rise) read back, its slot freed."""
        s = req.slot
        Q = self.Q.t[:, 0, 0, s].cpu().numpy()
        T = self.state.t[1, 0, 0, s].item()
        rows = [(req.first + n) % self.history for n in range(req.done)]
        hist = self.T[rows, s].cpu().numpy()
        T0 = req.params["T0"]
        delay = math.nan
        above = np.nonzero(hist > T0 + RISE)[0]
        if above.size:
            n = int(above[0])
            before = hist[n - 1] if n > 0 else T0
            frac = (T0 + RISE - before) / (hist[n] - before)
            delay = (n + frac) * self.dtval
        Y = Q[5:] / Q[0]
        req.result = {"Y": np.append(Y, 1.0 - Y.sum()), "T": T, "delay": delay, "sweeps": int(self.sweeps[s].item())}
        del self.slot[s]
        self._set(s, 300.0, 101325.0, np.eye(self.sim.ns)[-1])
        self.free.append(s)


# -- the solver requests --------------------------------------------------------------------------
def tube_solver(n, backend_cfg=None):
    c = Config()
    c["simulation"]["simulator"] = "euler"
    c["mixture"]["species"], c["mixture"]["eos"] = AIR, "cpg"
    a = c["adv-flux"]
    a["primary"], a["recon"], a["limiter"] = "hllc", "muscl", "koren"
    c["bcvals"]["walls"] = {"kind": "slp-adia-wall"}
    for k, v in (backend_cfg or {}).items():
        c["backend-mi355x"][k] = v
    c.validate()
    L = 1.0
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1, 2, 2], lengths=[L, 0.1, 0.1], periodic=[False] * 3,
                                bcnames=dict.fromkeys(range(1, 7), "walls"))
    mb = pg.get_solver(c, mesh, pg.get_backend("mi355x", c))
    blk = mb.blocks[0]
    x = blk.xcc.get()[..., 0]
    left = x <= 0.5 * L
    q = np.zeros(blk.cellshape + (mb.sim.ne,))
    q[..., 0] = np.where(left, SOD["pL"], SOD["pR"])
    q[..., 1] = np.where(left, SOD["uL"], SOD["uR"])
    q[..., 4] = np.where(left, SOD["pL"] / (SOD["rhoL"] * R_AIR), SOD["pR"] / (SOD["rhoR"] * R_AIR))
    mb.set_prims([q])
    # to t = 0.2 L / c_left at CFL 0.4 on the left state's sound speed and the fan's speed
    cL = math.sqrt(1.4 * SOD["pL"] / SOD["rhoL"])
    dt = 0.4 * (L / n) / (2.0 * cL)
    steps = int(math.ceil(0.2 * L / cL / dt))
    return mb, dt, steps


def tgv_solver(n, backend_cfg=None):
    c = Config()
    c["simulation"]["simulator"] = "navier-stokes"
    m = c["mixture"]
    m["species"], m["eos"], m["trans"] = AIR, "cpg", "constant"
    a = c["adv-flux"]
    a["primary"], a["recon"] = "kepaec", "constant"
    c["time-integration"]["integrator"] = "rk4"
    for k, v in (backend_cfg or {}).items():
        c["backend-mi355x"][k] = v
    c.validate()
    L = 2.0 * math.pi
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=[L] * 3, periodic=[True] * 3)
    mb = pg.get_solver(c, mesh, pg.get_backend("mi355x", c))
    blk = mb.blocks[0]
    x = blk.xcc.get()
    T0, p0 = 300.0, 101325.0
    rho0 = p0 / (R_AIR * T0)
    V0 = 0.1 * math.sqrt(1.4 * R_AIR * T0)
    q = np.zeros(blk.cellshape + (mb.sim.ne,))
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    q[..., 1] = V0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    q[..., 2] = -V0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    q[..., 0] = p0 + rho0 * V0 ** 2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    q[..., 4] = q[..., 0] / (rho0 * R_AIR)
    mb.set_prims([q])
    dt = 0.2 * (L / n) / (math.sqrt(1.4 * R_AIR * T0) + V0)
    return mb, dt


def solver_footprint(kind, n):
    """Device bytes of a tube's or a box's arrays before it is built: the registers, state, face
    arrays and geometry over the cells with their halos (two for MUSCL, one for KEEP), and the
    Navier-Stokes box's transport, gradients, centres and face centres."""
    if kind == "tube":
        cells, ne, regs, extra = (n + 4) * 5 * 5, 5, 3, 0
    else:
        cells, ne, regs, extra = (n + 2) ** 3, 5, 4, 3 + 1 + 12 + 9 + 3 + 9
    per = regs * ne + 3 * ne + 8 + 9 + 1 + 3 + extra
    return 8 * cells * per


class SolverSlot:
    """A tube or a box resident on the device as a solver of its own, stepping on a stream of its own."""

    def __init__(self, req, stream, backend_cfg=None):
        self.req, self.stream = req, stream
        p = req.params
        if req.kind == "tube":
            self.mb, self.dt, req.steps = tube_solver(p["n"], backend_cfg)
        else:
            self.mb, self.dt = tgv_solver(p["n"], backend_cfg)
            req.steps = p["steps"]
        req.done = 0

    def tick(self):
        torch = self.mb.backend.torch
        with torch.cuda.stream(self.stream):
            self.mb.step(self.dt)
        self.req.done += 1

    def retire(self):
        torch = self.mb.backend.torch
        self.stream.synchronize()
        with torch.cuda.stream(self.stream):
            g = self.mb.ng
            Q = self.mb.blocks[0].Q.get()
        if self.req.kind == "tube":
            self.req.result = {"rho": Q[g:-g, g, g, 0].copy()}
        else:
            inner = Q[g:-g, g:-g, g:-g]
            ke = 0.5 * (inner[..., 1] ** 2 + inner[..., 2] ** 2 + inner[..., 3] ** 2) / inner[..., 0]
            self.req.result = {"ke": float(ke.mean()), "Q": inner.copy()}
        self.mb.release()
        self.mb = None


# -- the loop -------------------------------------------------------------------------------------
class Server:
    """Admits arrived requests while slots and the HBM budget allow, advances every resident request
    one step per tick, and retires the finished ones; `run` serves a list of requests whose arrival
    times are seconds from its start."""

    def __init__(self, backend, budget=8 << 30, reactors=1024, reactor_dt=1e-7, solvers=16, streams=4,
                 backend_cfg=None):
        torch = backend.torch
        self.backend = backend
        self.budget, self.used = int(budget), 0
        self.pool = ReactorPool(backend, reactors, reactor_dt)
        self.pool_bytes = ReactorPool.footprint(reactors) * reactors
        self.used += self.pool_bytes
        self.maxsolvers = int(solvers)
        self.streams = [torch.cuda.Stream(device=backend.device) for _ in range(int(streams))]
        self.backend_cfg = backend_cfg
        self.solvers = []
        self.nstream = 0
        self.ticks = 0

    def footprint(self, req):
        if req.kind == "ignition":
            return 0  # a slot of the pool, whose memory is held from the start
        return solver_footprint(req.kind, req.params["n"])

    def _admit(self, req, now):
        if req.kind == "ignition":
            if not self.pool.free:
                return False
            self.pool.admit(req)
        else:
            need = self.footprint(req)
            if len(self.solvers) >= self.maxsolvers or self.used + need > self.budget:
                return False
            stream = self.streams[self.nstream % len(self.streams)]
            self.nstream += 1
            self.solvers.append(SolverSlot(req, stream, self.backend_cfg))
            req.bytes = need
            self.used += need
        req.admitted = now
        return True

    def run(self, requests, clock=time.perf_counter):
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting, done = [], []
        t0 = clock()
        nxt = 0
        while nxt < len(pending) or waiting or self.pool.slot or self.solvers:
            now = clock() - t0
            while nxt < len(pending) and pending[nxt].arrival <= now:
                waiting.append(pending[nxt])
                nxt += 1
            # first come, first served: a request that does not fit holds the ones behind it of its kind
            held = set()
            for req in list(waiting):
                kind = "ignition" if req.kind == "ignition" else "solver"
                if kind in held:
                    continue
                if self._admit(req, now):
                    waiting.remove(req)
                else:
                    held.add(kind)
            if not self.pool.slot and not self.solvers:
                if nxt < len(pending):
                    time.sleep(max(0.0, min(pending[nxt].arrival - (clock() - t0), 0.01)))
                continue
            if self.pool.slot:
                self.pool.tick()
            for s in self.solvers:
                s.tick()
            self.ticks += 1
            for req in self.pool.finished():
                self.pool.retire(req)
                req.finished = clock() - t0
                done.append(req)
            for s in [s for s in self.solvers if s.req.done >= s.req.steps]:
                s.retire()
                s.req.finished = clock() - t0
                self.used -= s.req.bytes
                self.solvers.remove(s)
                done.append(s.req)
        return done, clock() - t0


def poisson_stream(rate, count, seed=20240611, mix=(0.8, 0.15, 0.05), reactor_tend=2e-4):
    """`count` requests arriving at `rate` per second (exponential gaps, default_rng(seed)): 0D
    ignitions (T0 900-1400 K, p 1-40 atm log-uniform, phi 0.5-2), Sod tubes of 200-4000 cells and
    Taylor-Green boxes of 16^3-64^3, in the proportions `mix`."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        u = rng.random()
        if u < mix[0]:
            r = ignition(rng.uniform(900.0, 1400.0), 101325.0 * math.exp(rng.uniform(0.0, math.log(40.0))),
                         rng.uniform(0.5, 2.0), tend=reactor_tend)
        elif u < mix[0] + mix[1]:
            r = tube(rng.integers(200, 4001))
        else:
            r = taylor_green(rng.integers(16, 65))
        r.arrival, r.rid = t, rid
        out.append(r)
    return out


def percentiles(done):
    lat = np.array([r.latency for r in done])
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
