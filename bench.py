"""Benchmark: cell-updates/s per RK stage of the reacting H2/O2 case on one MI355X.

Workload: H2/O2-N2 (9 species, 28 reactions, the H/O subset of GRI-Mech 3.0), one periodic
128^3 block on [0, 1 cm]^3, 1 atm, stoichiometric H2-air at rest, T 1000 K with eight seeded
Gaussian hot spots clipped at 1500 K, MUSCL (Koren) + HLLC, coupled finite-rate chemistry,
TVD-RK3, FP64, fixed dt = 1 ns. A step is one RK3 time step (three stages); value = interior
cells x stages / wall time of the timed steps, with every array resident on the device and each
step replayed from its captured HIP graph.

Alongside: the dominant kernel's achieved rate against its roofline (device events on the
launch stream around each launch, over as many more steps launched eagerly), and the CPU oracle
timed on the host's cores on a bounded sample of the same case.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--cpu-n 128] [--cpu-seconds 10] [--no-cpu]
"""

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

HBM_PEAK = 8000.0  # GB/s, MI355X HBM3E
FP64_PEAK = 78.6  # TFLOP/s, MI355X FP64 vector
# FP64 flop per face of the MUSCL-Koren + HLLC sweep on the 9-species thermally perfect gas, from
# its gfx950 code: 247 multiplies, 209 adds and 317 fused multiply-adds (2 each; the divisions'
# refinement steps are among them), 1090 in all; min, max, compares and selects not counted
FLUX_FLOP_PER_FACE = 1090
DT = 1e-9
L = (0.01, 0.01, 0.01)


def case_config(support):
    return support.config("h2o2", recon="muscl", primary="hllc", limiter="koren", chem=True, integrator="tvd-rk3", dt=DT)


def batch_config(support):
    c = case_config(support)
    c["chem-integration"]["kind"] = "implicit"
    c.validate()
    return c


def b_alg(ne, rbar=5.0 / 3.0, wbar=1.0, tguess=2, G=10):
    """Compulsory bytes per cell per stage of a fused pass (tvd-rk3, Euler): the stage input read
    and the step's start read in two of three stages (rbar ne), the result written (wbar ne), T read
    and written as the EOS's start, and the geometry read (detJinv and one face's area vector per
    axis)."""
    return 8.0 * (rbar * ne + wbar * ne + tguess + G)


def algorithmic_bytes(name, ne, ng):
    """Compulsory HBM bytes per item of one launch of an entry point (doubles x 8)."""
    if name == "pga_stage":
        return b_alg(ne)
    if name == "pga_implicit":
        # per cell: Q and T in, the carried species' mean rates out (the matrices are the call's own work)
        return 8 * (ne + 1 + ne - 5)
    if name == "pga_rates":
        # per cell: Q and T in, the carried species' rates out
        return 8 * (ne + 1 + ne - 5)
    if name in ("pga_flux", "pga_faces"):
        # per face: Q and p, T, rhoinv of one cell (the stencil's others are neighbours' reads), S, F out
        return 8 * (ne + 3 + 3 + ne)
    if name == "pga_source":
        # per cell: Q and T in, dQ out
        return 8 * (ne + 1 + ne)
    if name == "pga_applyflux":
        # per cell: three low faces' F (the high faces are neighbours' lows), detJinv, dQ in and out
        return 8 * (3 * ne + 1 + 2 * ne)
    if name == "pga_combine":
        return 8 * 3 * ne
    if name == "pga_fromQ":
        return 8 * (ne + 1 + 8)
    return 0


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(support, n, seconds):
    """The oracle (C, OpenMP over the host's cores) on an n^3 block of the same case (the full 128^3
    by default): stages in turn (RK3's three, repeated) until `seconds` of them have run, the
    median stage reported."""
    from oracle import oracle as O
    from peregrine_amd.simulator import get_simulator

    sim = get_simulator(case_config(support))
    oc = O.OracleCase(sim.mixture, "muscl", "koren", "hllc")
    hb = support.HostBlock((n,) * 3, sim.ng, sim.ne, 3, lengths=L)
    q = support.h2_prims(hb.xcc, hb.ncell, lengths=L)
    O.lib.pgo_fromprims(oc.ref, O.arr(support.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state),
                        O.box((0, 0, 0), hb.cshape))
    b = support.oracle_block(hb, "tvd-rk3", True)
    O.lib.pgo_stage(oc.ref, ctypes.byref(b), 0, DT)  # warm-up
    times = []
    start = time.perf_counter()
    s = 1
    while time.perf_counter() - start < seconds or len(times) < 6:
        t0 = time.perf_counter()
        O.lib.pgo_stage(oc.ref, ctypes.byref(b), s % 3, DT)
        times.append(time.perf_counter() - t0)
        s += 1
    med = float(np.median(times))
    return dict(value=n**3 / med, unit="cell-updates/s", cores=O.threads(), kind="port", cpu=cpu_model(),
                sample=f"{n}^3 block of the same case, median of {len(times)} RK3 stages ({sum(times):.1f} s), "
                       "C oracle with OpenMP, one thread per core the process may run on")


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel, from the rocprofv3 FETCH_SIZE and WRITE_SIZE passes over this
    benchmark, each scaled by the same counter's reading of a plain copy of known size; None when no
    profile has been recorded."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.is_file():
        return None
    got = json.loads(path.read_text()).get(kernel)
    return None if got is None else got["bytes_per_launch"]


def batch(args, pg, support):
    """The uneven batch: args.batch periodic H2/O2 boxes with edges drawn by
    default_rng(20240611).integers(16, 97), MUSCL-Koren + HLLC, implicit chemistry (linearly implicit
    Euler, FP32 matrix-core LU with FP64 refinement), tvd-rk3. value: one solver stepping every box,
    each phase one launch over the table of blocks, each step replayed from its HIP graph; solves/s =
    boxes x steps / wall time. Beside it the same boxes stepped by a solver each, one after another
    (the plain per-case form, launched eagerly and replayed from a graph each), the table launched
    eagerly, and a launch per block (one graph, or eagerly on side streams); the largest gap of any
    box's Y and T between the batch and the box alone is printed beside them."""
    import torch

    dims = np.random.default_rng(20240611).integers(16, 97, size=(args.batch, 3))
    cfg = batch_config(support)

    def solver(dlist, graphs, batched):
        c = batch_config(support)
        c["backend-mi355x"]["graphs"] = graphs
        c["backend-mi355x"]["batched"] = batched
        mb = pg.get_solver(c, pg.mesher.BatchMesher(dlist, lengths=L), pg.get_backend("mi355x", c))
        mb.set_prims([support.h2_prims(b.xcc.get(), b.ncell, lengths=L) for b in mb.blocks])
        return mb

    def timed(solvers):
        # the warm-up's second step tunes the tiles and captures each graph
        for _ in range(max(args.warmup, 2)):
            for mb in solvers:
                mb.step(DT)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for mb in solvers:
                mb.step(DT)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def YT(mb):
        out = []
        for blk in mb.blocks:
            Q = blk.regs[0].get()[blk.interior]
            T = blk.state.get()[blk.interior][..., 1]
            out.append((Q[..., 5:] / Q[..., :1], T))
        return out

    n = len(dims)
    times, results = {}, {}
    for name, (graphs, batched) in (("table_graph", (True, True)), ("table_eager", (False, True)),
                                    ("per_block_graph", (True, False)), ("per_block_side_streams", (False, False))):
        mb = solver(dims, graphs, batched)
        assert mb.batched == batched, name
        times[name] = timed([mb])
        print(f"{name}: {args.steps} steps in {times[name]:.3f} s", file=sys.stderr, flush=True)
        if name == "table_graph":
            cells = mb.ncells
            results["batch"] = YT(mb)
        del mb
    for name, graphs in (("plain_eager", False), ("plain_graph", True)):
        each = [solver([d], graphs, False) for d in dims]
        times[name] = timed(each)
        print(f"{name}: {args.steps} steps in {times[name]:.3f} s", file=sys.stderr, flush=True)
        if name == "plain_eager":
            results["alone"] = [YT(mb)[0] for mb in each]
        del each
    gap = 0.0
    for (Yb, Tb), (Ya, Ta) in zip(results["batch"], results["alone"]):
        gap = max(gap, float(np.abs(Yb - Ya).max()), float(np.abs(Tb - Ta).max() / np.abs(Ta).max()))
    rate = {k: n * args.steps / t for k, t in times.items()}
    line = {
        "metric": "batched solves/s (H2-air 3D uneven batch)", "value": rate["table_graph"], "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": times["table_graph"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n} periodic h2o2 boxes, edges 16-96 (seed 20240611), muscl-koren + hllc, "
                               "implicit chemistry (fp32 mfma lu + fp64 refinement), tvd-rk3", "cells": int(cells)},
        "plain_per_case": rate["plain_eager"], "speedup": times["plain_eager"] / times["table_graph"],
        "speedup_vs_plain_graph": times["plain_graph"] / times["table_graph"],
        "variants": {k: round(v, 2) for k, v in rate.items()},
        "cell_updates_per_s": cells * cfg_stages(cfg) * args.steps / times["table_graph"],
        "max_gap_Y_T_batch_vs_alone": gap,
    }
    print(json.dumps(line), flush=True)


def serve(args, pg, support):
    """The serving stream: args.serve requests (80% 0D H2-air ignitions to 0.2 ms, 15% Sod tubes of
    200-4000 cells, 5% Taylor-Green boxes of 16^3-64^3, 20 steps) arriving as a Poisson stream at
    args.rate per second (default_rng(20240611)), served with continuous batching (value: solves/s
    over the run's wall time, with p50 and p99 latency), against the same requests served one at a
    time, one after another (plain per-case solves/s); the largest difference of any result between
    the two is printed beside them."""
    from peregrine_amd import serving

    backend = pg.get_backend("mi355x", support.config("air-cpg"))
    # warm-up: compile caches and first-use allocations on a few requests of each kind
    warm = serving.Server(backend, reactors=64)
    warm.run([serving.ignition(1200.0, 1e6, 1.0, tend=1e-6), serving.tube(200), serving.taylor_green(16, steps=2)])
    reqs = serving.poisson_stream(args.rate, args.serve)
    srv = serving.Server(backend, reactors=512)
    done, wall = srv.run(reqs)
    p50, p99 = serving.percentiles(done)
    alone = serving.Server(backend, reactors=1, solvers=1)
    plain = {}
    t0 = time.perf_counter()
    for r in reqs:
        one = serving.Request(r.kind, dict(r.params), rid=r.rid)
        (d,), _ = alone.run([one])
        plain[r.rid] = d
    t_plain = time.perf_counter() - t0
    gap = 0.0
    for r in done:
        a, b = r.result, plain[r.rid].result
        for k in a:
            if isinstance(a[k], (float, np.ndarray)) and k != "delay":
                x, y = np.asarray(a[k], float), np.asarray(b[k], float)
                gap = max(gap, float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)))
    kinds = {k: sum(1 for r in done if r.kind == k) for k in ("ignition", "tube", "tgv")}
    line = {
        "metric": "batched solves/s + p99 latency (serving stream)", "value": len(done) / wall, "unit": "solves/s",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{len(reqs)} requests, Poisson at {args.rate}/s (seed 20240611): {kinds}",
                   "reactor_slots": 512, "budget_bytes": srv.budget},
        "p50_s": p50, "p99_s": p99, "wall_s": wall, "plain_per_case": len(reqs) / t_plain,
        "speedup": t_plain / wall, "max_result_gap_alone_vs_mix": gap,
    }
    print(json.dumps(line), flush=True)


def tgv_case(support, n):
    """Taylor-Green at Mach 0.1 and Re 1600 on [0, 2 pi]^3: thermally perfect NASA-7 air, KEEP,
    Navier-Stokes with constant transport (Pr 0.71), RK4, dt at CFL 0.2 on the sound speed plus V0."""
    L3 = (2 * np.pi,) * 3
    R = 8314.46261815324 / support.AIR_NASA7["Air"]["MW"]
    cp = 1005.0
    c0 = np.sqrt(cp / (cp - R) * R * 300.0)
    V0 = 0.1 * c0
    rho0 = 101325.0 / (R * 300.0)
    mu = rho0 * V0 / 1600.0
    cfg = support.config("air-tpg", integrator="rk4", trans="constant")
    cfg["mixture"]["species"] = {"Air": {**support.AIR_NASA7["Air"], "mu0": mu, "kappa0": mu * cp / 0.71}}

    def prims(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        q = np.zeros(x.shape[:3] + (5,))
        q[..., 0] = 101325.0 + rho0 * V0**2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
        q[..., 1] = V0 * np.sin(X) * np.cos(Y) * np.cos(Z)
        q[..., 2] = -V0 * np.cos(X) * np.sin(Y) * np.cos(Z)
        q[..., 4] = q[..., 0] / (rho0 * R)
        return q

    return cfg, L3, prims, 0.2 * (2 * np.pi / n) / (c0 + V0)


def tgv(args, pg, support):
    """Taylor-Green 32^3 (the case the reference's CPU path runs), 100 rk4 steps: value = the steps'
    cell-updates/s per stage on the device (fused rk4 stage, each step replayed from its HIP graph);
    beside it the C oracle running the same 100 steps on the host's cores (the whole case, not a
    sample) and the largest gap between the two solutions, row-normalised."""
    import torch
    from oracle import oracle as O
    from peregrine_amd.simulator import get_simulator

    n, nsteps = 32, 100
    cfg, L3, prims, dt = tgv_case(support, n)
    mb = pg.get_solver(cfg, pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=L3, periodic=[True] * 3),
                       pg.get_backend("mi355x", cfg))
    blk = mb.blocks[0]
    mb.set_prims([prims(blk.xcc.get())])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(nsteps):
        mb.step(dt)
    torch.cuda.synchronize()
    first = time.perf_counter() - t0
    got = blk.regs[0].get()[blk.interior]
    # the steady rate: as many more steps, every one replayed from the graph
    t0 = time.perf_counter()
    for _ in range(nsteps):
        mb.step(dt)
    torch.cuda.synchronize()
    steady = time.perf_counter() - t0
    sim = get_simulator(cfg)
    oc = O.OracleCase(sim.mixture, "constant", None, "kepaec")
    hb = support.HostBlock((n,) * 3, sim.ng, sim.ne, 4, lengths=L3)
    O.lib.pgo_fromprims(oc.ref, O.arr(support.to_dev(prims(hb.xcc))), O.arr(hb.regs[0]), O.arr(hb.state),
                        O.box((0, 0, 0), hb.cshape))
    O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box((hb.ng,) * 3, hb.ncell))
    b = support.oracle_block(hb, "rk4", 0, viscous=True, alpha=getattr(sim, "alpha", 1.0))
    t0 = time.perf_counter()
    for _ in range(nsteps):
        O.lib.pgo_step(oc.ref, ctypes.byref(b), dt)
    cpu = time.perf_counter() - t0
    want = hb.regs[0][hb.interior()].transpose(3, 2, 1, 0)
    scale = np.abs(want).reshape(-1, want.shape[-1]).max(axis=0)
    gap = float((np.abs(got - want).reshape(-1, want.shape[-1]).max(axis=0) / scale).max())
    cells = n**3
    line = {
        "metric": "cell-updates/s per RK stage (FP64, Taylor-Green 32^3)", "value": cells * 4 * nsteps / steady,
        "unit": "cell-updates/s", "n_gpus": 1, "steps": nsteps, "warmup": nsteps, "ms_per_step": steady / nsteps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "taylor-green 32^3, air nasa-7 tpg, navier-stokes constant transport, keep, rk4, "
                               "100 steps", "path": "fused" if mb.fused else "per-kernel",
                   "hip_graph_replays": mb.replays},
        "first_100_steps_ms_per_step": first / nsteps * 1e3,
        "cpu_baseline": dict(value=cells * 4 * nsteps / cpu, unit="cell-updates/s", cores=O.threads(), kind="port",
                             sample="the whole case: 100 rk4 steps of 32^3, C oracle with OpenMP",
                             ms_per_step=cpu / nsteps * 1e3),
        "max_row_gap_vs_oracle_100_steps": gap,
    }
    print(json.dumps(line), flush=True)


def layer_case(support):
    """The transcritical mixing layer: O2 at 120 K under N2 at 300 K across a tanh layer in y, 60 bar,
    Peng-Robinson with Chung transport, MUSCL-Koren + Rusanov, Navier-Stokes, tvd-rk3; 10 um cells,
    periodic in x and z, adiabatic slip walls at both y ends; a shear and seeded perturbations."""
    cfg = support.config("o2n2-pr", recon="muscl", primary="rusanov", limiter="koren", trans="chung", dt=2e-9)
    for f in (2, 3):
        cfg["bcvals"][f"wall{f}"] = {"kind": "slp-adia-wall"}

    def prims(xcc, ne, seed=7):
        rng = np.random.default_rng(seed)
        y = xcc[..., 1]
        Ly = xcc[..., 1].max() + xcc[..., 1].min()
        s = np.tanh((y - 0.5 * Ly) / (0.1 * Ly))
        q = np.zeros(xcc.shape[:3] + (ne,))
        q[..., 0] = 6e6
        q[..., 1] = 10.0 * s + 0.5 * np.sin(2 * np.pi * xcc[..., 0] / Ly + rng.uniform(0, 6.3))
        q[..., 2] = 0.5 * np.sin(2 * np.pi * xcc[..., 2] / Ly + rng.uniform(0, 6.3))
        q[..., 4] = 210.0 - 90.0 * s
        q[..., 5] = 0.5 * (1.0 - s) * 0.999 + 0.0005
        return q

    return cfg, prims


def layer(args, pg, support):
    """The transcritical mixing layer at 256 x 128 x 64: cell-updates/s per stage over the timed
    steps (each replayed from its HIP graph), and the C oracle on the host's cores on a bounded
    64 x 32 x 16 sample of the same case."""
    import torch
    from oracle import oracle as O
    from peregrine_amd.simulator import get_simulator

    ncell = (256, 128, 64)
    lengths = tuple(1e-5 * n for n in ncell)
    cfg, prims = layer_case(support)
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1 for n in ncell], lengths=lengths, periodic=[True, False, True],
                                bcnames={3: "wall2", 4: "wall3"})
    mb = pg.get_solver(cfg, mesh, pg.get_backend("mi355x", cfg))
    blk = mb.blocks[0]
    mb.set_prims([prims(blk.xcc.get(), mb.sim.ne)])
    for _ in range(max(args.warmup, 2)):
        mb.step(2e-9)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mb.step(2e-9)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cells = int(np.prod(ncell))
    cpu = None
    if not args.no_cpu:
        small = (64, 32, 16)
        sim = get_simulator(cfg)
        oc = O.OracleCase(sim.mixture, "muscl", "koren", "rusanov")
        hb = support.HostBlock(small, sim.ng, sim.ne, 3, lengths=tuple(1e-5 * n for n in small), nstate=9)
        q = prims(hb.xcc, sim.ne)
        O.lib.pgo_fromprims(oc.ref, O.arr(support.to_dev(q)), O.arr(hb.regs[0]), O.arr(hb.state),
                            O.box((0, 0, 0), hb.cshape))
        O.lib.pgo_fromQ(oc.ref, O.arr(hb.regs[0]), O.arr(hb.state), O.box((hb.ng,) * 3, hb.ncell))
        b = support.oracle_block(hb, "tvd-rk3", 0, viscous=True, alpha=getattr(sim, "alpha", 1.0), walls=(2, 3))
        times, start = [], time.perf_counter()
        while time.perf_counter() - start < args.cpu_seconds or len(times) < 3:
            t1 = time.perf_counter()
            O.lib.pgo_step(oc.ref, ctypes.byref(b), 2e-9)
            times.append(time.perf_counter() - t1)
        med = float(np.median(times))
        cpu = dict(value=int(np.prod(small)) * 3 / med, unit="cell-updates/s", cores=O.threads(), kind="port",
                   cpu=cpu_model(), sample=f"64 x 32 x 16 block of the same case, median of {len(times)} tvd-rk3 steps, "
                                           "C oracle with OpenMP")
    line = {
        "metric": "cell-updates/s per RK stage (FP64, transcritical O2/N2 layer)",
        "value": cells * 3 * args.steps / elapsed, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "o2/n2 mixing layer 256x128x64, 60 bar, peng-robinson, chung, muscl-koren + rusanov, "
                               "navier-stokes, slip walls in y, tvd-rk3", "cells": cells,
                   "path": "fused" if mb.fused else "per-kernel", "hip_graph_replays": mb.replays},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cfg_stages(cfg):
    return {"rk1": 1, "rk2": 2, "tvd-rk3": 3, "rk4": 4, "rk34": 4}[cfg["time-integration"]["integrator"]]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--per-kernel", action="store_true", help="the reference's launch sequence instead of the fused stage")
    ap.add_argument("--k-chunk", type=int, default=None)
    ap.add_argument("--stage-slabs", type=int, default=None, help="k-slabs of the fused stage")
    ap.add_argument("--k-march", type=int, default=None,
                    help="k-planes a thread of the march walks (the k sweep inside the stage's launch; 0: split)")
    ap.add_argument("--no-graphs", action="store_true", help="launch every step eagerly (no HIP graph replay)")
    ap.add_argument("--implicit", action="store_true",
                    help="implicit chemistry (FP32 matrix-core LU, FP64 refinement) in place of the coupled source")
    ap.add_argument("--batch", type=int, default=0,
                    help="the uneven batch instead: this many periodic H2/O2 cases of 16-96 cells per axis "
                         "(seeded) stepped together, against the same cases stepped one by one")
    ap.add_argument("--serve", type=int, default=0,
                    help="the serving stream instead: this many requests arriving as a Poisson stream")
    ap.add_argument("--layer", action="store_true",
                    help="the transcritical O2/N2 mixing layer at 256x128x64 instead")
    ap.add_argument("--tgv", action="store_true",
                    help="the Taylor-Green 32^3 rk4 case instead, 100 steps, against the oracle's same steps on the host")
    ap.add_argument("--rate", type=float, default=400.0, help="the serving stream's arrivals per second")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only a plain device copy of 512 MiB, five times: the reference for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU: this build has no multi-GPU path (the block sharding is design only)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy:
        n = 512 * 2**20 // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 5}), flush=True)
        return

    import peregrine_amd as pg
    import support

    if args.batch:
        return batch(args, pg, support)
    if args.serve:
        return serve(args, pg, support)
    if args.tgv:
        return tgv(args, pg, support)
    if args.layer:
        return layer(args, pg, support)
    cfg = batch_config(support) if args.implicit else case_config(support)
    cfg["backend-mi355x"]["fused-stage"] = not args.per_kernel
    if args.k_chunk:
        cfg["backend-mi355x"]["k-chunk"] = args.k_chunk
    if args.stage_slabs is not None:
        cfg["backend-mi355x"]["stage-slabs"] = args.stage_slabs
    if args.k_march is not None:
        cfg["backend-mi355x"]["k-march"] = args.k_march
    cfg["backend-mi355x"]["graphs"] = not args.no_graphs
    backend = pg.get_backend("mi355x", cfg)
    n = args.n
    mesh = pg.mesher.CubeMesher(blkdims=[n + 1] * 3, lengths=L, periodic=[True] * 3)
    mb = pg.get_solver(cfg, mesh, backend)
    blk = mb.blocks[0]
    mb.set_prims([support.h2_prims(blk.xcc.get(), blk.ncell, lengths=L)])
    # the warm-up's second step captures the step's HIP graph; the timed steps replay it
    for _ in range(max(args.warmup, 2)):
        mb.step(DT)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mb.step(DT)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    nst = mb.integ.nstages
    ncells = mb.ncells
    value = ncells * nst * args.steps / elapsed
    # the kernels' own times from device events around each launch, over as many more steps (a
    # timed step launches eagerly, each launch between its two events)
    mb.time_kernels(True)
    for _ in range(args.steps):
        mb.step(DT)
    times = mb.kernel_times()
    # the dominant entry point by total time over the timed steps
    total = {k: c * ms for k, (c, ms) in times.items()}
    dom = max(total, key=total.get)
    count, ms = times[dom]
    ne = mb.sim.ne
    # items per launch: a stage's faces (three axes) or cells over the launches a stage makes
    per_stage = {"pga_flux": 3 * (n + 1) * n * n, "pga_faces": 3 * (n + 1) * n * n}.get(dom, n**3)
    items = per_stage * nst * args.steps / count
    per = algorithmic_bytes(dom, ne, mb.ng)
    achieved = per * items / (ms * 1e-3) / 1e9
    kname = {"pga_flux": "k_flux", "pga_faces": "k_faces", "pga_stage": "k_stage"}.get(dom, dom)
    # the flux sweep is bound by FP64 VALU issue and latency (3 waves per SIMD), not by HBM: its
    # bytes against HBM and its flops against FP64 are both reported, and the bound names the one
    # whose peak it is nearer
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK, unit="GB/s", frac=round(achieved / HBM_PEAK, 4),
                traffic=measured_traffic(kname), traffic_source="profiles/traffic.json: rocprofv3 FETCH_SIZE and "
                "WRITE_SIZE passes of this benchmark, scaled by a plain copy's readings",
                kernel=dom, kernel_ms=round(ms, 4), launches=count, bytes_per_item=per, items_per_launch=items)
    if dom in ("pga_flux", "pga_faces"):
        tf = FLUX_FLOP_PER_FACE * items / (ms * 1e-3) / 1e12
        roof["fp64"] = dict(achieved=round(tf, 2), peak=FP64_PEAK, unit="TFLOP/s", frac=round(tf / FP64_PEAK, 4),
                            flop_per_item=FLUX_FLOP_PER_FACE)
        roof["limiter"] = "FP64 VALU issue and latency (3 waves per SIMD at 152 VGPRs); neither peak is reached"
    # the whole stage against the compulsory bytes of a fused pass (b_alg per cell and stage)
    if not mb.sim.viscous:
        ba = b_alg(ne)
        st = ba * ncells * nst * args.steps / elapsed / 1e9
        roof["stage"] = dict(achieved=round(st, 1), peak=HBM_PEAK, unit="GB/s", frac=round(st / HBM_PEAK, 4),
                             bytes_per_cell=ba, stage_ms=round(elapsed / args.steps / nst * 1e3, 4))
    breakdown = {k: {"launches": c, "mean_ms": round(t, 4), "share": round(c * t / sum(total.values()), 3)}
                 for k, (c, t) in sorted(times.items(), key=lambda kv: -kv[1][0] * kv[1][1])}
    cpu = None if args.no_cpu else cpu_baseline(support, args.cpu_n, args.cpu_seconds)
    line = {
        "metric": "cell-updates/s per RK stage (FP64, H2-air 3D)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"h2o2-9sp-28rx {n}^3 periodic, muscl-koren + hllc, "
                               f"{'implicit' if args.implicit else 'coupled'} chemistry, tvd-rk3",
                   "cells": ncells, "stages_per_step": nst, "path": "fused" if mb.fused else "per-kernel",
                   "hip_graph_replays": mb.replays,
                   "k_chunk": cfg["backend-mi355x"]["k-chunk"],
                   "stage_slabs": cfg["backend-mi355x"]["stage-slabs"],
                   "k_march": cfg["backend-mi355x"]["k-march"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": breakdown,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
