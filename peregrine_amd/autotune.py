"""The tile shapes of the launches over a table of blocks, chosen by timing them on the case itself.

A batch's blocks fall in size classes by their interior's extent along i (the contiguous index a
tile's lanes run along). Per class, the autotuner times the stage's divergence-and-combine launch
with each candidate column tile (width along i and k-planes per thread) and the k-sweep with each
candidate tile width, holding the other classes at their current choice, and keeps the fastest. A
block-at-a-time launch takes 64-wide tiles: on a 16-cell-wide block three quarters of their lanes
would idle. The choice changes only the speed: a cell's arithmetic is the same in every tile.

Choices persist in a JSON file (backend-mi355x.autotune-cache, or
~/.cache/peregrine_amd/tiles.json) keyed by the device, the case's kernel variant and the batch's
class census, so a case run again starts with its tiles.

The reference's counterpart is a fixed launch policy (peregrine/backend/base/kernels/policy.mako:8-13,
docs/config.md:190-197 for its hand-set tile options); it has no tuner.
"""

import ctypes
import json
import os
from pathlib import Path

from peregrine_amd import native

# size classes by interior cells along i: (name, upper bound inclusive)
CLASSES = (("narrow", 24), ("small", 48), ("medium", 80), ("wide", 1 << 30))
STAGE_TILES = ((64, 1), (32, 1), (16, 1), (64, 2), (32, 2), (16, 2))
FLUXK_TILES = (64, 32, 16)
DEFAULT = {"stage": (64, 1), "fluxk": 64}


def size_class(ni):
    for name, top in CLASSES:
        if ni <= top:
            return name
    return CLASSES[-1][0]


def cache_path(cfg):
    p = cfg["backend-mi355x"]["autotune-cache"]
    return Path(p) if p else Path(os.path.expanduser("~")) / ".cache" / "peregrine_amd" / "tiles.json"


def _key(solver, census):
    torch = solver.backend.torch
    d = solver.case.desc
    dev = torch.cuda.get_device_name(solver.backend.device)
    variant = f"ns{d.ns}-eos{d.eos}-recon{d.recon}-lim{d.limiter}-flux{d.primary}-chem{int(d.mechanism >= 0)}"
    classes = ",".join(f"{k}:{census[k]}" for k in sorted(census))
    return f"{dev}|{variant}|{classes}"


def _load(path):
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _time(solver, launch, reps=5):
    """Mean ms of launch() over reps, after one untimed run, by events on the launch stream."""
    torch = solver.backend.torch
    launch()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        launch()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(solver):
    """{class: {"stage": (wi, kc), "fluxk": ki}} for the solver's blocks: from the cache, or timed now.
    Timing runs the stage into the unused derivative register with a scratch copy of each block's
    state, so the solution and its state are untouched."""
    census = {}
    for blk in solver.blocks:
        c = size_class(blk.ncell[0])
        census[c] = census.get(c, 0) + 1
    path = cache_path(solver.cfg)
    cache = _load(path)
    key = _key(solver, census)
    if key in cache:
        return {c: {"stage": tuple(v["stage"]), "fluxk": int(v["fluxk"])} for c, v in cache[key].items()}
    choice = {c: dict(DEFAULT) for c in census}
    scratch = [solver.backend.alloc(blk.state.shape, name="tune-state", lead=solver.ng) for blk in solver.blocks]
    for blk, s in zip(solver.blocks, scratch):
        s.t.copy_(blk.state.t)
    out = solver.integ.nregs - 1  # the derivative register, which the fused stage never writes
    (terms,) = solver._combine_terms(0)
    src = native.ints([t[0] for t in terms])
    w = (ctypes.c_double * len(terms))(*[t[1] for t in terms])
    withdt = native.ints([t[2] for t in terms])
    L = solver.case.L

    def measure(kind, trial):
        batch = solver._make_batch(trial, regs=(0, out), states=scratch)
        try:
            if kind == "stage":
                return _time(solver, lambda: native.check(
                    L.pga_batch_stage(batch.handle, 0, 1, len(terms), src, w, withdt, solver.dt.ptr,
                                      solver.backend.stream), "pga_batch_stage", L))
            return _time(solver, lambda: native.check(L.pga_batch_flux(batch.handle, 0, 2, solver.backend.stream),
                                                      "pga_batch_flux", L))
        finally:
            batch.close()

    for c in sorted(census):
        for kind, cands in (("stage", STAGE_TILES), ("fluxk", FLUXK_TILES)):
            best = None
            for cand in cands:
                trial = {k: dict(v) for k, v in choice.items()}
                trial[c][kind] = cand
                ms = measure(kind, trial)
                if best is None or ms < best[0]:
                    best = (ms, cand)
            choice[c][kind] = best[1]
    solver.synchronize()
    del scratch
    cache[key] = {c: {"stage": list(v["stage"]), "fluxk": v["fluxk"]} for c, v in choice.items()}
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(cache, indent=1, sort_keys=True))
    except OSError:
        pass
    return choice
