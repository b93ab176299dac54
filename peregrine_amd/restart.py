"""Results and restarts: the solution as primitives per block, with the step and the time, the
variable names and the case's config, written and read back (``writers/restart.py``,
``readers/restart.py:10-54``, ``docs/files.md``).

The reference keeps a result in an HDF5 file; neither h5py nor libhdf5 is on this image, so a
result here is numpy's npz container laid out as the reference lays out its file: one group per
block (``results_<block>/<variable>``, the interior only, each variable an array indexed [k, j, i]
as the reference's datasets are) and the attributes ``nrt``, ``tme``, ``privars``, ``variables``
and ``config`` (the case's config as YAML). A result is a complete restart: the solution is all an
integrator carries between steps.

Two departures from the reference's files: a block's group is keyed by its number in this
partition and its arrays are written in the block's current axes (the reference keys by the
grid's original block and undoes a reorientation), so a result restarts the partition that wrote
it, reoriented blocks included, and not another partition of the grid; and a carried species the
file lacks starts at zero, where the reference takes the variable's default.
"""

import numpy as np
import yaml


def exported(sim):
    """The variables a result holds: rho, the primitives and the last species, which the others
    leave (``simulator/euler/simulator.py:55-59``)."""
    names = sim.mixture.spnames
    return ["rho", "p", "u", "v", "w", "T", *names[:-1], names[-1]]


def privars(sim):
    return ["p", "u", "v", "w", "T", *sim.mixture.spnames[:-1]]


def block_variables(solver, blk):
    """{variable: interior array [i, j, k]} of a block, from its solution and state."""
    ng = solver.ng
    Q, st = (a.get()[ng:-ng, ng:-ng, ng:-ng] for a in (blk.Q, blk.state))
    rho = Q[..., 0]
    out = {"rho": rho, "p": st[..., 0], "u": Q[..., 1] / rho, "v": Q[..., 2] / rho, "w": Q[..., 3] / rho,
           "T": st[..., 1]}
    names = solver.sim.mixture.spnames
    Ys = [Q[..., 5 + n] / rho for n in range(len(names) - 1)]
    for name, Y in zip(names[:-1], Ys):
        out[name] = Y
    out[names[-1]] = 1.0 - sum(Ys) if Ys else np.ones_like(rho)
    return out


class RestartWriter:
    """Writes the solver's solution as a numbered result, in single or double precision."""

    def __init__(self, solver, basename="q.{n:08d}.npz", precision="double", cfg=None):
        if precision not in ("single", "double"):
            raise ValueError(f"a result is single or double, not {precision!r}")
        self.basename, self.cfg = basename, cfg
        self.dtype = np.float32 if precision == "single" else np.float64

    def write(self, solver, nrt, tme, directory="."):
        from pathlib import Path

        names = exported(solver.sim)
        arrays = {}
        for blk in solver.blocks:
            data = block_variables(solver, blk)
            for name in names:
                arrays[f"results_{blk.nblki:06d}/{name}"] = np.ascontiguousarray(data[name].T, dtype=self.dtype)
        text = yaml.safe_dump(self.cfg.to_dict(), sort_keys=False) if self.cfg is not None else ""
        attrs = dict(nrt=np.int64(nrt), tme=np.float64(tme), privars=np.array(privars(solver.sim)),
                     variables=np.array(names), config=np.array(text))
        path = Path(directory) / self.basename.format(n=nrt, t=tme)
        with open(path, "wb") as f:
            np.savez(f, **{f"attrs/{k}": v for k, v in attrs.items()}, **arrays)
        return path


class RestartReader:
    """A result read back: its step, time and names, the config it was written with, and the
    primitives it fills a solver's blocks with."""

    def __init__(self, fname):
        self.fname = str(fname)
        with np.load(self.fname, allow_pickle=False) as f:
            self.nrt, self.tme = int(f["attrs/nrt"]), float(f["attrs/tme"])
            self.privars = [str(s) for s in f["attrs/privars"]]
            self.variables = [str(s) for s in f["attrs/variables"]]
            self.text = str(f["attrs/config"])

    @property
    def cfg(self):
        from peregrine_amd.config import Config

        return Config.from_dict(yaml.safe_load(self.text)) if self.text else None

    def fill(self, solver):
        """Each block's primitives over its interior from the file; the halos keep what they held
        (the case's initial values) until the exchange and the walls replace them."""
        names = privars(solver.sim)
        ng = solver.ng
        with np.load(self.fname, allow_pickle=False) as f:
            for blk in solver.blocks:
                q = blk.q.get()
                for n, name in enumerate(names):
                    key = f"results_{blk.nblki:06d}/{name}"
                    if key in f:
                        q[ng:-ng, ng:-ng, ng:-ng, n] = f[key].T
                    elif n < 5:
                        raise ValueError(f"{self.fname} holds no {name}, which a case starts from")
                    else:
                        q[ng:-ng, ng:-ng, ng:-ng, n] = 0.0
                blk.q.set(q)
