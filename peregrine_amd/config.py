"""Case configuration with the reference's sections and keys.

Mirrors ``peregrine/config.py:12-98``: every key has a default and a key not
listed is refused when set, so a misspelled setting fails at once. The
reference's per-backend sections are kept; ``backend-mi355x`` is this
package's own section (tile shape, fused stage on/off).
"""

import yaml

from peregrine_amd.misc import LockedDict


class ConfigError(Exception):
    def __init__(self, setting, msg):
        super().__init__(f"invalid setting {setting}: {msg}")


_DEFAULTS = {
    "simulation": {
        "simulator": "navier-stokes",
        "precision": "double",
        "niter": 1,
        "controller": "fixed",
        "dt": 1e-3,
        "dt-max": 1e-3,
        "cfl": 0.1,
    },
    "mixture": {
        "species": None,
        "eos": "cpg",
        "trans": None,
        "diffusion": "lewis",
        "mixing-rule": "wilke",
        "chemistry": None,
        "bath": None,
        "T-range": None,
        "refit-tol": 1e-3,
        "refit-max-deg": 6,
    },
    "chem-integration": {
        "kind": None,
        "max-substeps": 200,
        "bulk-substeps": 1,
        "stiff-bins": 8,
        "equilibrium-tolerance": 1e-7,
        "implicit-parts": 1,
        "max-sweeps": 8,
        "refine-tolerance": 1e-13,
    },
    "ics": {"p": 101325.0, "u": 0.0, "v": 0.0, "w": 0.0, "T": 300.0, "Y": {}},
    "time-integration": {"integrator": "tvd-rk3"},
    "adv-flux": {
        "formulation": "single",
        "primary": "kepaec",
        "recon": "constant",
        "limiter": None,
        "secondary": None,
        "correction": None,
        "switch": None,
        "values": {},
    },
    "diff-flux": {"formula": "alpha-damping", "values": {}},
    "backend-openmp": {"cxx": "c++", "tile-cells": 128, "tile-elements": 1024, "constform": "literal"},
    "backend-hip": {
        "tile-cells": 128,
        "tile-elements": 1024,
        "launch-threads": 256,
        "launch-waves": 2,
        "constform": "literal",
    },
    # this package's device: which GPU, whether a stage runs as the fused pass or as the
    # per-kernel sequence, the fused stage's k-planes per thread, the k-slabs a fused stage is
    # run in (a slab's face arrays stay in the 256 MB Infinity Cache between the flux sweeps and
    # the divergence that reads them), the side streams several blocks' stages overlap on when
    # launched eagerly, whether steps replay one captured HIP graph, and the march: the k sweep
    # done inside the stage's launch, a thread walking k-march planes of its column (0: off),
    # and jit: a case the main library is not compiled for (species count, mechanism structure)
    # gets a library of its own, built on first use and cached (0: such a case is refused)
    # (the kernels' block shapes are compiled in: 64 lanes along i, waves along j or k)
    "backend-mi355x": {
        "device": 0,
        "fused-stage": True,
        "k-chunk": 1,
        "stage-slabs": 1,
        "block-streams": 4,
        "graphs": True,
        "k-march": 0,
        "batched": "auto",
        "autotune": True,
        "autotune-cache": None,
        "jit": True,
    },
    "halo-exchange": {"kind": "hoststaged"},
}


class Config(LockedDict):
    def __init__(self):
        super().__init__()
        for section, entries in _DEFAULTS.items():
            self[section] = LockedDict(
                {k: (dict(v) if isinstance(v, dict) else v) for k, v in entries.items()}
            ).lock()
        # the case's own names: not locked
        self["bcvals"] = {}
        self["plugins"] = {}
        self.lock()

    @classmethod
    def from_dict(cls, given):
        config = cls()
        for section, entries in (given or {}).items():
            if section not in config:
                raise ConfigError(section, "no such section")
            for key, value in (entries or {}).items():
                config[section][key] = value
        config.validate()
        return config

    def to_dict(self):
        return {s: dict(e) if isinstance(e, dict) else e for s, e in self.items()}

    def validate(self):
        self["simulation"]["dt"] = float(self["simulation"]["dt"])
        if self["simulation"]["precision"] != "double":
            raise ConfigError("simulation.precision", "this build computes the stage in double only")


def read_cfg(fname="./peregrine.yaml"):
    with open(fname) as f:
        return Config.from_dict(yaml.safe_load(f.read()))


def write_cfg(cfg, fname):
    with open(fname, "w") as f:
        yaml.safe_dump(cfg.to_dict(), f, sort_keys=False)
