"""peregrine_amd: Peregrine's per-RK-stage right-hand side on MI355X (gfx950).

The case API mirrors the reference (``peregrine/__init__.py``): ``Config``,
``get_backend``, ``mesher.CubeMesher``, ``get_solver``; the stage runs as
hand-written HIP kernels behind the C ABI in ``include/peregrine_amd.h``.
"""

from peregrine_amd import mesher
from peregrine_amd.backend import get_backend
from peregrine_amd.config import Config, ConfigError, read_cfg, write_cfg
from peregrine_amd.solver import Solver, get_solver

__version__ = "0.1.0"

__all__ = ["Config", "ConfigError", "Solver", "get_backend", "get_solver", "mesher", "read_cfg", "write_cfg"]
