"""flowx_amd - MI355X-native Position Based Fluids step / reset / readback /
surface path behind the C ABI of include/flowx.h (libflowx_hip.so, gfx950).

The directory is named flow-x_amd; load it as the package ``flowx_amd`` with
``__graft_entry__.load_package()`` (or add a spec loader of your own)."""

from . import _lib
from .batch import BatchSolver
from .cache import FrameCache, config_hash
from .collision import ColliderGrid, grid_geometry
from .config import (
    DomainSettings,
    SolverConfig,
    adjoint_config,
    animated_config,
    batch_configs,
    dam_break_config,
    lattice_config,
    resolve_config,
    scale_config,
    seed_positions,
)
from .sph import Solver
from .surface import MarchingCubes, Surface, SurfaceConfig, dedup_triangles, mesh_handoff, resolve
from .timeline import Timeline

__all__ = [
    "BatchSolver", "ColliderGrid", "DomainSettings", "FrameCache", "MarchingCubes", "Timeline", "Solver", "SolverConfig", "Surface", "SurfaceConfig",
    "adjoint_config", "animated_config", "batch_configs", "config_hash", "dam_break_config", "dedup_triangles", "grid_geometry", "mesh_handoff",
    "lattice_config", "resolve", "resolve_config", "scale_config", "seed_positions",
]
