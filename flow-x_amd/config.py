"""Solver configuration, substep sizing and seeding (host side).

Mirrors the reference's config resolution and seeding so that a scene set up
for the Blender add-on resolves to the same particle lattice, cell grid and
parameters here:

* ``SolverConfig`` / ``substep_dt``  - solver/sph.py:195-249
* ``resolve_config``                 - solver/sph.py:252-303 (the caps are
  parameters that default to the reference's MAX_PARTICLES, MAX_CELLS and
  MAX_CELLS_PER_AXIS, solver/sph.py:88-96)
* ``seed_counts`` / ``seed_positions`` - solver/sph.py:306-347
* ``DomainSettings``                 - the property defaults of
  domain/__init__.py:31-209 (FlowXDomainSettings)

Blender's ``mathutils.Vector`` stores float32 components, so domain bounds,
their differences and the seeding lattice origin are rounded to float32 where
the reference passes them through a Vector; everything else is float64 as in
Python.  Values reach the GPU as the float32 ``FxParams`` block.
"""

import math
from dataclasses import dataclass, field, replace

import numpy as np

from ._lib import FxParams

GRAVITY = -9.81                  # solver/sph.py:83
BOUNDARY_DAMPING = 0.35          # solver/sph.py:86
MAX_PARTICLES = 16384            # solver/sph.py:90
MAX_CELLS = 262144               # solver/sph.py:95
MAX_CELLS_PER_AXIS = 128         # solver/sph.py:96
CFL_FACTOR = 0.25                # solver/sph.py:111
MAX_CELL_SCAN = 512              # shaders/sph_common.glsl:30
UNCAPPED = 1 << 30
FX_FLAG_MASS_WEIGHTED_DELTA = 1  # include/flowx.h


def f32(x):
    """Round like a mathutils.Vector component (float32), as a Python float."""
    return float(np.float32(x))


def vec32(v):
    return tuple(f32(c) for c in v)


@dataclass
class DomainSettings:
    """FlowXDomainSettings defaults (domain/__init__.py:31-209)."""

    resolution: int = 32
    fluid_level: float = 50.0
    smoothing_radius: float = 0.1
    rest_density: float = 1000.0
    pbf_iterations: int = 4
    pbf_relaxation: float = 100.0
    viscosity: float = 0.1
    pbf_scorr_k: float = 0.1
    surface_tension: float = 0.0
    surface_resolution: int = 48
    surface_iso: float = 0.5
    max_substeps: int = 24


@dataclass
class Caps:
    max_particles: int = MAX_PARTICLES
    max_cells: int = MAX_CELLS
    max_cells_per_axis: int = MAX_CELLS_PER_AXIS


REFERENCE_CAPS = Caps()
NO_CAPS = Caps(UNCAPPED, UNCAPPED, UNCAPPED)


@dataclass
class SolverConfig:
    """Resolved run parameters (solver/sph.py:195-249 SolverConfig)."""

    lo: tuple
    hi: tuple
    fill_fraction: float
    smoothing_radius: float
    spacing: float
    mass: float
    rest_density: float
    relaxation: float
    scorr_strength: float
    surface_tension: float
    viscosity: float
    max_substeps: int
    iterations: int
    cell_size: float
    cell_dims: tuple
    particle_count: int = 0
    lattice: tuple = None           # explicit (nx, ny, nz) seeding lattice, or None
    seed: int = 0
    collider_dims: tuple = (1, 1, 1)
    collider_voxel: float = 0.0
    max_cell_scan: int = MAX_CELL_SCAN
    flags: int = 0                  # FX_FLAG_* (0 = reference semantics)
    extra: dict = field(default_factory=dict)

    @property
    def cell_count(self):
        return self.cell_dims[0] * self.cell_dims[1] * self.cell_dims[2]

    @property
    def sorted_count(self):
        """Padded sort size of the reference's bitonic network (solver/sph.py:353)."""
        return max(2, 1 << (max(1, self.particle_count) - 1).bit_length())

    @property
    def particle_radius(self):
        return self.spacing * 0.5

    @property
    def size(self):
        return tuple(f32(h - l) for l, h in zip(self.lo, self.hi))

    def substep_dt(self, frame_dt):
        """(substeps, dt) for one frame - solver/sph.py:228-249."""
        gravity_dt = CFL_FACTOR * math.sqrt(self.smoothing_radius / abs(GRAVITY))
        kinematic = self.viscosity / max(self.rest_density, 1e-6)
        viscous_dt = 0.125 * self.smoothing_radius**2 / max(kinematic, 1e-12)
        limit = min(gravity_dt, viscous_dt)
        steps = min(max(1, math.ceil(frame_dt / limit)), self.max_substeps)
        return steps, min(frame_dt / steps, limit)

    def params(self, dt):
        """The float32 FxParams block for a substep of length dt
        (the push-constant block of solver/sph.py:433-461)."""
        p = FxParams()
        p.n = int(self.particle_count)
        p.cells[:] = list(self.cell_dims)
        p.n_cells = int(self.cell_count)
        p.lo[:] = list(self.lo)
        p.cell_size = self.cell_size
        p.hi[:] = list(self.hi)
        p.particle_radius = self.particle_radius
        p.h = self.smoothing_radius
        p.mass = self.mass
        p.rest_density = self.rest_density
        p.relaxation = self.relaxation
        p.viscosity = self.viscosity
        p.dt = dt
        p.gravity = GRAVITY
        p.boundary_damping = BOUNDARY_DAMPING
        p.scorr_k = self.scorr_strength
        p.surface_tension = self.surface_tension
        p.collider_dims[:] = list(self.collider_dims)
        p.collider_voxel = self.collider_voxel
        p.iterations = int(self.iterations)
        p.max_cell_scan = int(self.max_cell_scan)
        p.flags = int(self.flags)
        return p


def seed_counts(size, fill_height, spacing):
    """solver/sph.py:306-313: particles per axis, inset by one spacing."""
    margin = spacing
    return (
        max(1, int((size[0] - margin) / spacing)),
        max(1, int((size[1] - margin) / spacing)),
        max(1, int(max(fill_height - margin, 0.0) / spacing)),
    )


def _cell_grid(size, smoothing_radius, caps):
    """Cell size and dims (solver/sph.py:286-299)."""
    cell = max(smoothing_radius, max(size) / caps.max_cells_per_axis)
    dims = (1, 1, 1)
    for _ in range(8):
        dims = tuple(min(caps.max_cells_per_axis, max(1, math.ceil(axis / cell))) for axis in size)
        if dims[0] * dims[1] * dims[2] <= caps.max_cells:
            break
        cell *= 1.5
    return cell, dims


def resolve_config(settings, lo, hi, caps=REFERENCE_CAPS):
    """solver/sph.py:252-303 _resolve_config for a domain with world bounds
    (lo, hi) and FlowXDomainSettings-like `settings`."""
    lo, hi = vec32(lo), vec32(hi)
    size = tuple(f32(b - a) for a, b in zip(lo, hi))
    h = settings.smoothing_radius
    spacing = h * 0.5
    fill = settings.fluid_level / 100.0
    fill_height = size[2] * fill
    for _ in range(8):
        nx, ny, nz = seed_counts(size, fill_height, spacing)
        if nx * ny * nz <= caps.max_particles:
            break
        spacing *= (nx * ny * nz / caps.max_particles) ** (1 / 3)
        h = spacing * 2.0
    cell, dims = _cell_grid(size, h, caps)
    cfg = SolverConfig(
        lo=lo, hi=hi, fill_fraction=fill, smoothing_radius=h, spacing=spacing,
        mass=settings.rest_density * spacing**3, rest_density=settings.rest_density,
        relaxation=settings.pbf_relaxation, scorr_strength=settings.pbf_scorr_k,
        surface_tension=settings.surface_tension, viscosity=settings.viscosity,
        max_substeps=settings.max_substeps, iterations=settings.pbf_iterations,
        cell_size=cell, cell_dims=dims,
    )
    cfg.extra["max_particles"] = caps.max_particles
    n = int(np.prod(seed_counts(size, fill_height, spacing)))
    cfg.particle_count = min(n, caps.max_particles)
    return cfg


def lattice_config(lo, hi, h, lattice, *, iterations=4, viscosity=0.1, surface_tension=0.0,
                   scorr_k=0.1, relaxation=100.0, rest_density=1000.0, max_substeps=24,
                   caps=NO_CAPS, seed=0):
    """A synthetic scene on an explicit nx x ny x nz lattice of spacing h/2
    starting half a spacing inside lo, with the reference's jitter and RNG."""
    lo, hi = vec32(lo), vec32(hi)
    size = tuple(f32(b - a) for a, b in zip(lo, hi))
    spacing = h * 0.5
    cell, dims = _cell_grid(size, h, caps)
    nx, ny, nz = lattice
    return SolverConfig(
        lo=lo, hi=hi, fill_fraction=nz * spacing / size[2], smoothing_radius=h, spacing=spacing,
        mass=rest_density * spacing**3, rest_density=rest_density, relaxation=relaxation,
        scorr_strength=scorr_k, surface_tension=surface_tension, viscosity=viscosity,
        max_substeps=max_substeps, iterations=iterations, cell_size=cell, cell_dims=dims,
        particle_count=nx * ny * nz, lattice=(nx, ny, nz), seed=seed,
    )


def _lattice_positions(base, spacing, counts, seed, limit=None):
    """Jittered lattice: loops k -> j -> i, draws x, y, z from
    random.Random(seed) (numpy's MT19937 with the same init_by_array seeding
    and 53-bit doubles), uniform(a, b) = a + (b - a) * random()."""
    nx, ny, nz = counts
    n = nx * ny * nz
    if limit is not None:
        n = min(n, limit)
    jitter = spacing * 0.1
    a, b = -jitter, jitter
    u = np.random.RandomState([seed]).random_sample(3 * n).reshape(n, 3)
    draws = a + (b - a) * u
    lin = np.arange(n, dtype=np.int64)
    i = lin % nx
    j = (lin // nx) % ny
    k = lin // (nx * ny)
    out = np.empty((n, 4), np.float32)
    out[:, 0] = (base[0] + i * spacing) + draws[:, 0]
    out[:, 1] = (base[1] + j * spacing) + draws[:, 1]
    out[:, 2] = (base[2] + k * spacing) + draws[:, 2]
    out[:, 3] = 1.0
    return out


def lattice_base(lo, spacing):
    """config.lo + Vector((s, s, s)) * 0.5 in float32 (solver/sph.py:331)."""
    half = np.float32(np.float32(spacing) * np.float32(0.5))
    return tuple(float(np.float32(np.float32(c) + half)) for c in lo)


def seed_positions(cfg):
    """solver/sph.py:316-347 _seed_positions -> (N, 4) float32, w = 1."""
    base = lattice_base(cfg.lo, cfg.spacing)
    if cfg.lattice is not None:
        counts = cfg.lattice
        limit = None
    else:
        size = cfg.size
        counts = seed_counts(size, size[2] * cfg.fill_fraction, cfg.spacing)
        limit = cfg.extra.get("max_particles", MAX_PARTICLES)
    pos = _lattice_positions(base, cfg.spacing, counts, cfg.seed, limit)
    cfg.particle_count = len(pos)
    return pos


# ---- the benchmark scenes (2 m domain at lo = (-1, -1, 0), the demo placement)

DOMAIN_LO = (-1.0, -1.0, 0.0)
DOMAIN_HI = (1.0, 1.0, 2.0)


def dam_break_config(literal_update=False):
    """25,600-particle dam break: h = 0.1, 40x40x16 lattice (40% fill), 4
    iterations, a 0.4 x 0.4 x 1.0 m box collider at the centre on a 20^3
    voxel grid, 64-resolution surface."""
    cfg = lattice_config(DOMAIN_LO, DOMAIN_HI, 0.1, (40, 40, 16), iterations=4)
    cfg.extra.update(colliders=[("box", (-0.2, -0.2, 0.5), (0.2, 0.2, 1.5))], collider_resolution=20,
                     surface_resolution=64, frames=48, fps=24.0)
    return cfg if literal_update else converging_update(cfg)


def scale_config(literal_update=False):
    """1.64M-particle scene: h = 0.025, 160x160x64 lattice (40% fill), 4
    iterations, cohesion on (sigma = 0.001), XSPH viscosity 0.01, a
    0.3 m pillar and a 1.0 x 0.1 x 0.5 m half-wall on an 80^3 collider grid,
    256-resolution surface every 4 substeps."""
    cfg = lattice_config(DOMAIN_LO, DOMAIN_HI, 0.025, (160, 160, 64), iterations=4, viscosity=0.01,
                         surface_tension=0.001)
    cfg.extra.update(
        colliders=[("cylinder", (0.4, 0.4, 0.0), 0.3, 2.0), ("box", (-1.0, -0.45, 0.0), (0.0, -0.35, 0.5))],
        collider_resolution=80, surface_resolution=256, surface_every=4, substeps=240, fps=24.0,
    )
    return cfg if literal_update else converging_update(cfg)


def animated_config(literal_update=False):
    """~4.1M particles: h = 0.018, 222x222x83 lattice, 6 iterations, a 0.3 m
    sphere dropped ballistically from z = 1.8 m re-voxelized on a 111^3 grid
    every frame, 384-resolution surface every frame."""
    cfg = lattice_config(DOMAIN_LO, DOMAIN_HI, 0.018, (222, 222, 83), iterations=6)
    cfg.extra.update(sphere=((0.0, 0.0, 1.8), 0.3), collider_resolution=111, surface_resolution=384,
                     frames=120, fps=24.0)
    return cfg if literal_update else converging_update(cfg)


def adjoint_config(literal_update=False):
    """200k-particle differentiable run: h = 0.05, 80x80x31 lattice, 4
    iterations, a 0.4 m box collider, 64-resolution loss grid, 60 substeps."""
    cfg = lattice_config(DOMAIN_LO, DOMAIN_HI, 0.05, (80, 80, 31), iterations=4)
    cfg.extra.update(colliders=[("box", (-0.2, -0.2, 0.0), (0.2, 0.2, 0.4))], collider_resolution=40,
                     surface_resolution=64, substeps=60, fps=24.0)
    return cfg if literal_update else converging_update(cfg)


def batch_configs(count=64, seed=0):
    """Independent scenes with particle counts log-uniform in [1e4, 4e5],
    fill uniform in [0.2, 0.6], h uniform in [0.03, 0.08] and 0-3 random
    boxes each (numpy PCG64(seed)).  Each scene is a cube of side L sized so
    an (L/s) x (L/s) x (fill L/s) lattice holds the requested count."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes = []
    for _ in range(count):
        target = float(np.exp(rng.uniform(np.log(1e4), np.log(4e5))))
        fill = float(rng.uniform(0.2, 0.6))
        h = float(rng.uniform(0.03, 0.08))
        s = h * 0.5
        side = s * (target / fill) ** (1.0 / 3.0)
        nxy = max(2, int(round(side / s)))
        side = nxy * s
        nz = max(1, int(round(fill * side / s)))
        lo, hi = (-side / 2, -side / 2, 0.0), (side / 2, side / 2, side)
        cfg = converging_update(lattice_config(lo, hi, h, (nxy, nxy, nz), iterations=4))
        boxes = []
        for _b in range(int(rng.integers(0, 4))):
            c = rng.uniform([-side / 2, -side / 2, 0.0], [side / 2, side / 2, side * 0.6])
            e = rng.uniform(0.05, 0.25, size=3) * side
            boxes.append(("box", tuple(c - e / 2), tuple(c + e / 2)))
        cfg.extra.update(colliders=boxes, collider_resolution=max(4, int(round(side / h))), fps=24.0)
        scenes.append(cfg)
    return scenes


def with_dims(cfg, **kw):
    return replace(cfg, **kw)


def converging_update(cfg):
    """Switch a config to the mass-weighted position correction with s_corr
    off.  The reference's literal update (shaders/sph_delta.glsl:81-84, no
    particle mass) overshoots the Newton step by 1/m and its s_corr term
    (k = 0.1) is orders of magnitude larger than lambda at these particle
    masses, so every benchmark scene diverges within its first substep under
    the literal update.  The kernels are the same; one multiply differs."""
    cfg.flags |= FX_FLAG_MASS_WEIGHTED_DELTA
    cfg.scorr_strength = 0.0
    cfg.extra["update"] = "mass-weighted, s_corr off"
    return cfg
