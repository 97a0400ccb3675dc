"""Fluid surface: density splat + marching cubes, both on the GPU.

Replaces solver/surface.py's update() (solver/surface.py:201-251): the splat
of shaders/surface_splat.glsl over the surface lattice, then - instead of the
read-back and the pure-Python marching cubes of solver/marching_cubes.py -
a GPU classify / scan / emit marching cubes whose vertices and triangles stay
on the device until the caller wants them (the bmesh rebuild of
solver/surface.py:278-301 is the consumer).
"""

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import FxSurface
from .config import f32, vec32

MAX_SAMPLES = 1 << 20        # solver/surface.py:49
SURFACE_CELL_RADIUS = 2      # solver/surface.py:53, shaders/surface_splat.glsl:29


@dataclass
class SurfaceConfig:
    """solver/surface.py:87-94."""

    dims: tuple
    spacing: float
    sample_count: int
    kernel_radius: float
    iso: float
    lo: tuple

    def struct(self):
        s = FxSurface()
        s.dims[:] = list(self.dims)
        s.lo[:] = list(self.lo)
        s.spacing = self.spacing
        s.kernel_radius = self.kernel_radius
        return s


def resolve(cfg, surface_resolution=48, iso=0.5, max_samples=MAX_SAMPLES):
    """solver/surface.py:97-140 resolve(): lattice one kernel radius past the
    domain on every side, spacing = longest / resolution (x1.25 while the
    sample count exceeds max_samples)."""
    size = tuple(f32(b - a) for a, b in zip(cfg.lo, cfg.hi))
    longest = max(size)
    spacing = max(longest / max(surface_resolution, 1), 1e-6)

    def kernel_radius(sp):
        return min(max(cfg.smoothing_radius, sp * 1.5), SURFACE_CELL_RADIUS * cfg.cell_size)

    dims = (2, 2, 2)
    for _ in range(8):
        margin = kernel_radius(spacing)
        m2 = f32(2 * margin)
        extent = tuple(f32(a + m2) for a in size)
        dims = tuple(max(2, int(-(-axis // spacing)) + 1) for axis in extent)
        if dims[0] * dims[1] * dims[2] <= max_samples:
            break
        spacing *= 1.25
    kr = kernel_radius(spacing)
    m = f32(kr)
    lo = tuple(f32(c - m) for c in vec32(cfg.lo))
    return SurfaceConfig(dims=dims, spacing=spacing, sample_count=dims[0] * dims[1] * dims[2],
                         kernel_radius=kr, iso=iso, lo=lo)


class MarchingCubes:
    """GPU marching cubes with reusable device buffers (grown on demand)."""

    def __init__(self, device="cuda"):
        self.device = torch.device(device)
        self.work = torch.empty(0, dtype=torch.uint8, device=self.device)
        self.verts = torch.empty((1, 3), dtype=torch.float32, device=self.device)
        self.tris = torch.empty((1, 3), dtype=torch.int32, device=self.device)
        self.counts = torch.zeros(2, dtype=torch.int64, device=self.device)
        self.cases = None

    def extract(self, field, dims, origin, spacing, iso, want_cases=False):
        """(verts (V,3) f32, tris (T,3) int32[, cases uint8]) on the device.
        Samples >= iso are inside (solver/marching_cubes.py:191-280)."""
        d = (ctypes.c_int32 * 3)(*[int(c) for c in dims])
        o = (ctypes.c_double * 3)(*[float(c) for c in origin])
        need = _lib.lib().fx_mc_workspace_bytes(d)
        if self.work.numel() < need:
            self.work = torch.empty(int(need * 1.25) + 256, dtype=torch.uint8, device=self.device)
        ncubes = max(0, dims[0] - 1) * max(0, dims[1] - 1) * max(0, dims[2] - 1)
        cases_ptr = None
        if want_cases:
            self.cases = torch.zeros(max(ncubes, 1), dtype=torch.uint8, device=self.device)
            cases_ptr = ctypes.c_void_p(self.cases.data_ptr())
        stream = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        field = field.contiguous()
        for _ in range(2):
            rc = _lib.lib().fx_mc_extract(
                ctypes.c_void_p(field.data_ptr()), d, o, ctypes.c_double(spacing), ctypes.c_float(iso),
                ctypes.c_void_p(self.verts.data_ptr()), ctypes.c_int64(self.verts.shape[0]),
                ctypes.c_void_p(self.tris.data_ptr()), ctypes.c_int64(self.tris.shape[0]),
                ctypes.c_void_p(self.counts.data_ptr()), cases_ptr,
                ctypes.c_void_p(self.work.data_ptr()), ctypes.c_size_t(self.work.numel()), stream)
            _lib.check(rc, "fx_mc_extract")
            nv, nt = (int(c) for c in self.counts.cpu())
            if nv <= self.verts.shape[0] and nt <= self.tris.shape[0]:
                break
            self.verts = torch.empty((max(nv, 1) + nv // 4, 3), dtype=torch.float32, device=self.device)
            self.tris = torch.empty((max(nt, 1) + nt // 4, 3), dtype=torch.int32, device=self.device)
        out = (self.verts[:nv], self.tris[:nt])
        if want_cases:
            out = out + (self.cases[:ncubes],)
        return out


def dedup_triangles(tris, nverts):
    """The triangles bmesh accepts, in order (solver/surface.py:278-301): a
    triangle over a vertex set that already has one, in any winding, is
    dropped.  tris: (T, 3) int32 device tensor; returns a (T', 3) tensor."""
    tris = tris.contiguous()
    n = int(tris.shape[0])
    dev = tris.device
    work = torch.empty(int(_lib.lib().fx_mc_dedup_workspace_bytes(ctypes.c_int64(n))), dtype=torch.uint8, device=dev)
    out = torch.empty((max(n, 1), 3), dtype=torch.int32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(_lib.lib().fx_mc_dedup(
        ctypes.c_void_p(tris.data_ptr()), ctypes.c_int64(n), ctypes.c_int64(int(nverts)),
        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(count.data_ptr()),
        ctypes.c_void_p(work.data_ptr()), ctypes.c_size_t(work.numel()), stream), "fx_mc_dedup")
    return out[: int(count.item())]


def mesh_handoff(verts, tris, domain_matrix=None):
    """Host buffers for the add-on's mesh rebuild (solver/surface.py:255-301):
    world-space vertices, the de-duplicated triangles as Blender's loop
    arrays (``mesh.vertices.foreach_set("co", ...)``,
    ``mesh.loops.foreach_set("vertex_index", ...)``,
    ``mesh.polygons.foreach_set("loop_start" / "loop_total", ...)``), the
    smooth flags, and the surface object's parent-inverse matrix (the inverse
    of the domain's world matrix, :266-268), which cancels the parent
    transform so the world-space vertices stay where they are."""
    kept = dedup_triangles(tris, verts.shape[0]) if tris.shape[0] else tris
    v = verts.detach().cpu().numpy().astype(np.float32)
    t = kept.cpu().numpy().astype(np.int32)
    m = np.eye(4) if domain_matrix is None else np.asarray(domain_matrix, np.float64)
    return {
        "co": v.reshape(-1),
        "vertex_index": t.reshape(-1),
        "loop_start": np.arange(len(t), dtype=np.int32) * 3,
        "loop_total": np.full(len(t), 3, np.int32),
        "use_smooth": np.ones(len(t), bool),
        "matrix_parent_inverse": np.linalg.inv(m),
        "vertices": len(v),
        "triangles": len(t),
    }


class Surface:
    """Per-run surface state: lattice, field buffer, marching cubes."""

    def __init__(self, solver, surface_resolution=48, iso=0.5, max_samples=MAX_SAMPLES):
        self.solver = solver
        self.config = resolve(solver.cfg, surface_resolution, iso, max_samples)
        self.field = torch.zeros(self.config.sample_count, dtype=torch.float32, device=solver.device)
        self.mc = MarchingCubes(solver.device)
        self.vertices = 0
        self.triangles = 0

    def splat(self):
        """Density field at every sample (shaders/surface_splat.glsl)."""
        s = self.solver
        stream = ctypes.c_void_p(torch.cuda.current_stream(s.device).cuda_stream)
        _lib.check(
            _lib.lib().fx_splat(ctypes.byref(s.params_for(s.last_dt)), ctypes.byref(self.config.struct()),
                                ctypes.byref(s.state), ctypes.c_void_p(self.field.data_ptr()), stream),
            "fx_splat")
        return self.field

    def update(self):
        """Splat and extract: (verts, tris) device tensors of this frame's mesh."""
        self.splat()
        c = self.config
        verts, tris = self.mc.extract(self.field, c.dims, c.lo, c.spacing, c.iso)
        self.vertices, self.triangles = int(verts.shape[0]), int(tris.shape[0])
        return verts, tris

    def stats(self):
        c = self.config
        return {"dims": c.dims, "samples": c.sample_count, "spacing": c.spacing, "iso": c.iso,
                "vertices": self.vertices, "triangles": self.triangles}
