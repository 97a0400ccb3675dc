"""Collider occupancy grid: geometry and GPU voxelization of analytic shapes.

Geometry follows collision/__init__.py:245-260 (_domain_grid_geometry): the
grid shares the domain's low corner as origin, voxel = longest axis /
resolution, dims = max(1, round(axis / voxel)).  Occupancy is decided at voxel
centres origin + (i + 0.5) * voxel and stored uint8 at (k*ny + j)*nx + i, the
union over every collider (collision/__init__.py:225-242, :353-366).  The
reference voxelizes Blender meshes by BVH ray parity; this path voxelizes the
analytic shapes its benchmark scenes use (boxes, a vertical cylinder, a
sphere) on the GPU, fast enough to redo every frame for a moving collider.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .config import f32, vec32


def grid_geometry(lo, hi, resolution):
    """(origin, voxel, dims) - collision/__init__.py:245-255."""
    lo, hi = vec32(lo), vec32(hi)
    size = tuple(f32(b - a) for a, b in zip(lo, hi))
    longest = max(size)
    if longest <= 0.0:
        return lo, 0.0, (0, 0, 0)
    voxel = longest / resolution
    dims = tuple(max(1, round(axis / voxel)) for axis in size)
    return lo, voxel, dims


def _f3(v):
    return (ctypes.c_float * 3)(*[float(c) for c in v])


def _i3(v):
    return (ctypes.c_int32 * 3)(*[int(c) for c in v])


class ColliderGrid:
    """Device uint8 occupancy grid for one domain."""

    def __init__(self, lo, hi, resolution, device="cuda"):
        self.origin, self.voxel, self.dims = grid_geometry(lo, hi, resolution)
        nx, ny, nz = self.dims
        self.grid = torch.zeros(nz * ny * nx, dtype=torch.uint8, device=device)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.grid.device).cuda_stream)

    def clear(self):
        self.grid.zero_()

    def add(self, shape, accumulate=True):
        """OR one shape into the grid.  shape: ("box", lo, hi) |
        ("sphere", centre, r) | ("cylinder", base, r, height) |
        ("mesh", MeshCollider)."""
        kind = shape[0]
        d, o, v = _i3(self.dims), _f3(self.origin), ctypes.c_float(self.voxel)
        g = ctypes.c_void_p(self.grid.data_ptr())
        acc = 1 if accumulate else 0
        if kind == "box":
            boxes = (ctypes.c_float * 6)(*[float(c) for c in (*shape[1], *shape[2])])
            _lib.call("fx_voxelize_boxes", d, o, v, boxes, 1, g, acc, self._stream())
        elif kind == "sphere":
            _lib.call("fx_voxelize_sphere", d, o, v, _f3(shape[1]), ctypes.c_float(shape[2]), g, acc,
                      self._stream())
        elif kind == "cylinder":
            _lib.call("fx_voxelize_cylinder", d, o, v, _f3(shape[1]), ctypes.c_float(shape[2]),
                      ctypes.c_float(shape[3]), g, acc, self._stream())
        elif kind == "mesh":
            m = shape[1]
            flag = torch.zeros(1, dtype=torch.int32, device=self.grid.device)
            _lib.call("fx_voxelize_mesh", d, o, v, ctypes.c_void_p(m.verts.data_ptr()), int(m.verts.shape[0]),
                      ctypes.c_void_p(m.tris.data_ptr()), int(m.tris.shape[0]), _f3(m.bounds_lo), _f3(m.bounds_hi),
                      ctypes.c_float(m.ray_eps), g, acc, ctypes.c_void_p(flag.data_ptr()), self._stream())
            if int(flag.item()):
                raise RuntimeError("mesh collider: a voxel column crossed more than 128 triangles")
        else:
            raise ValueError(f"unknown collider shape {kind!r}")

    def rebuild(self, shapes):
        """Clear, then union every shape (the solver grid of
        collision/__init__.py:225-242)."""
        self.clear()
        for shape in shapes:
            self.add(shape, accumulate=True)

    def occupancy(self):
        nx, ny, nz = self.dims
        return self.grid.cpu().numpy().reshape(nz, ny, nx)


def occupancy_cpu(origin, voxel, dims, shapes):
    """numpy reference of the voxel-centre test (used by host-side callers that
    have no GPU, e.g. cache hashing of shape lists).  Inputs are rounded to
    float32 as the C ABI receives them (origin, voxel size, shape
    parameters); the voxel centres and the tests are float64, as on the GPU."""
    origin = vec32(origin)
    voxel = f32(voxel)
    shapes = [(s[0],) + tuple(vec32(a) if isinstance(a, (tuple, list)) else f32(a) for a in s[1:])
              for s in shapes]
    nx, ny, nz = dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = origin[0] + (i + 0.5) * voxel
    y = origin[1] + (j + 0.5) * voxel
    z = origin[2] + (k + 0.5) * voxel
    occ = np.zeros((nz, ny, nx), bool)
    for shape in shapes:
        if shape[0] == "box":
            a, b = shape[1], shape[2]
            occ |= (x > a[0]) & (x < b[0]) & (y > a[1]) & (y < b[1]) & (z > a[2]) & (z < b[2])
        elif shape[0] == "sphere":
            c, r = shape[1], shape[2]
            occ |= (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < r * r
        elif shape[0] == "cylinder":
            c, r, hgt = shape[1], shape[2], shape[3]
            occ |= ((x - c[0]) ** 2 + (y - c[1]) ** 2 < r * r) & (z > c[2]) & (z < c[2] + hgt)
    return occ.astype(np.uint8)


class MeshCollider:
    """A triangle mesh placed in the world, ready for fx_voxelize_mesh.

    verts_local: (V, 3) local-space vertices; tris: (F, 3) vertex indices
    (polygons fanned into triangles by the caller); matrix: 4 x 4 object-to-
    world transform.  The world bounds are those of the transformed local
    bounding box (the object's world_bounds), and the ray restart step is
    the reference's 1e-4 local units expressed along the world +z ray
    (collision/__init__.py:263-273, :336-346)."""

    def __init__(self, verts_local, tris, matrix=None, device="cuda"):
        v = np.asarray(verts_local, np.float64).reshape(-1, 3)
        m = np.eye(4) if matrix is None else np.asarray(matrix, np.float64)
        world = v @ m[:3, :3].T + m[:3, 3]
        lo, hi = v.min(0), v.max(0)
        corners = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
        cw = corners @ m[:3, :3].T + m[:3, 3]
        self.bounds_lo, self.bounds_hi = cw.min(0), cw.max(0)
        inv3 = np.linalg.inv(m[:3, :3])
        self.ray_eps = 1e-4 / float(np.linalg.norm(inv3 @ np.array([0.0, 0.0, 1.0])))
        self.verts_world = world.astype(np.float32)
        self.tris_np = np.asarray(tris, np.int32).reshape(-1, 3)
        self.verts = torch.as_tensor(self.verts_world).to(device)
        self.tris = torch.as_tensor(self.tris_np).to(device)
