"""Oracle - TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference's PBF passes, grid build, marching cubes and
seeder (oracle.c / pbf_passes.inc, built by oracle/Makefile into
liboracle.so).  Only tests/, the smoke check in __graft_entry__.py and the
CPU-baseline leg of bench.py import this package; the product path never does.

Parity pins (what checks each function):
  * orc_mt_random / orc_seed_lattice: CPython's random.Random(0) draws and
    uniform() (tests/test_oracle_cpu.py).
  * orc_bitonic: the hand-derived known answer [0, 1, 3, 2] for four equal
    keys; sortedness and multiset equality for random keys.
  * orc_cell_range: closed form from sorted keys (brute force).
  * orc_mc_tables / orc_mc_extract: the golden vectors that
    tests/golden/make_mc_golden.py wrote from solver/marching_cubes.py, and
    the properties of scripts/test_marching_cubes.py.
  * orc_*_f32 / orc_*_f64 passes: closed forms and invariants
    (tests/test_oracle_cpu.py): a single particle's density equals
    mass*poly6(0); lattice density sums; free fall v = g*dt; sum of
    position corrections is zero for mutual pairs; XSPH of a uniform
    velocity field is zero; brute-force O(N^2) neighbour sums equal the
    grid path; clamp and push-out against an analytic box; FP32 within
    1e-5 of FP64.
  * orc_normal: the two-particle closed form of n and curvature (the
    neighbour's density) and an O(N^2) float64 brute force.
  * orc_lambda: lambda itself (the gradient sums of its denominator) by
    O(N^2) float64 brute force.
  * orc_delta: the two-particle closed form, s_corr on and off, with and
    without the mass weighting.
  * orc_xsph: the two-particle closed form with its viscosity factor.
  * orc_splat: O(N * S) float64 brute force over every particle and sample,
    with the collider carve-out.
    Per-pass GLSL-on-Blender output itself is not available anywhere (no
    Blender, no GL context), so agreement with the reference's own
    executable is not pinned beyond these checks.
"""

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None


class FxParams(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("cells", ctypes.c_int32 * 3),
        ("n_cells", ctypes.c_int32),
        ("lo", ctypes.c_float * 3),
        ("cell_size", ctypes.c_float),
        ("hi", ctypes.c_float * 3),
        ("particle_radius", ctypes.c_float),
        ("h", ctypes.c_float),
        ("mass", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("relaxation", ctypes.c_float),
        ("viscosity", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("boundary_damping", ctypes.c_float),
        ("scorr_k", ctypes.c_float),
        ("surface_tension", ctypes.c_float),
        ("collider_dims", ctypes.c_int32 * 3),
        ("collider_voxel", ctypes.c_float),
        ("iterations", ctypes.c_int32),
        ("max_cell_scan", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


class FxSurface(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32 * 3),
        ("lo", ctypes.c_float * 3),
        ("spacing", ctypes.c_float),
        ("kernel_radius", ctypes.c_float),
    ]


def build():
    """Compile liboracle.so (gcc, OpenMP)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        _LIB = ctypes.CDLL(str(path))
    return _LIB


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def params_struct(params):
    """FxParams from an object with the same field names, or pass through."""
    if isinstance(params, FxParams):
        return params
    out = FxParams()
    for name, ctype in FxParams._fields_:
        value = getattr(params, name)
        if hasattr(ctype, "_length_"):
            getattr(out, name)[:] = list(value)
        else:
            setattr(out, name, value)
    return out


def surface_struct(surf):
    if isinstance(surf, FxSurface):
        return surf
    out = FxSurface()
    out.dims[:] = list(surf.dims)
    out.lo[:] = list(surf.lo)
    out.spacing = surf.spacing
    out.kernel_radius = surf.kernel_radius
    return out


def _dt(precision):
    return np.float32 if precision == "f32" else np.float64


def _c4(a, precision):
    return np.ascontiguousarray(a, dtype=_dt(precision)).reshape(-1, 4)


def num_threads():
    return lib().orc_num_threads()


# ---- passes ---------------------------------------------------------------


def predict(params, pos4, vel4, normal4, precision="f32"):
    p = params_struct(params)
    pos4, vel4, normal4 = (_c4(a, precision).copy() for a in (pos4, vel4, normal4))
    pred4 = np.zeros_like(pos4)
    getattr(lib(), f"orc_predict_{precision}")(
        ctypes.byref(p), _ptr(pos4), _ptr(vel4), _ptr(normal4), _ptr(pred4)
    )
    return vel4, pred4


def grid_key(params, pred4, padded, precision="f32"):
    p = params_struct(params)
    pred4 = _c4(pred4, precision)
    keys = np.zeros(padded, np.int32)
    idx = np.zeros(padded, np.int32)
    getattr(lib(), f"orc_grid_key_{precision}")(
        ctypes.byref(p), _ptr(pred4), ctypes.c_int(padded), _ptr(keys), _ptr(idx)
    )
    return keys, idx


def bitonic(keys, idx):
    keys = np.ascontiguousarray(keys, np.int32).copy()
    idx = np.ascontiguousarray(idx, np.int32).copy()
    lib().orc_bitonic(ctypes.c_int(len(keys)), _ptr(keys), _ptr(idx))
    return keys, idx


def stable_sort(keys, idx, max_key):
    keys = np.ascontiguousarray(keys, np.int32).copy()
    idx = np.ascontiguousarray(idx, np.int32).copy()
    rc = lib().orc_stable_sort(ctypes.c_int(len(keys)), _ptr(keys), _ptr(idx), ctypes.c_int(max_key))
    if rc != 0:
        raise MemoryError("orc_stable_sort")
    return keys, idx


def cell_range(n_cells, keys):
    keys = np.ascontiguousarray(keys, np.int32)
    start = np.zeros(n_cells, np.int32)
    end = np.zeros(n_cells, np.int32)
    lib().orc_cell_range(ctypes.c_int(n_cells), ctypes.c_int(len(keys)), _ptr(keys), _ptr(start), _ptr(end))
    return start, end


def padded_count(n):
    """solver/sph.py:353: max(2, next power of two >= n)."""
    return max(2, 1 << (max(1, n) - 1).bit_length())


def build_grid(params, pred4, sort_mode="bitonic", precision="f32"):
    """(keys, idx, start, end) over the padded P slots (sort_mode: bitonic | stable)."""
    p = params_struct(params)
    P = padded_count(p.n)
    pred4 = _c4(pred4, precision)
    keys = np.zeros(P, np.int32)
    idx = np.zeros(P, np.int32)
    start = np.zeros(p.n_cells, np.int32)
    end = np.zeros(p.n_cells, np.int32)
    mode = 0 if sort_mode == "bitonic" else 1
    rc = getattr(lib(), f"orc_build_grid_{precision}")(
        ctypes.byref(p), _ptr(pred4), ctypes.c_int(P), _ptr(keys), _ptr(idx), _ptr(start), _ptr(end), ctypes.c_int(mode)
    )
    if rc != 0:
        raise MemoryError("orc_build_grid")
    return keys, idx, start, end


def _grid_args(idx, start, end):
    return (
        np.ascontiguousarray(idx, np.int32),
        np.ascontiguousarray(start, np.int32),
        np.ascontiguousarray(end, np.int32),
    )


def lambda_(params, pred4, idx, start, end, precision="f32"):
    p = params_struct(params)
    pred4 = _c4(pred4, precision)
    idx, start, end = _grid_args(idx, start, end)
    out = np.zeros_like(pred4)
    getattr(lib(), f"orc_lambda_{precision}")(ctypes.byref(p), _ptr(pred4), _ptr(idx), _ptr(start), _ptr(end), _ptr(out))
    return out


def delta(params, pred4, lambda4, idx, start, end, precision="f32"):
    p = params_struct(params)
    pred4, lambda4 = _c4(pred4, precision), _c4(lambda4, precision)
    idx, start, end = _grid_args(idx, start, end)
    out = np.zeros_like(pred4)
    getattr(lib(), f"orc_delta_{precision}")(
        ctypes.byref(p), _ptr(pred4), _ptr(lambda4), _ptr(idx), _ptr(start), _ptr(end), _ptr(out)
    )
    return out


def apply_delta(params, pred4, delta4, precision="f32"):
    p = params_struct(params)
    pred4 = _c4(pred4, precision).copy()
    getattr(lib(), f"orc_apply_delta_{precision}")(ctypes.byref(p), _ptr(pred4), _ptr(_c4(delta4, precision)))
    return pred4


def velocity(params, pred4, pos4, precision="f32"):
    p = params_struct(params)
    pred4, pos4 = _c4(pred4, precision), _c4(pos4, precision)
    out = np.zeros_like(pred4)
    getattr(lib(), f"orc_velocity_{precision}")(ctypes.byref(p), _ptr(pred4), _ptr(pos4), _ptr(out))
    return out


def xsph(params, pred4, vel4, lambda4, idx, start, end, precision="f32"):
    p = params_struct(params)
    pred4, vel4, lambda4 = (_c4(a, precision) for a in (pred4, vel4, lambda4))
    idx, start, end = _grid_args(idx, start, end)
    out = np.zeros_like(pred4)
    getattr(lib(), f"orc_xsph_{precision}")(
        ctypes.byref(p), _ptr(pred4), _ptr(vel4), _ptr(lambda4), _ptr(idx), _ptr(start), _ptr(end), _ptr(out)
    )
    return out


def normal(params, pred4, lambda4, idx, start, end, precision="f32"):
    p = params_struct(params)
    pred4, lambda4 = _c4(pred4, precision), _c4(lambda4, precision)
    idx, start, end = _grid_args(idx, start, end)
    out = np.zeros_like(pred4)
    getattr(lib(), f"orc_normal_{precision}")(
        ctypes.byref(p), _ptr(pred4), _ptr(lambda4), _ptr(idx), _ptr(start), _ptr(end), _ptr(out)
    )
    return out


def finalize(params, vel4, delta4, pred4, collider, precision="f32"):
    """Returns (pos4, vel4)."""
    p = params_struct(params)
    vel4 = _c4(vel4, precision).copy()
    delta4, pred4 = _c4(delta4, precision), _c4(pred4, precision)
    col = None if collider is None else np.ascontiguousarray(collider, np.uint8)
    pos4 = np.zeros_like(pred4)
    getattr(lib(), f"orc_finalize_{precision}")(
        ctypes.byref(p), _ptr(vel4), _ptr(delta4), _ptr(pred4), _ptr(col), _ptr(pos4)
    )
    return pos4, vel4


def splat(params, surf, pos4, idx, start, end, collider, precision="f32"):
    p = params_struct(params)
    s = surface_struct(surf)
    pos4 = _c4(pos4, precision)
    idx, start, end = _grid_args(idx, start, end)
    col = None if collider is None else np.ascontiguousarray(collider, np.uint8)
    field = np.zeros(int(np.prod(list(s.dims))), _dt(precision))
    getattr(lib(), f"orc_splat_{precision}")(
        ctypes.byref(p), ctypes.byref(s), _ptr(pos4), _ptr(idx), _ptr(start), _ptr(end), _ptr(col), _ptr(field)
    )
    return field


class State:
    """Reference-layout state (AoS-4 arrays, padded keys) for substep()."""

    def __init__(self, params, pos4, vel4=None, precision="f32", sort_mode="bitonic"):
        self.params = params_struct(params)
        self.precision = precision
        self.sort_mode = sort_mode
        n = self.params.n
        self.pos = _c4(pos4, precision).copy()
        self.vel = np.zeros_like(self.pos) if vel4 is None else _c4(vel4, precision).copy()
        self.pred = self.pos.copy()
        self.lam = np.zeros_like(self.pos)
        self.delta = np.zeros_like(self.pos)
        self.normal = np.zeros_like(self.pos)
        P = padded_count(n)
        self.keys, self.idx, self.start, self.end = build_grid(
            self.params, self.pred, sort_mode, precision
        )
        self.P = P

    def substep(self, collider=None):
        col = None if collider is None else np.ascontiguousarray(collider, np.uint8)
        rc = getattr(lib(), f"orc_substep_{self.precision}")(
            ctypes.byref(self.params),
            _ptr(self.pos), _ptr(self.vel), _ptr(self.pred), _ptr(self.lam), _ptr(self.delta), _ptr(self.normal),
            ctypes.c_int(self.P), _ptr(self.keys), _ptr(self.idx), _ptr(self.start), _ptr(self.end),
            _ptr(col), ctypes.c_int(0 if self.sort_mode == "bitonic" else 1),
        )
        if rc != 0:
            raise MemoryError("orc_substep")


# ---- RNG / seeding --------------------------------------------------------


def mt_random(seed, n):
    out = np.zeros(n, np.float64)
    lib().orc_mt_random(ctypes.c_uint32(seed), ctypes.c_long(n), _ptr(out))
    return out


def seed_lattice(base, spacing, jitter, counts, seed=0):
    nx, ny, nz = counts
    out = np.zeros((nx * ny * nz, 4), np.float32)
    b = np.asarray(base, np.float64)
    lib().orc_seed_lattice(
        _ptr(b), ctypes.c_double(spacing), ctypes.c_double(jitter), ctypes.c_int(nx), ctypes.c_int(ny),
        ctypes.c_int(nz), ctypes.c_uint32(seed), _ptr(out)
    )
    return out


# ---- marching cubes -------------------------------------------------------


def mc_tables():
    tri = np.zeros((256, 16), np.int8)
    mask = np.zeros(256, np.uint16)
    walk = np.zeros((12, 3), np.int32)
    lib().orc_mc_tables(_ptr(tri), _ptr(mask), _ptr(walk))
    return tri, mask, walk


def mc_extract(field, dims, origin, spacing, iso, with_cases=False):
    """(verts f64 (V,3), tris int32 (T,3)[, cases uint8]) in reference order."""
    field = np.ascontiguousarray(field, np.float32)
    d = np.asarray(dims, np.int32)
    o = np.asarray(origin, np.float64)
    counts = np.zeros(2, np.int64)
    ncubes = max(0, (int(d[0]) - 1)) * max(0, (int(d[1]) - 1)) * max(0, (int(d[2]) - 1))
    cases = np.zeros(max(ncubes, 1), np.uint8)
    L = lib()
    fn = L.orc_mc_extract
    vcap = tcap = 0
    verts = np.zeros((1, 3), np.float64)
    tris = np.zeros((1, 3), np.int32)
    for _ in range(2):
        rc = fn(
            _ptr(field), _ptr(d), _ptr(o), ctypes.c_double(spacing), ctypes.c_double(iso),
            _ptr(verts), ctypes.c_long(vcap), _ptr(tris), ctypes.c_long(tcap), _ptr(counts), _ptr(cases)
        )
        if rc != 0:
            raise MemoryError("orc_mc_extract")
        nv, nt = int(counts[0]), int(counts[1])
        if nv <= vcap and nt <= tcap:
            break
        vcap, tcap = nv, nt
        verts = np.zeros((max(nv, 1), 3), np.float64)
        tris = np.zeros((max(nt, 1), 3), np.int32)
    out = (verts[:nv], tris[:nt])
    if with_cases:
        out = out + (cases[:ncubes],)
    return out


def _set_argtypes():
    # default ctypes int return types are what every entry point returns
    pass


os.environ.setdefault("OMP_PROC_BIND", "false")


# ---- mesh collider voxelization (numpy) -------------------------------------


def voxelize_mesh(dims, origin, voxel, verts_world, tris, bounds_lo, bounds_hi, ray_eps):
    """collision/__init__.py:263-273 _point_inside and :304-367 _rebuild_grid,
    restated with the ray in world space: for each voxel centre inside the
    object's index range (_index_range, :290-294), cast upward along +z; a
    ray cast returns the nearest triangle hit at or beyond its origin
    (inclusive point-in-triangle test in the xy projection, edge-on triangles
    never hit); restart ray_eps beyond each hit; at most 64 casts; odd hit
    count = occupied.  Returns a uint8 (nz, ny, nx) grid."""
    nx, ny, nz = (int(d) for d in dims)
    out = np.zeros((nz, ny, nx), np.uint8)
    V = np.asarray(verts_world, np.float32).astype(np.float64)
    T = np.asarray(tris, np.int64).reshape(-1, 3)
    if len(T) == 0:
        return out
    rng = []
    for a, n in enumerate((nx, ny, nz)):
        lo = max(0, int(np.floor((float(np.float32(bounds_lo[a])) - float(np.float32(origin[a]))) / float(np.float32(voxel)))))
        hi = min(n, int(np.ceil((float(np.float32(bounds_hi[a])) - float(np.float32(origin[a]))) / float(np.float32(voxel)))))
        rng.append((lo, hi))
    ox, oy, oz = (float(np.float32(c)) for c in origin)
    vox = float(np.float32(voxel))
    eps = float(np.float32(ray_eps))
    A, B, C = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    area = (B[:, 0] - A[:, 0]) * (C[:, 1] - A[:, 1]) - (B[:, 1] - A[:, 1]) * (C[:, 0] - A[:, 0])
    for j in range(*rng[1]):
        y = oy + (j + 0.5) * vox
        for i in range(*rng[0]):
            x = ox + (i + 0.5) * vox
            w0 = (B[:, 0] - x) * (C[:, 1] - y) - (B[:, 1] - y) * (C[:, 0] - x)
            w1 = (C[:, 0] - x) * (A[:, 1] - y) - (C[:, 1] - y) * (A[:, 0] - x)
            w2 = (A[:, 0] - x) * (B[:, 1] - y) - (A[:, 1] - y) * (B[:, 0] - x)
            pos = (area > 0) & (w0 >= 0) & (w1 >= 0) & (w2 >= 0)
            neg = (area < 0) & (w0 <= 0) & (w1 <= 0) & (w2 <= 0)
            hit = pos | neg
            if not hit.any():
                continue
            zs = np.sort((w0[hit] * A[hit, 2] + w1[hit] * B[hit, 2] + w2[hit] * C[hit, 2]) / area[hit])
            for k in range(*rng[2]):
                start = oz + (k + 0.5) * vox
                count = 0
                while count < 64:
                    nxt = zs[zs >= start]
                    if len(nxt) == 0:
                        break
                    count += 1
                    start = nxt[0] + eps
                out[k, j, i] = count & 1
    return out
