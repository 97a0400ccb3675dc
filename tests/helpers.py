"""Shared helpers for the parity tests (GPU calls go through the C ABI)."""

import ctypes

import numpy as np


def rel_err(a, b, scale):
    """max |a - b| / max(|b|, scale): the 1e-5 relative bound with a scale
    floor for quantities that cross zero."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), scale)))


def rms(x):
    x = np.asarray(x, np.float64)
    return float(np.sqrt(np.mean(x * x))) if x.size else 0.0


class GpuState:
    """Device buffers for one FxState, filled from numpy arrays."""

    def __init__(self, fx, params, collider=None, **arrays):
        import torch

        self.torch = torch
        self.fx = fx
        self.params = params
        n, c = params.n, params.n_cells
        dev = "cuda"
        f4 = lambda name: torch.as_tensor(np.asarray(arrays[name], np.float32).reshape(n, 4)).to(dev).contiguous() \
            if name in arrays else torch.zeros((n, 4), dtype=torch.float32, device=dev)
        self.pos, self.vel, self.pred = f4("pos"), f4("vel"), f4("pred")
        self.lam, self.delta, self.normal = f4("lam"), f4("delta"), f4("normal")
        i32 = lambda name, size, fill=0: torch.as_tensor(np.asarray(arrays[name], np.int32)).to(dev).contiguous() \
            if name in arrays else torch.full((size,), fill, dtype=torch.int32, device=dev)
        self.key = i32("key", n)
        self.sorted_key = i32("sorted_key", n)
        self.sorted_idx = i32("sorted_idx", n)
        self.cell_start = i32("cell_start", c, -1)
        self.cell_end = i32("cell_end", c, -1)
        self.collider = None
        if collider is not None:
            self.collider = torch.as_tensor(np.asarray(collider, np.uint8)).to(dev).contiguous()
        need = fx._lib.lib().fx_workspace_bytes(ctypes.byref(params))
        self.work = torch.zeros(int(need), dtype=torch.uint8, device=dev)
        s = fx._lib.FxState()
        P = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        s.pos, s.vel, s.pred, s.lambda_, s.delta, s.normal = (P(self.pos), P(self.vel), P(self.pred), P(self.lam),
                                                             P(self.delta), P(self.normal))
        s.key, s.sorted_key, s.sorted_idx = P(self.key), P(self.sorted_key), P(self.sorted_idx)
        s.cell_start, s.cell_end = P(self.cell_start), P(self.cell_end)
        s.collider = P(self.collider)
        s.work, s.work_bytes = P(self.work), self.work.numel()
        self.state = s

    def stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    def run(self, name, params=None):
        p = params or self.params
        L = self.fx._lib
        if name in L.PASS_IDS:
            rc = L.lib().fx_pass(L.PASS_IDS[name], ctypes.byref(p), ctypes.byref(self.state), self.stream())
        else:
            rc = getattr(L.lib(), f"fx_{name}")(ctypes.byref(p), ctypes.byref(self.state), self.stream())
        L.check(rc, name)
        self.torch.cuda.synchronize()

    def get(self, name):
        return getattr(self, name).cpu().numpy()


# ---- closed triangle meshes for the mesh-collider tests ----------------------


def box_mesh(lo, hi):
    """Closed axis-aligned box: 8 vertices, 12 outward-wound triangles."""
    import numpy as np

    (x0, y0, z0), (x1, y1, z1) = lo, hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float64)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
    t = []
    for a, b, c, d in quads:
        t += [(a, b, c), (a, c, d)]
    return v, np.array(t, np.int32)


def uv_sphere(centre, r, nu=48, nv=24):
    """Closed UV sphere (poles as single vertices), fanned into triangles."""
    import numpy as np

    v = [(0.0, 0.0, -r)]
    for j in range(1, nv):
        th = np.pi * j / nv - np.pi / 2
        for i in range(nu):
            ph = 2 * np.pi * i / nu
            v.append((r * np.cos(th) * np.cos(ph), r * np.cos(th) * np.sin(ph), r * np.sin(th)))
    v.append((0.0, 0.0, r))
    v = np.array(v) + np.asarray(centre)
    t = []
    ring = lambda j, i: 1 + (j - 1) * nu + (i % nu)
    for i in range(nu):
        t.append((0, ring(1, i + 1), ring(1, i)))
        t.append((len(v) - 1, ring(nv - 1, i), ring(nv - 1, i + 1)))
    for j in range(1, nv - 1):
        for i in range(nu):
            a, b, c, d = ring(j, i), ring(j, i + 1), ring(j + 1, i + 1), ring(j + 1, i)
            t += [(a, b, c), (a, c, d)]
    return v, np.array(t, np.int32)


def rotation(axis, angle):
    import numpy as np

    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    x, y, z = axis
    c, s = np.cos(angle), np.sin(angle)
    return np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                     [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                     [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
