"""The PBF solver driver: step / reset / readback on the GPU.

A drop-in for the hot path of solver/sph.py: allocation (_allocate,
:350-369), the grid build (_build_grid, :482-498), the substep (_substep,
:501-514), the frame step (_step, :517-545), reset (_seed / reseed,
:585-662), cached-state upload (_apply_cached, :700-733) and the position /
velocity / density read-backs (gpu_util.read_texture, solver/gpu_util.py:62-89).

Device state lives in torch tensors on the ROCm device (AoS-4 float32, the
reference's RGBA32F texture layout); every pass runs in libflowx_hip.so.  The
Blender timeline semantics of solver/sph.py:773-855 are not part of this path.
"""

import ctypes
import time
from collections import deque

import numpy as np
import torch

from . import _lib, autotune
from ._lib import FxState
from .collision import ColliderGrid
from .config import seed_positions

TIMING_WINDOW = 30  # solver/sph.py:104


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Solver:
    """GPU-resident PBF state for one scene.

    keep_scratch: also maintain the canonical copies of the reference's
    scratch textures (predicted, lambda, delta, normal, keys, cell ranges) after
    every substep, for callers that read them back; the step itself only
    needs positions and velocities.
    """

    def __init__(self, cfg, device="cuda", keep_scratch=True, colliders=None):
        if not torch.cuda.is_available():
            raise RuntimeError("Solver needs a ROCm GPU (torch.cuda.is_available() is False)")
        _lib.lib()  # fail loudly now if the HIP library is missing
        autotune.apply()  # tuning.json from a previous autotune run, if any
        self.cfg = cfg
        self.device = torch.device(device)
        self.keep_scratch = keep_scratch
        self.colliders = list(colliders if colliders is not None else cfg.extra.get("colliders", []))
        self.collider = None
        self.substeps = 0
        self.last_dt = 0.0
        self.timings = deque(maxlen=TIMING_WINDOW)
        self._seeded = False
        self._build_collider()

    # ---- allocation ------------------------------------------------------
    def _build_collider(self):
        res = self.cfg.extra.get("collider_resolution")
        if self.colliders and res:
            self.collider = ColliderGrid(self.cfg.lo, self.cfg.hi, res, self.device)
            self.collider.rebuild(self.colliders)
            self.cfg.collider_dims = self.collider.dims
            self.cfg.collider_voxel = self.collider.voxel
        else:
            self.collider = None
            self.cfg.collider_dims = (1, 1, 1)
            self.cfg.collider_voxel = 0.0

    def _allocate(self, n):
        dev, f32, i32 = self.device, torch.float32, torch.int32
        c = self.cfg.cell_count
        self.pos = torch.zeros((n, 4), dtype=f32, device=dev)
        self.vel = torch.zeros((n, 4), dtype=f32, device=dev)
        self.pred = torch.zeros((n, 4), dtype=f32, device=dev)
        self.normal = torch.zeros((n, 4), dtype=f32, device=dev)
        if self.keep_scratch:
            self.lam = torch.zeros((n, 4), dtype=f32, device=dev)
            self.delta = torch.zeros((n, 4), dtype=f32, device=dev)
            self.key = torch.zeros(n, dtype=i32, device=dev)
            self.sorted_key = torch.zeros(n, dtype=i32, device=dev)
            self.sorted_idx = torch.zeros(n, dtype=i32, device=dev)
            self.cell_start = torch.full((c,), -1, dtype=i32, device=dev)
            self.cell_end = torch.full((c,), -1, dtype=i32, device=dev)
        else:
            self.lam = self.delta = self.key = self.sorted_key = self.sorted_idx = None
            self.cell_start = self.cell_end = None
        need = _lib.lib().fx_workspace_bytes(ctypes.byref(self.params_for(0.0)))
        self.work = torch.zeros(int(need), dtype=torch.uint8, device=dev)
        s = FxState()
        s.pos, s.vel, s.pred = _ptr(self.pos), _ptr(self.vel), _ptr(self.pred)
        s.lambda_, s.delta, s.normal = _ptr(self.lam), _ptr(self.delta), _ptr(self.normal)
        s.key, s.sorted_key, s.sorted_idx = _ptr(self.key), _ptr(self.sorted_key), _ptr(self.sorted_idx)
        s.cell_start, s.cell_end = _ptr(self.cell_start), _ptr(self.cell_end)
        s.collider = _ptr(self.collider.grid) if self.collider is not None else None
        s.work, s.work_bytes = _ptr(self.work), self.work.numel()
        self.state = s

    def params_for(self, dt):
        return self.cfg.params(dt)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    # ---- reset / load ----------------------------------------------------
    def reset(self, positions=None):
        """Seed (solver/sph.py:585-611 _seed): jittered lattice, zero velocity,
        pred = positions, lambda = 0, grid built from pred (a zero-length
        step)."""
        pos = seed_positions(self.cfg) if positions is None else np.asarray(positions, np.float32)
        self.cfg.particle_count = len(pos)
        self._allocate(len(pos))
        self.load_state(pos, None)
        self.substeps = 0
        self.timings.clear()
        self._seeded = True

    def load_state(self, positions, velocities=None):
        """_apply_cached (solver/sph.py:700-733): install positions and
        velocities (N x 4 float32, or N x 3), predicted = positions, lambda = 0,
        then rebuild the grid from predicted."""
        n = self.cfg.particle_count
        p = torch.as_tensor(np.asarray(positions, np.float32)).reshape(n, -1)
        self.pos[:, : p.shape[1]].copy_(p[:, :4].to(self.device))
        if velocities is None:
            self.vel.zero_()
        else:
            v = torch.as_tensor(np.asarray(velocities, np.float32)).reshape(n, -1)
            self.vel[:, : v.shape[1]].copy_(v[:, :4].to(self.device))
        _lib.check(_lib.lib().fx_load_state(ctypes.byref(self.params_for(0.0)), ctypes.byref(self.state),
                                            self._stream()), "fx_load_state")
        self._seeded = True

    # ---- stepping ----------------------------------------------------------
    def build_grid(self, dt=0.0):
        """_build_grid (solver/sph.py:482-498) from the current predicted positions."""
        _lib.check(_lib.lib().fx_build_grid(ctypes.byref(self.params_for(dt)), ctypes.byref(self.state),
                                            self._stream()), "fx_build_grid")

    def substep(self, dt):
        """_substep (solver/sph.py:501-514)."""
        self._require_seeded()
        self.last_dt = dt
        _lib.check(_lib.lib().fx_substep(ctypes.byref(self.params_for(dt)), ctypes.byref(self.state),
                                         self._stream()), "fx_substep")

    def substeps_n(self, dt, count):
        self._require_seeded()
        self.last_dt = dt
        _lib.check(_lib.lib().fx_step(ctypes.byref(self.params_for(dt)), ctypes.byref(self.state),
                                      int(count), self._stream()), "fx_step")

    def capture(self, dt, substeps):
        """Capture `substeps` substeps of length dt in a HIP graph and return
        it; graph.replay() then runs them with one host call.  They are
        recorded as one fx_step call, so consecutive substeps hand their state
        on by slot as an eager substeps_n call does.  What the recorded
        substeps find in the workspace (whether the normals of the previous
        substep are already computed) is decided by the kernels on every
        replay, so load_state / reset between replays are safe.  One warm-up
        substep runs eagerly first (capture itself executes nothing), so the
        state advances by one substep here."""
        self._require_seeded()
        self.last_dt = dt
        torch.cuda.synchronize(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.substep(dt)
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.substeps_n(dt, substeps)
        return graph

    def step(self, frame_dt):
        """_step (solver/sph.py:517-545) without the Blender outputs: the
        frame's substeps, timed end to end (the wall time includes the GPU
        work because the caller reads results back)."""
        started = time.perf_counter()
        substeps, dt = self.cfg.substep_dt(frame_dt)
        self.substeps = substeps
        self.substeps_n(dt, substeps)
        torch.cuda.current_stream(self.device).synchronize()
        self.timings.append((time.perf_counter() - started) * 1000.0)
        return substeps, dt

    def _require_seeded(self):
        if not self._seeded:
            raise RuntimeError("Solver.reset() or load_state() must run before stepping")

    # ---- read-back (solver/gpu_util.py:62-89) ------------------------------
    # Every read-back checks that what it returns is finite (check=False
    # skips it, e.g. to inspect a run that has diverged).
    @staticmethod
    def _finite(a, what, check):
        if check and not np.all(np.isfinite(a)):
            bad = int(np.count_nonzero(~np.all(np.isfinite(a.reshape(len(a), -1)), axis=1)))
            raise RuntimeError(f"{what}: {bad} of {len(a)} particles are not finite (the run diverged)")
        return a

    def read_positions(self, check=True):
        """N x 4 float32 (w = 1), as read_texture(positions_img) returns."""
        return self._finite(self.pos.cpu().numpy(), "positions", check)

    def read_velocities(self, check=True):
        """N x 4 float32 (w = 0)."""
        return self._finite(self.vel.cpu().numpy(), "velocities", check)

    def read_density(self, check=True):
        """lambda_img.x: density from the last Jacobi lambda pass
        (scripts/smoke_test.py:117-134)."""
        if self.lam is None:
            raise RuntimeError("read_density needs keep_scratch=True")
        return self._finite(self.lam[:, 0].cpu().numpy(), "densities", check)

    def stats(self):
        """solver/sph.py:862-886 stats(), minus the Blender timeline fields."""
        c = self.cfg
        t = self.timings
        return {
            "particles": c.particle_count,
            "cells": c.cell_count,
            "cell_dims": c.cell_dims,
            "smoothing_radius": c.smoothing_radius,
            "spacing": c.spacing,
            "substeps": self.substeps,
            "step_ms": sum(t) / len(t) if t else None,
        }
