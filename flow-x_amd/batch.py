"""Batched scenes: many independent PBF scenes in one launch per pass.

No reference counterpart (the add-on runs one scene, solver/sph.py:501-514).
Scenes of uneven particle counts are concatenated; one counting sort over
scene-offset cell keys orders all of them, and every pass is a single launch
whose threads look up their scene's parameters in a device table
(fx_batch_* in include/flowx.h).  Each scene advances by its own dt per
substep.  Substeps can be captured in a HIP graph (``capture``) and replayed,
which removes the per-launch host overhead.
"""

import ctypes

import numpy as np
import torch

from . import _lib, autotune
from ._lib import FxScene
from .collision import ColliderGrid
from .config import seed_positions


class BatchSolver:
    def __init__(self, configs, device="cuda", fps=24.0, positions=None):
        """positions: optional list of per-scene (N_i x 4 float32, w = 1)
        starting positions (default: each scene's seeded lattice); scene i's
        particle_count becomes N_i."""
        if not torch.cuda.is_available():
            raise RuntimeError("BatchSolver needs a ROCm GPU")
        _lib.lib()
        autotune.apply()  # tuning.json from a previous autotune run, if any
        self.device = torch.device(device)
        self.configs = list(configs)
        self.fps = fps
        if positions is None:
            positions = [seed_positions(c) for c in self.configs]
        else:
            if len(positions) != len(self.configs):
                raise ValueError("positions: one array per scene")
            positions = [np.ascontiguousarray(np.asarray(p, np.float32).reshape(-1, 4)) for p in positions]
            for c, p in zip(self.configs, positions):
                c.particle_count = len(p)
        self.counts = [len(p) for p in positions]
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        n = int(self.offsets[-1])
        self.pos = torch.as_tensor(np.concatenate(positions)).to(self.device).contiguous()
        self.vel = torch.zeros((n, 4), dtype=torch.float32, device=self.device)
        # colliders: each scene's grid voxelized on the GPU into its own slice
        grids, col_off, total = [], [], 0
        for c in self.configs:
            res = c.extra.get("collider_resolution")
            if c.extra.get("colliders") and res:
                g = ColliderGrid(c.lo, c.hi, res, self.device)
                g.rebuild(c.extra["colliders"])
                c.collider_dims, c.collider_voxel = g.dims, g.voxel
                grids.append(g.grid)
                col_off.append(total)
                total += g.grid.numel()
            else:
                c.collider_dims, c.collider_voxel = (1, 1, 1), 0.0
                col_off.append(-1)
        self.colliders = (torch.cat(grids) if grids else torch.zeros(1, dtype=torch.uint8, device=self.device))
        self.dts = [c.substep_dt(1.0 / fps)[1] for c in self.configs]
        self.scenes = (FxScene * len(self.configs))()
        for i, c in enumerate(self.configs):
            self.scenes[i].params = c.params(self.dts[i])
            self.scenes[i].collider_offset = col_off[i]
        need = _lib.lib().fx_batch_workspace_bytes(self.scenes, len(self.configs))
        if need == 0:
            raise RuntimeError("fx_batch_workspace_bytes rejected the scenes: "
                               + _lib.lib().fx_last_error().decode())
        self.work = torch.zeros(int(need), dtype=torch.uint8, device=self.device)
        self.graph = None
        self.load_state()

    @property
    def particle_count(self):
        return int(self.offsets[-1])

    def _args(self):
        return (self.scenes, len(self.configs), ctypes.c_void_p(self.pos.data_ptr()),
                ctypes.c_void_p(self.vel.data_ptr()), ctypes.c_void_p(self.colliders.data_ptr()),
                ctypes.c_void_p(self.work.data_ptr()), ctypes.c_size_t(self.work.numel()),
                ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))

    def load_state(self):
        _lib.check(_lib.lib().fx_batch_load_state(*self._args()), "fx_batch_load_state")

    def substep(self):
        """One substep of every scene (each with its own dt)."""
        _lib.check(_lib.lib().fx_batch_substep(*self._args()), "fx_batch_substep")

    def capture(self, substeps):
        """Capture `substeps` batched substeps in a HIP graph (replay with
        replay())."""
        torch.cuda.synchronize(self.device)
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):  # warm-up outside the graph
            self.substep()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(substeps):
                self.substep()
        self.graph = g
        self.graph_substeps = substeps
        return g

    def replay(self):
        self.graph.replay()

    def scene_positions(self, i):
        a, b = self.offsets[i], self.offsets[i + 1]
        return self.pos[a:b].cpu().numpy()

    def scene_velocities(self, i):
        a, b = self.offsets[i], self.offsets[i + 1]
        return self.vel[a:b].cpu().numpy()
