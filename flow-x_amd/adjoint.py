"""Differentiable substeps behind torch.autograd.Function.

No reference counterpart.  ``Substep`` advances (x, v) one PBF substep with the
hand-written HIP adjoint of include/flowx.h (fx_adjoint_*): the neighbour
structure is recorded per substep, the substep's inputs are its checkpoint,
and backward recomputes the Jacobi iterates before running the reverse
passes.  Gradients flow to the positions, the velocities, gravity and the
collider translation.  ``Splat`` is the density field of the surface splat and
its adjoint, used for the loss.  float32 and float64 are both supported; the
finite-difference checks run in float64.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import FxSurface
from .collision import ColliderGrid
from .surface import resolve


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class DiffSolver:
    """Device state shared by the autograd functions of one scene."""

    def __init__(self, cfg, dtype=torch.float64, list_skin=0.5, device="cuda", surface_resolution=None):
        if not torch.cuda.is_available():
            raise RuntimeError("DiffSolver needs a ROCm GPU")
        _lib.lib()
        self.cfg = cfg
        self.device = torch.device(device)
        self.dtype = dtype
        self.precision = 8 if dtype == torch.float64 else 4
        self.list_skin = float(list_skin)
        _, self.dt = cfg.substep_dt(1.0 / cfg.extra.get("fps", 24.0))
        res = cfg.extra.get("collider_resolution")
        self.collider = None
        if cfg.extra.get("colliders") and res:
            self.collider = ColliderGrid(cfg.lo, cfg.hi, res, self.device)
            self.collider.rebuild(cfg.extra["colliders"])
            cfg.collider_dims, cfg.collider_voxel = self.collider.dims, self.collider.voxel
        self.params = cfg.params(self.dt)
        need = _lib.lib().fx_adjoint_workspace_bytes(ctypes.byref(self.params), self.precision)
        if need == 0:
            raise RuntimeError("fx_adjoint_workspace_bytes rejected the parameters")
        self.work = torch.zeros(int(need), dtype=torch.uint8, device=self.device)
        self.grad_params = torch.zeros(4, dtype=torch.float64, device=self.device)
        # set on the device by any backward substep whose adjoints are not all
        # finite; read once, by the backward of a rollout's first substep
        self.nonfinite = torch.zeros((), dtype=torch.bool, device=self.device)
        self.surface = resolve(cfg, surface_resolution or cfg.extra.get("surface_resolution", 64), 0.5,
                               max_samples=1 << 30)
        self.sample_mask = self._sample_mask(np.zeros(3))

    def _sample_mask(self, translation):
        """Samples inside a collider voxel (the splat's carve-out,
        shaders/surface_splat.glsl:88-90), in the collider frame at
        `translation`; piecewise constant, so it carries no gradient."""
        s = self.surface
        if self.collider is None:
            return None
        nx, ny, nz = s.dims
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        pts = [np.float32(s.lo[a]) + np.float32([i, j, k][a]).astype(np.float32) * np.float32(s.spacing)
               for a in range(3)]
        vox = np.float32(self.collider.voxel)
        idx = [np.floor((pts[a] - np.float32(translation[a]) - np.float32(self.cfg.lo[a])) / vox).astype(np.int64)
               for a in range(3)]
        kx, ky, kz = self.collider.dims
        inside = (idx[0] >= 0) & (idx[1] >= 0) & (idx[2] >= 0) & (idx[0] < kx) & (idx[1] < ky) & (idx[2] < kz)
        occ = self.collider.occupancy()
        mask = np.zeros((nz, ny, nx), np.uint8)
        mask[inside] = occ[idx[2][inside], idx[1][inside], idx[0][inside]]
        return torch.as_tensor(mask.reshape(-1)).to(self.device)

    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def new_decisions(self):
        """A substep record (branches + neighbour lists) for forward_raw mode 1."""
        nbytes = _lib.lib().fx_adjoint_record_bytes(ctypes.byref(self.params))
        return torch.empty(int(nbytes) // 4, dtype=torch.int32, device=self.device)

    def forward_raw(self, x, v, gravity, translation, decisions=None, mode=0):
        """mode: 0 decide, 1 decide and record into `decisions`, 2 replay them."""
        x_out = torch.empty_like(x)
        v_out = torch.empty_like(v)
        tr = (ctypes.c_double * 3)(*[float(t) for t in translation])
        _lib.check(_lib.lib().fx_adjoint_forward(
            ctypes.byref(self.params), self.precision, ctypes.c_double(float(gravity)), tr,
            ctypes.c_double(self.list_skin), _p(self.collider.grid if self.collider else None),
            _p(decisions), int(mode if decisions is not None else 0),
            _p(x), _p(v), _p(x_out), _p(v_out), _p(self.work), ctypes.c_size_t(self.work.numel()),
            self.stream()), "fx_adjoint_forward")
        return x_out, v_out

    def backward_raw(self, x, v, gravity, translation, gx_out, gv_out, decisions=None):
        gx = torch.zeros_like(x)
        gv = torch.zeros_like(v)
        self.grad_params.zero_()
        tr = (ctypes.c_double * 3)(*[float(t) for t in translation])
        _lib.check(_lib.lib().fx_adjoint_backward(
            ctypes.byref(self.params), self.precision, ctypes.c_double(float(gravity)), tr,
            ctypes.c_double(self.list_skin), _p(self.collider.grid if self.collider else None), _p(decisions),
            _p(x), _p(v), _p(gx_out.contiguous()), _p(gv_out.contiguous()), _p(gx), _p(gv),
            _p(self.grad_params), _p(self.work), ctypes.c_size_t(self.work.numel()), self.stream()),
            "fx_adjoint_backward")
        return gx, gv, self.grad_params.clone()

    def overflowed(self):
        """Whether a neighbour list exceeded its capacity since the last call
        (the device flag is cleared by reading it)."""
        flag = ctypes.c_int32(0)
        _lib.check(_lib.lib().fx_adjoint_overflow(ctypes.byref(self.params), self.precision, _p(self.work),
                                                  ctypes.byref(flag), self.stream()), "fx_adjoint_overflow")
        return bool(flag.value)

    def splat_raw(self, x):
        s = self.surface
        field = torch.empty(s.sample_count, dtype=self.dtype, device=self.device)
        _lib.check(_lib.lib().fx_adjoint_splat(
            ctypes.byref(self.params), ctypes.byref(s.struct()), self.precision, _p(x), _p(self.sample_mask),
            _p(field), _p(self.work), ctypes.c_size_t(self.work.numel()), self.stream()), "fx_adjoint_splat")
        return field

    def splat_backward_raw(self, x, gfield):
        gx = torch.zeros_like(x)
        s = self.surface
        _lib.check(_lib.lib().fx_adjoint_splat_backward(
            ctypes.byref(self.params), ctypes.byref(s.struct()), self.precision, _p(x), _p(self.sample_mask),
            _p(gfield.contiguous()), _p(gx), self.stream()), "fx_adjoint_splat_backward")
        return gx


class NonFiniteError(RuntimeError):
    """A differentiable rollout produced a non-finite state, loss or gradient."""


def _all_finite(*ts):
    """Device-side: whether every element of every tensor is finite (no sync)."""
    ok = torch.ones((), dtype=torch.bool, device=ts[0].device)
    for t in ts:
        ok = ok & torch.isfinite(t).all()
    return ok


class Substep(torch.autograd.Function):
    """(x, v, gravity, translation) -> (x', v'); x, v: N x 4 tensors.
    decisions/mode: see DiffSolver.forward_raw; backward replays them.
    first: this is the first substep of a rollout, so its backward is the
    last of the chain - it raises NonFiniteError if any backward substep of
    the chain produced a non-finite adjoint (flagged on the device as they
    run, read back once here)."""

    @staticmethod
    def forward(ctx, x, v, gravity, translation, solver, decisions=None, mode=0, first=False):
        g = float(gravity)
        tr = [float(t) for t in translation]
        x_out, v_out = solver.forward_raw(x.contiguous(), v.contiguous(), g, tr, decisions, mode)
        ctx.save_for_backward(x, v)  # the substep's checkpoint
        ctx.solver, ctx.g, ctx.tr, ctx.dec, ctx.first = solver, g, tr, decisions, first
        ctx.gdtype = gravity.dtype, translation.dtype
        return x_out, v_out

    @staticmethod
    def backward(ctx, gx_out, gv_out):
        x, v = ctx.saved_tensors
        if gx_out is None:
            gx_out = torch.zeros_like(x)
        if gv_out is None:
            gv_out = torch.zeros_like(v)
        solver = ctx.solver
        gx, gv, gp = solver.backward_raw(x, v, ctx.g, ctx.tr, gx_out, gv_out, ctx.dec)
        solver.nonfinite |= ~_all_finite(gx, gv, gp)
        if ctx.first:
            bad = bool(solver.nonfinite.item())
            solver.nonfinite.zero_()
            if bad:
                raise NonFiniteError("the adjoint of the rollout is not finite (the substep map amplified the "
                                     "output adjoint past the floating-point range): shorten the horizon")
        return gx, gv, gp[0].to(ctx.gdtype[0]), gp[1:4].to(ctx.gdtype[1]), None, None, None, None


class Splat(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, solver):
        ctx.save_for_backward(x)
        ctx.solver = solver
        return solver.splat_raw(x.contiguous())

    @staticmethod
    def backward(ctx, gfield):
        (x,) = ctx.saved_tensors
        return ctx.solver.splat_backward_raw(x, gfield), None


def rollout(solver, x0, v0, gravity, translation, substeps, decisions=None, mode=0, check_overflow=True):
    """`substeps` differentiable substeps.  With a list `decisions` and mode 1
    the per-substep records (branches and neighbour lists) are written into
    the list: entry k is reused when present (records allocated ahead, e.g.
    with solver.new_decisions(), keep allocation out of the loop), else a new
    record is appended.  Mode 2 replays a list recorded earlier.  Raises if a
    neighbour list overflowed its capacity, since the gradient of a truncated
    list is not that of the substep, and raises NonFiniteError if the final
    state is not finite.  The backward of the rollout raises NonFiniteError
    when any substep's adjoint is not finite."""
    x, v = x0, v0
    solver.nonfinite.zero_()
    for k in range(substeps):
        dec = None
        if decisions is not None:
            if mode == 1:
                if k < len(decisions):
                    dec = decisions[k]
                else:
                    dec = solver.new_decisions()
                    decisions.append(dec)
            elif mode == 2:
                dec = decisions[k]
        x, v = Substep.apply(x, v, gravity, translation, solver, dec, mode, k == 0)
    if check_overflow and solver.overflowed():
        raise RuntimeError("adjoint neighbour lists overflowed their capacity (local density too high)")
    if substeps > 0 and not bool(_all_finite(x, v).item()):
        raise NonFiniteError(f"the rollout's state is not finite after {substeps} substeps")
    return x, v


def gaussian_target(solver, centre, sigma=0.2, amplitude=1.0):
    """Target density grid: a Gaussian blob of width sigma at `centre`."""
    s = solver.surface
    nx, ny, nz = s.dims
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = [s.lo[a] + [i, j, k][a] * s.spacing for a in range(3)]
    r2 = sum((p[a] - centre[a]) ** 2 for a in range(3))
    t = amplitude * np.exp(-r2 / (2 * sigma**2))
    return torch.as_tensor(t.reshape(-1), dtype=solver.dtype, device=solver.device)


def loss_fn(solver, x, target, com_target, com_weight=1.0):
    """MSE between the splat density grid and the target, plus a
    centre-of-mass term."""
    field = Splat.apply(x, solver)
    com = x[:, :3].mean(0)
    loss = ((field - target) ** 2).mean() + com_weight * ((com - com_target) ** 2).sum()
    if not bool(torch.isfinite(loss).item()):
        raise NonFiniteError("the loss is not finite")
    return loss
