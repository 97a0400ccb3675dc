"""What a workspace holds is recorded in the workspace itself (its header:
grid built for (n, n_cells), normals computed for the next substep), never
in host state keyed by its address.  The normal pass reads the previous
grid (shaders/sph_normal.glsl:15-21) and the splat gathers through it
(shaders/surface_splat.glsl:8-9), so a workspace that never held a grid must
be refused - also when it sits at the address of one that did - and a
captured substep must decide afresh, on every replay, whether the normals
are already there."""

import ctypes

import numpy as np
import pytest

from helpers import GpuState

pytestmark = pytest.mark.gpu


def _cohesion_params(fx):
    cfg = fx.dam_break_config()
    cfg.surface_tension = 0.001
    from flowx_amd.collision import grid_geometry, occupancy_cpu

    origin, voxel, dims = grid_geometry(cfg.lo, cfg.hi, cfg.extra["collider_resolution"])
    col = occupancy_cpu(origin, voxel, dims, cfg.extra["colliders"]).reshape(-1)
    cfg.collider_dims, cfg.collider_voxel = dims, voxel
    _, dt = cfg.substep_dt(1.0 / 24.0)
    return cfg, cfg.params(dt), col, dt


def _splat(fx, g, p, cfg):
    import torch

    surf = fx.resolve(cfg, 32, 0.5)
    field = torch.zeros(surf.sample_count, dtype=torch.float32, device="cuda")
    rc = fx._lib.lib().fx_splat(ctypes.byref(p), ctypes.byref(surf.struct()), ctypes.byref(g.state),
                                ctypes.c_void_p(field.data_ptr()), g.stream())
    fx._lib.check(rc, "fx_splat")
    return field


def test_workspace_without_a_grid_is_refused(fx):
    """A zeroed workspace and one full of garbage: a cohesion substep and a
    splat return an error naming fx_load_state instead of reading offsets."""
    import torch

    cfg, p, col, _ = _cohesion_params(fx)
    pos = fx.seed_positions(cfg)
    for fill in ("zeros", "garbage"):
        g = GpuState(fx, p, collider=col, pos=pos)
        if fill == "garbage":
            g.work.copy_(torch.randint(0, 256, g.work.shape, dtype=torch.uint8, device="cuda"))
        with pytest.raises(RuntimeError, match="fx_load_state"):
            g.run("substep")
        with pytest.raises(RuntimeError, match="fx_load_state"):
            _splat(fx, g, p, cfg)
        # loading the state makes the same workspace usable
        g.run("load_state")
        g.run("substep")
        assert np.all(np.isfinite(g.get("pos")))


def test_reused_address_does_not_inherit_the_grid(fx):
    """A workspace that held a grid and fused normals is freed; the caching
    allocator hands its address to a new, zeroed workspace.  Stepping the new
    one without fx_load_state is refused: nothing about the old one is kept
    by address."""
    import torch

    cfg, p, col, _ = _cohesion_params(fx)
    pos = fx.seed_positions(cfg)
    a = GpuState(fx, p, collider=col, pos=pos)
    a.run("load_state")
    for _ in range(2):
        a.run("substep")  # the second substep took its normals from the first's XSPH pass
    addr = a.work.data_ptr()
    del a
    b = GpuState(fx, p, collider=col, pos=pos)
    print(f"reused address: {b.work.data_ptr() == addr}")
    with pytest.raises(RuntimeError, match="fx_load_state"):
        b.run("substep")
    # with an explicit step count too
    L = fx._lib
    rc = L.lib().fx_step(ctypes.byref(p), ctypes.byref(b.state), 3, b.stream())
    assert rc != 0 and "fx_load_state" in L.lib().fx_last_error().decode()
    torch.cuda.synchronize()


def test_normals_follow_the_workspace_not_the_call(fx):
    """The fused normals are taken only for the normal buffer and constants
    they were made for: with another normal buffer (a second FxState over the
    same workspace) the normal pass runs, and its result equals the fused
    normals to the fused-normals bar (1e-6 relative)."""
    from helpers import rel_err, rms

    cfg, p, col, _ = _cohesion_params(fx)
    pos = fx.seed_positions(cfg)
    a = GpuState(fx, p, collider=col, pos=pos)
    a.run("load_state")
    for _ in range(3):
        a.run("substep")
    b = GpuState(fx, p, collider=col, pos=pos)  # its own buffers, and a copy of a's workspace:
    b.work.copy_(a.work)  # the grid travels with it; the normals were made for a.normal
    b.pos.copy_(a.pos)
    b.vel.copy_(a.vel)
    a.run("substep")
    b.run("substep")
    na, nb = a.get("normal"), b.get("normal")
    assert rel_err(nb[:, :3], na[:, :3], rms(na[:, :3])) <= 1e-6
    assert rel_err(nb[:, 3], na[:, 3], rms(na[:, 3])) <= 1e-5


def test_captured_cohesion_substeps_replay_after_load_state(fx):
    """Cohesion on: capture, load the seeded state again, replay.  The first
    replayed substep must compute its normals (the load forgot the fused
    ones), later replays take them from the previous replay's XSPH pass -
    every replay equals eager substeps_n calls bit for bit, with the
    program's own eager work between replays."""
    import torch

    cfg = fx.dam_break_config()
    cfg.surface_tension = 0.001
    _, dt = cfg.substep_dt(1.0 / 24.0)
    seed = fx.seed_positions(cfg)
    eager = fx.Solver(cfg, keep_scratch=True)
    eager.reset()
    graphed = fx.Solver(cfg, keep_scratch=True)
    graphed.reset()
    graph = graphed.capture(dt, 3)  # advances graphed by a warm-up substep
    for _ in range(2):
        graph.replay()  # the captured state moves on; load_state below resets it
    graphed.load_state(seed, None)
    eager.load_state(seed, None)
    for r in range(3):
        graph.replay()
        eager.substeps_n(dt, 3)
        torch.cuda.synchronize()
        _ = eager.pos.sum().item()
        for name in ("pos", "vel", "normal", "lam"):
            assert np.array_equal(getattr(graphed, name).cpu().numpy(), getattr(eager, name).cpu().numpy()), (r, name)
