"""Parity on the benchmark workloads themselves, against the oracle.

* The 1.64M-particle scale scene (cohesion, XSPH 0.01, pillar + half-wall on
  80^3 voxels), from an oracle-developed state: grid keys, sorted order and
  cell ranges bit-exact; predict and finalize bit-exact; lambda, delta,
  XSPH and normal within 1e-5; one fused substep within 1e-5 h or, where
  the oracle's own float32 result moves further when only the order of its
  sums or its precision changes, within that distance; three frames of it
  against the oracle by density-error, height and speed statistics; the GPU
  cylinder + box voxelizer bit-equal to the voxel-centre test.
* The 4.1M-particle animated scene: the sphere voxelizer at 111^3 over the
  drop frames, the grid build and a lambda pass against the oracle, one
  fused substep within 1e-5 h.
* The 64-scene batch set (seed 0): two batched substeps equal per-scene
  plain runs bit for bit.
* The reference's own smoke-test scene (scripts/smoke_test.py: default
  domain at the origin, reference caps, a 0.5 m cube collider) under the
  reference's literal update and under the converging update.

Bars as in test_gpu_parity.py; where float32 rounding of the oracle itself
is above a bar, the test states the float64 error instead and requires the
GPU to be no further from the float64 oracle than the float32 oracle is.
"""

import numpy as np
import pytest

from helpers import GpuState, rel_err, rms

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _collider(cfg):
    from flowx_amd.collision import grid_geometry, occupancy_cpu

    origin, voxel, dims = grid_geometry(cfg.lo, cfg.hi, cfg.extra["collider_resolution"])
    col = occupancy_cpu(origin, voxel, dims, cfg.extra["colliders"]).reshape(-1)
    cfg.collider_dims, cfg.collider_voxel = dims, voxel
    return col, (origin, voxel, dims)


def _stable_grid(oracle, p, pred, n):
    keys, idx, start, end = oracle.build_grid(p, pred, "stable")
    return keys[:n], idx[:n], start, end


@pytest.fixture(scope="module")
def scale(fx, oracle):
    """The scale scene advanced two substeps by the float32 oracle (stable
    order): a developed state with lambda, normals and a previous grid."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    cfg = fx.scale_config()
    col, geo = _collider(cfg)
    _, dt = cfg.substep_dt(1.0 / 24.0)
    p = cfg.params(dt)
    st = oracle.State(p, fx.seed_positions(cfg), precision="f32", sort_mode="stable")
    for _ in range(2):
        st.substep(col)
    return dict(cfg=cfg, p=p, dt=dt, col=col, geo=geo, st=st, n=cfg.particle_count)


def test_scale_collider_voxelizer_bit_exact(fx, scale):
    """Cylinder (r 0.3 m) + half-wall box on the 80^3 grid: the GPU
    voxelizer against the voxel-centre test (collision/__init__.py:353-366)."""
    from flowx_amd.collision import ColliderGrid

    cfg = scale["cfg"]
    g = ColliderGrid(cfg.lo, cfg.hi, cfg.extra["collider_resolution"], "cuda")
    g.rebuild(cfg.extra["colliders"])
    got = g.grid.cpu().numpy()
    assert got.sum() > 0 and np.array_equal(got, scale["col"])


def test_scale_grid_bit_exact(fx, oracle, scale):
    p, st, n = scale["p"], scale["st"], scale["n"]
    keys, idx, start, end = _stable_grid(oracle, p, st.pred, n)
    g = GpuState(fx, p, pred=st.pred, pos=st.pos)
    g.run("build_grid")
    assert np.array_equal(g.get("sorted_key"), keys)
    assert np.array_equal(g.get("sorted_idx"), idx)
    assert np.array_equal(g.get("cell_start"), start)
    assert np.array_equal(g.get("cell_end"), end)
    k_ref, _ = oracle.grid_key(p, st.pred, n)
    assert np.array_equal(g.get("key"), k_ref)


def test_scale_predict_and_finalize_bit_exact(fx, oracle, scale):
    """Predict with cohesion (the oracle's normals of this state) and the
    finalize against the pillar and the half-wall: bit for bit."""
    p, st, n, col = scale["p"], scale["st"], scale["n"], scale["col"]
    _, idx, start, end = _stable_grid(oracle, p, st.pred, n)
    normal = oracle.normal(p, st.pred, st.lam, idx, start, end)
    vel_ref, pred_ref = oracle.predict(p, st.pos, st.vel, normal)
    g = GpuState(fx, p, pos=st.pos, vel=st.vel, normal=normal)
    g.run("predict")
    assert np.array_equal(g.get("vel"), vel_ref) and np.array_equal(g.get("pred"), pred_ref)
    # finalize: the oracle's own last correction of this state, plus particles
    # pushed into the pillar and the wall
    pred = st.pred.copy()
    rng = np.random.default_rng(5)
    k = rng.choice(n, 20000, replace=False)
    pred[k[:10000], :3] = np.c_[rng.uniform(0.15, 0.65, (10000, 2)), rng.uniform(0.0, 1.0, 10000)]
    pred[k[10000:], :3] = np.c_[rng.uniform(-1.0, 0.0, 10000), rng.uniform(-0.45, -0.35, 10000),
                                rng.uniform(0.0, 0.5, 10000)]
    pos_ref, vel_fin = oracle.finalize(p, st.vel, st.delta, pred, col)
    assert np.abs(pos_ref[k, :3] - pred[k, :3]).max() > 0.01  # pushed out
    g = GpuState(fx, p, collider=col, pos=st.pos, vel=st.vel, pred=pred, delta=st.delta)
    g.run("finalize")
    assert np.array_equal(g.get("pos"), pos_ref) and np.array_equal(g.get("vel"), vel_fin)


def _report(name, out, ref, ref64, scale_floor):
    e = rel_err(out, ref, scale_floor)
    e64_gpu = rel_err(out, ref64, scale_floor)
    e64_orc = rel_err(ref, ref64, scale_floor)
    print(f"{name}: err {e:.2e} vs oracle f32; vs f64: gpu {e64_gpu:.2e}, oracle f32 {e64_orc:.2e}")
    return e, e64_gpu, e64_orc


def test_scale_neighbour_passes(fx, oracle, scale):
    """lambda (density bit-exact), delta, XSPH and normal on 1.64M
    particles, 1e-5 relative with an rms scale floor.  Curvature is the
    ratio -lap / |n| and |n| is a sum that cancels inside the fluid, so its
    float32 value carries the cancellation's rounding: it is held to the
    float64 oracle, no further than twice the float32 oracle's own error
    (or 1e-5)."""
    p, st, n = scale["p"], scale["st"], scale["n"]
    keys, idx, start, end = _stable_grid(oracle, p, st.pred, n)
    lam = oracle.lambda_(p, st.pred, idx, start, end)
    lam64 = oracle.lambda_(p, st.pred, idx, start, end, precision="f64")
    g = GpuState(fx, p, pred=st.pred, sorted_key=keys, sorted_idx=idx)
    g.run("lambda")
    out = g.get("lam")
    assert np.array_equal(out[:, 0], lam[:, 0])
    e, _, _ = _report("lambda", out[:, 1], lam[:, 1], lam64[:, 1], rms(lam64[:, 1]))
    assert e <= TOL

    d = oracle.delta(p, st.pred, lam, idx, start, end)
    d64 = oracle.delta(p, st.pred, lam, idx, start, end, precision="f64")
    g = GpuState(fx, p, pred=st.pred, lam=lam, sorted_key=keys, sorted_idx=idx)
    g.run("delta")
    e, _, _ = _report("delta", g.get("delta")[:, :3], d[:, :3], d64[:, :3], rms(d64[:, :3]))
    assert e <= TOL

    x = oracle.xsph(p, st.pred, st.vel, st.lam, idx, start, end)
    x64 = oracle.xsph(p, st.pred, st.vel, st.lam, idx, start, end, precision="f64")
    g = GpuState(fx, p, pred=st.pred, vel=st.vel, lam=st.lam, sorted_key=keys, sorted_idx=idx)
    g.run("xsph")
    e, _, _ = _report("xsph", g.get("delta")[:, :3], x[:, :3], x64[:, :3], max(rms(x64[:, :3]), 1e-12))
    assert e <= TOL

    nm = oracle.normal(p, st.pred, st.lam, idx, start, end)
    nm64 = oracle.normal(p, st.pred, st.lam, idx, start, end, precision="f64")
    g = GpuState(fx, p, pred=st.pred, lam=st.lam, sorted_key=keys, sorted_idx=idx)
    g.run("normal")
    out = g.get("normal")
    e, _, _ = _report("normal", out[:, :3], nm[:, :3], nm64[:, :3], rms(nm64[:, :3]))
    assert e <= TOL
    _, k_gpu, k_orc = _report("curvature", out[:, 3], nm[:, 3], nm64[:, 3], rms(nm64[:, 3]))
    assert k_gpu <= max(TOL, 2 * k_orc)


@pytest.mark.parametrize("sigma", [0.0, 0.001])
def test_scale_fused_substep(fx, oracle, scale, sigma):
    """One fused substep (fx_substep: normal, predict, grid, 4 x (lambda,
    delta+apply), velocity, XSPH, finalize) from the developed state against
    the oracle's pass-by-pass substep, the grid and sorted order identical.
    Without cohesion: positions within 1e-5 h, densities within 1e-5, or
    within the distance between the oracle run in the reference's bitonic
    order and in stable order (at 1.64M particles four Jacobi iterations
    carry a few particles' rounding past 1e-5 h, and reordering the same
    float32 sums moves them further: 4.9e-5 h against the GPU's 2.7e-5).  With
    cohesion (sigma 0.001, the benchmark's), after a state load: the normal
    pass then divides by densities clamped to 1e-6 (lambda_img is zeroed by
    the load, solver/sph.py:716-724), and predict's cohesion kick takes the
    direction n / |n| of a sum that cancels inside the fluid
    (shaders/sph_predict.glsl:33-46), so float32 rounding picks that
    direction for interior particles - the float32 oracle itself is then far
    from the float64 one, and the GPU is held to no more than twice the
    float32 oracle's distance from float64."""
    cfg, p0, st, col, n = scale["cfg"], scale["p"], scale["st"], scale["col"], scale["n"]
    p = type(p0).from_buffer_copy(p0)
    p.surface_tension = sigma
    ref = oracle.State(p, st.pos, st.vel, precision="f32", sort_mode="stable")
    g = GpuState(fx, p, collider=col, pos=st.pos, vel=st.vel)
    g.run("load_state")
    # the oracle state's normal pass reads the previous grid: the one built
    # from pred = pos by the load, as in the GPU state
    ref.substep(col)
    g.run("substep")
    h = cfg.smoothing_radius
    gp = g.get("pos")[:, :3]
    e_pos = rel_err(gp, ref.pos[:, :3], h)
    e_vel = rel_err(g.get("vel")[:, :3], ref.vel[:, :3], h / scale["dt"])
    e_rho = rel_err(g.get("lam")[:, 0], ref.lam[:, 0], p.rest_density)
    far = int(np.count_nonzero(np.abs(gp - ref.pos[:, :3]).max(1) > TOL * h))
    print(f"scale substep, sigma {sigma}: pos err {e_pos:.2e} h-relative ({far} of {n} particles beyond 1e-5 h), "
          f"vel {e_vel:.2e}, rho {e_rho:.2e}")
    assert np.array_equal(g.get("sorted_idx"), ref.idx[:n])
    ref64 = oracle.State(p, st.pos, st.vel, precision="f64", sort_mode="stable")
    ref64.substep(col)
    e64_gpu = rel_err(gp, ref64.pos[:, :3], h)
    e64_orc = rel_err(ref.pos[:, :3], ref64.pos[:, :3], h)
    far64 = int(np.count_nonzero(np.abs(ref.pos[:, :3] - ref64.pos[:, :3]).max(1) > TOL * h))
    print(f"  vs f64: gpu {e64_gpu:.2e}, oracle f32 {e64_orc:.2e} ({far64} oracle f32 particles beyond 1e-5 h)")
    # the reference's own summation order (its bitonic network) against the
    # stable order: how far the same float32 passes land when only the order
    # of each neighbour sum changes
    refb = oracle.State(p, st.pos, st.vel, precision="f32", sort_mode="bitonic")
    refb.substep(col)
    e_order = rel_err(refb.pos[:, :3], ref.pos[:, :3], h)
    e_vel_order = rel_err(refb.vel[:, :3], ref.vel[:, :3], h / scale["dt"])
    e_rho_order = rel_err(refb.lam[:, 0], ref.lam[:, 0], p.rest_density)
    print(f"  oracle bitonic order vs stable order: pos {e_order:.2e} h, vel {e_vel_order:.2e}, rho {e_rho_order:.2e}")
    if sigma == 0.0:
        assert e_pos <= max(TOL, e_order) and e_rho <= max(TOL, e_rho_order)
        assert e_vel <= max(10 * TOL, e_vel_order)
    else:
        assert e_pos <= TOL or e64_gpu <= 2 * e64_orc


def test_scale_run_statistics(fx, oracle, scale):
    """Three frames (12 substeps, cohesion on, one fx_step call: the resident
    hand-off between substeps) of the 1.64M scene from the developed state,
    against the float32 oracle in the reference's bitonic order: the run is
    chaotic, so it is compared as long runs are, by the density-constraint
    error and fluid-height statistics - mean density and
    mean |rho / rho0 - 1| within 5e-3, mean and 90th-percentile height
    within 0.02 / 0.05 h, rms speed within 1%."""
    import torch

    cfg, p, st, col = scale["cfg"], scale["p"], scale["st"], scale["col"]
    steps = 12
    ref = oracle.State(p, st.pos, st.vel, precision="f32", sort_mode="bitonic")
    for _ in range(steps):
        ref.substep(col)
    solver = fx.Solver(cfg, keep_scratch=True)
    solver.reset(st.pos)
    solver.load_state(st.pos, st.vel)
    solver.substeps_n(scale["dt"], steps)
    torch.cuda.synchronize()
    gp, gv, gl = solver.read_positions(), solver.read_velocities(), solver.lam.cpu().numpy()
    h, rho0 = cfg.smoothing_radius, p.rest_density
    d_rho = abs(gl[:, 0].mean() - ref.lam[:, 0].mean()) / rho0
    d_err = abs(np.abs(gl[:, 0] / rho0 - 1).mean() - np.abs(ref.lam[:, 0] / rho0 - 1).mean())
    d_zmean = abs(gp[:, 2].mean() - ref.pos[:, 2].mean()) / h
    d_z90 = abs(np.percentile(gp[:, 2], 90) - np.percentile(ref.pos[:, 2], 90)) / h
    v_gpu, v_ref = rms(gv[:, :3]), rms(ref.vel[:, :3])
    print(f"scale, {steps} substeps: mean rho {gl[:, 0].mean():.2f} / {ref.lam[:, 0].mean():.2f}, |C| diff "
          f"{d_err:.2e}, mean height diff {d_zmean:.2e} h, p90 height diff {d_z90:.2e} h, rms speed "
          f"{v_gpu:.4f} / {v_ref:.4f}")
    assert d_rho < 5e-3 and d_err < 5e-3 and d_zmean < 2e-2 and d_z90 < 5e-2
    assert abs(v_gpu - v_ref) <= 0.01 * v_ref


# ---- the animated scene ------------------------------------------------------------


def test_animated_sphere_voxelizer_over_the_drop(fx):
    """The 0.3 m sphere dropped from z = 1.8 m, re-voxelized on the 111^3
    grid at every frame of its fall: bit-equal to the voxel-centre test."""
    from flowx_amd.collision import ColliderGrid, occupancy_cpu

    cfg = fx.animated_config()
    (c0, r) = cfg.extra["sphere"]
    g = ColliderGrid(cfg.lo, cfg.hi, cfg.extra["collider_resolution"], "cuda")
    assert tuple(g.dims) == (111, 111, 111)
    for frame in (0, 5, 10, 13, 15, 20):
        t = frame / cfg.extra["fps"]
        shape = ("sphere", (c0[0], c0[1], max(c0[2] - 0.5 * 9.81 * t * t, r)), r)
        g.rebuild([shape])
        ref = occupancy_cpu(g.origin, g.voxel, g.dims, [shape]).reshape(-1)
        assert ref.sum() > 0 and np.array_equal(g.grid.cpu().numpy(), ref), frame


def test_animated_grid_lambda_and_substep(fx, oracle):
    """4.1M particles, 6 iterations, the sphere at its start height: one
    oracle substep from the seeded lattice; on its state the GPU grid build
    is bit-exact and the lambda pass within 1e-5; the GPU's own fused
    substep from the same seed is within 1e-5 h of the oracle's."""
    from flowx_amd.collision import ColliderGrid

    cfg = fx.animated_config()
    (c0, r) = cfg.extra["sphere"]
    cg = ColliderGrid(cfg.lo, cfg.hi, cfg.extra["collider_resolution"], "cuda")
    cg.rebuild([("sphere", c0, r)])
    cfg.collider_dims, cfg.collider_voxel = cg.dims, cg.voxel
    col = cg.grid.cpu().numpy()
    _, dt = cfg.substep_dt(1.0 / 24.0)
    p = cfg.params(dt)
    pos0 = fx.seed_positions(cfg)
    n = len(pos0)
    st = oracle.State(p, pos0, precision="f32", sort_mode="stable")
    st.substep(col)
    keys, idx, start, end = _stable_grid(oracle, p, st.pred, n)
    g = GpuState(fx, p, pred=st.pred, pos=st.pos)
    g.run("build_grid")
    assert np.array_equal(g.get("sorted_key"), keys) and np.array_equal(g.get("sorted_idx"), idx)
    assert np.array_equal(g.get("cell_start"), start) and np.array_equal(g.get("cell_end"), end)
    del g
    lam = oracle.lambda_(p, st.pred, idx, start, end)
    g = GpuState(fx, p, pred=st.pred, sorted_key=keys, sorted_idx=idx)
    g.run("lambda")
    out = g.get("lam")
    assert np.array_equal(out[:, 0], lam[:, 0])
    assert rel_err(out[:, 1], lam[:, 1], rms(lam[:, 1])) <= TOL
    del g
    g = GpuState(fx, p, collider=col, pos=pos0)
    g.run("load_state")
    g.run("substep")
    e_pos = rel_err(g.get("pos")[:, :3], st.pos[:, :3], cfg.smoothing_radius)
    print(f"animated substep: pos err {e_pos:.2e} h-relative")
    assert e_pos <= TOL


# ---- the batch set ------------------------------------------------------------------


def test_full_batch_set_equals_plain_runs(fx):
    """The 64 scenes of the benchmark set (seed 0; 10k-400k particles, 0-3
    boxes each), two batched substeps: every scene bit-identical to its own
    plain run."""
    import torch

    cfgs = fx.batch_configs(count=64, seed=0)
    batch = fx.BatchSolver(cfgs)
    for _ in range(2):
        batch.substep()
    torch.cuda.synchronize()
    pos = batch.pos.cpu().numpy()
    vel = batch.vel.cpu().numpy()
    offsets = batch.offsets
    del batch
    torch.cuda.empty_cache()
    for i, cfg in enumerate(fx.batch_configs(count=64, seed=0)):
        s = fx.Solver(cfg, keep_scratch=False)
        s.reset()
        _, dt = cfg.substep_dt(1.0 / 24.0)
        s.substeps_n(dt, 2)
        a, b = offsets[i], offsets[i + 1]
        assert np.array_equal(pos[a:b], s.read_positions()), f"scene {i} positions"
        assert np.array_equal(vel[a:b], s.read_velocities()), f"scene {i} velocities"
        del s


# ---- the reference's own smoke scene --------------------------------------------------


def _smoke_config(fx, literal):
    """scripts/smoke_test.py: flowx.domain_add at the cursor (the origin: a 2 m
    cube, domain/__init__.py:280-309), default settings, a 0.5 m cube
    collider at the origin (mesh.primitive_cube_add size 0.5), the
    reference's caps: 15,360 particles after coarsening (h = 0.1187 m)."""
    from flowx_amd.config import REFERENCE_CAPS, DomainSettings, converging_update, resolve_config

    settings = DomainSettings()
    cfg = resolve_config(settings, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), REFERENCE_CAPS)
    cfg.extra.update(colliders=[("box", (-0.25, -0.25, -0.25), (0.25, 0.25, 0.25))],
                     collider_resolution=settings.resolution)
    return cfg if literal else converging_update(cfg)


def test_smoke_scene_density_check(fx, oracle):
    """The smoke test's PBF check (scripts/smoke_test.py:107-134): mean
    lambda-pass density within 15% of rest after 3 frames.  It holds under
    the converging update.  Under the reference's literal update (no mass in
    the correction, s_corr 0.1) the scene blows apart in its first frame - on
    the GPU and in the oracle alike (float32 and float64) - so the check
    fails there; the test pins that both agree on it."""
    import torch

    res = {}
    for literal in (False, True):
        cfg = _smoke_config(fx, literal)
        solver = fx.Solver(cfg)
        solver.reset()
        assert cfg.particle_count == 15360
        for _ in range(3):
            solver.step(1.0 / 24.0)
        rho = solver.read_density(check=False)
        gpu_err = abs(rho.mean() - cfg.rest_density) / cfg.rest_density
        col, _ = _collider(cfg)
        steps, dt = cfg.substep_dt(1.0 / 24.0)
        st = oracle.State(cfg.params(dt), fx.seed_positions(cfg), precision="f32", sort_mode="bitonic")
        for _ in range(3 * steps):
            st.substep(col)
        orc_err = abs(st.lam[:, 0].mean() - cfg.rest_density) / cfg.rest_density
        vmax = float(np.abs(solver.read_velocities(check=False)[:, :3]).max())
        print(f"smoke scene, {'literal' if literal else 'converging'} update: mean density error "
              f"GPU {gpu_err * 100:.2f}%, oracle {orc_err * 100:.2f}%; max |v| {vmax:.3g} m/s")
        res[literal] = (gpu_err, orc_err)
        del solver
        torch.cuda.empty_cache()
    assert res[False][0] <= 0.15 and res[False][1] <= 0.15
    assert res[True][0] > 0.15 and res[True][1] > 0.15


@pytest.mark.parametrize("scene", ["dam_break", "smoke"])
def test_literal_update_four_iterations(fx, oracle, scene):
    """The reference's literal pass semantics (flags 0, s_corr 0.1) with its
    4 Jacobi iterations: one fused substep from the seeded lattice against
    the oracle's.  The literal update overshoots the Newton step by 1/m, so
    each iteration amplifies the previous one's float32 rounding; the bar is
    1e-5 h or, where the float32 oracle is itself further than that from the
    float64 oracle, no further from float64 than twice the float32 oracle."""
    if scene == "dam_break":
        cfg = fx.dam_break_config(literal_update=True)
    else:
        cfg = _smoke_config(fx, True)
    assert cfg.flags == 0 and cfg.scorr_strength == 0.1 and cfg.iterations == 4
    col, _ = _collider(cfg)
    pos0 = fx.seed_positions(cfg)
    _, dt = cfg.substep_dt(1.0 / 24.0)
    p = cfg.params(dt)
    ref = oracle.State(p, pos0, precision="f32", sort_mode="stable")
    ref64 = oracle.State(p, pos0, precision="f64", sort_mode="stable")
    ref.substep(col)
    ref64.substep(col)
    g = GpuState(fx, p, collider=col, pos=pos0)
    g.run("load_state")
    g.run("substep")
    h = cfg.smoothing_radius
    e = rel_err(g.get("pos")[:, :3], ref.pos[:, :3], h)
    e64_gpu = rel_err(g.get("pos")[:, :3], ref64.pos[:, :3], h)
    e64_orc = rel_err(ref.pos[:, :3], ref64.pos[:, :3], h)
    e_rho = rel_err(g.get("lam")[:, 0], ref.lam[:, 0], p.rest_density)
    print(f"literal update, {scene}, 4 iterations: pos err {e:.2e} h vs oracle f32; vs f64: gpu {e64_gpu:.2e}, "
          f"oracle f32 {e64_orc:.2e}; rho err {e_rho:.2e}; oracle mean rho {ref.lam[:, 0].mean():.1f}")
    assert e <= TOL or e64_gpu <= 2 * e64_orc


@pytest.mark.parametrize("scene", ["dam_break", "smoke"])
def test_literal_update_every_pass_of_a_substep(fx, oracle, scene):
    """The literal substep's passes one at a time: the oracle runs the
    substep pass by pass (predict, grid, 4 x (lambda, delta, apply),
    velocity, XSPH, finalize) and every GPU pass, given that pass's oracle
    inputs, is within 1e-5 of the oracle's output (density bit-exact,
    streaming passes bit-exact) - each of the four Jacobi iterations
    included, although their composition amplifies rounding."""
    if scene == "dam_break":
        cfg = fx.dam_break_config(literal_update=True)
    else:
        cfg = _smoke_config(fx, True)
    col, _ = _collider(cfg)
    pos0 = fx.seed_positions(cfg)
    n = len(pos0)
    _, dt = cfg.substep_dt(1.0 / 24.0)
    p = cfg.params(dt)
    vel0 = np.zeros_like(pos0)
    zero = np.zeros_like(pos0)
    vel_p, pred = oracle.predict(p, pos0, vel0, zero)
    g = GpuState(fx, p, pos=pos0, vel=vel0)
    g.run("predict")
    assert np.array_equal(g.get("pred"), pred) and np.array_equal(g.get("vel"), vel_p)
    keys, idx, start, end = _stable_grid(oracle, p, pred, n)
    lam = None
    for it in range(cfg.iterations):
        lam = oracle.lambda_(p, pred, idx, start, end)
        g = GpuState(fx, p, pred=pred, sorted_key=keys, sorted_idx=idx)
        g.run("lambda")
        out = g.get("lam")
        assert np.array_equal(out[:, 0], lam[:, 0]), it
        e_l = rel_err(out[:, 1], lam[:, 1], rms(lam[:, 1]))
        d = oracle.delta(p, pred, lam, idx, start, end)
        g = GpuState(fx, p, pred=pred, lam=lam, sorted_key=keys, sorted_idx=idx)
        g.run("delta")
        e_d = rel_err(g.get("delta")[:, :3], d[:, :3], rms(d[:, :3]))
        nxt = oracle.apply_delta(p, pred, d)
        g = GpuState(fx, p, pred=pred, delta=d)
        g.run("apply_delta")
        assert np.array_equal(g.get("pred"), nxt), it
        print(f"literal {scene}, iteration {it}: lambda err {e_l:.2e}, delta err {e_d:.2e} "
              f"(rms correction {rms(d[:, :3]) / cfg.smoothing_radius:.2e} h)")
        assert e_l <= TOL and e_d <= TOL
        pred = nxt
    v = oracle.velocity(p, pred, pos0)
    g = GpuState(fx, p, pred=pred, pos=pos0)
    g.run("velocity")
    assert np.array_equal(g.get("vel"), v)
    x = oracle.xsph(p, pred, v, lam, idx, start, end)
    g = GpuState(fx, p, pred=pred, vel=v, lam=lam, sorted_key=keys, sorted_idx=idx)
    g.run("xsph")
    assert rel_err(g.get("delta")[:, :3], x[:, :3], max(rms(x[:, :3]), 1e-12)) <= TOL
    pos_ref, vel_ref = oracle.finalize(p, v, x, pred, col)
    g = GpuState(fx, p, collider=col, pos=pos0, vel=v, pred=pred, delta=x)
    g.run("finalize")
    assert np.array_equal(g.get("pos"), pos_ref) and np.array_equal(g.get("vel"), vel_ref)
