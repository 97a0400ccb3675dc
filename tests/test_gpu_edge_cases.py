"""Edge cases of the reference's passes on the GPU, against the oracle: a cell
holding more than MAX_CELL_SCAN (512) particles (shaders/sph_lambda.glsl:
49-56 stops each cell's scan after 512 slots), predictions outside the
domain (cell_coord clamps them into the edge cells,
shaders/sph_common.glsl:50-54), and scenes of one and of zero particles."""

import numpy as np
import pytest

from helpers import GpuState, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _params(fx, n, lattice=(12, 12, 6)):
    cfg = fx.lattice_config((-0.5, -0.5, 0.0), (0.5, 0.5, 1.0), 0.1, lattice, iterations=4)
    from flowx_amd.config import converging_update

    cfg = converging_update(cfg)
    cfg.particle_count = n
    _, dt = cfg.substep_dt(1.0 / 24.0)
    return cfg, cfg.params(dt), dt


def test_cell_beyond_max_cell_scan(fx, oracle):
    """700 particles packed into one 0.1 m cell beside a lattice: the dense
    cell's scan stops after 512 slots in both, with the same ascending-index
    order, so density is bit-exact and lambda within the bar."""
    rng = np.random.default_rng(11)
    cfg, _, _ = _params(fx, 0)
    lattice = fx.seed_positions(cfg)
    clump = np.concatenate([rng.uniform([0.01, 0.01, 0.41], [0.09, 0.09, 0.49], (700, 3)),
                            np.ones((700, 1))], 1).astype(np.float32)
    pos = np.concatenate([lattice, clump])
    cfg, p, dt = _params(fx, len(pos))
    keys, idx, start, end = oracle.build_grid(p, pos, "stable")
    counts = (end - start)[start >= 0]
    assert counts.max() > 512
    n = len(pos)
    ref = oracle.lambda_(p, pos, idx[:n], start, end)
    g = GpuState(fx, p, pred=pos, sorted_key=keys[:n], sorted_idx=idx[:n])
    g.run("lambda")
    out = g.get("lam")
    assert np.array_equal(out[:, 0], ref[:, 0])
    scale = float(np.sqrt(np.mean(ref[:, 1].astype(np.float64) ** 2)))
    assert rel_err(out[:, 1], ref[:, 1], scale) <= TOL
    # and one fused substep (stable order) against the oracle's.  Four Jacobi
    # iterations over 700 overlapping particles amplify float32 rounding, so
    # the bar is stated against float64: the GPU may not be further from the
    # float64 oracle than the float32 oracle is (and both stay near 1e-5 h)
    st = oracle.State(p, pos, precision="f32", sort_mode="stable")
    st64 = oracle.State(p, pos, precision="f64", sort_mode="stable")
    st.substep(None)
    st64.substep(None)
    g2 = GpuState(fx, p, pos=pos)
    g2.run("load_state")
    g2.run("substep")
    h = cfg.smoothing_radius
    e = rel_err(g2.get("pos")[:, :3], st.pos[:, :3], h)
    e_gpu64 = rel_err(g2.get("pos")[:, :3], st64.pos[:, :3], h)
    e_orc64 = rel_err(st.pos[:, :3], st64.pos[:, :3], h)
    print(f"dense cell: max cell {counts.max()}, substep pos err vs f32 oracle {e:.2e}, "
          f"vs f64: gpu {e_gpu64:.2e}, f32 oracle {e_orc64:.2e}")
    assert e_gpu64 <= max(2 * e_orc64, TOL) and e <= 1e-4


def test_cell_beyond_max_cell_scan_large_scene(fx, oracle):
    """The same 700-particle cell inside a 204,800-particle lattice, where
    the neighbour passes take the large-scene walks: a lane whose stencil row
    holds the dense cell leaves the stream / row-batch / LDS-tile walk for the
    per-cell path that stops each cell's scan after 512 slots.  The lambda
    pass against the oracle (density bit-exact, lambda within the bar), and
    two substeps bit-identical across the five walks (FxTuning.walk 0-4:
    walk 0 records lambda's pair masks for delta, and a particle whose row
    holds the dense cell is not recorded, so its delta walks the full
    stencil)."""
    import torch

    from flowx_amd import _lib
    from flowx_amd.config import converging_update

    rng = np.random.default_rng(12)
    cfg = converging_update(fx.lattice_config((-1.0, -1.0, 0.0), (1.0, 1.0, 2.0), 0.025, (80, 80, 32),
                                              iterations=4, viscosity=0.01))
    lattice = fx.seed_positions(cfg)
    clump = np.concatenate([rng.uniform([0.0026, 0.0026, 0.1026], [0.0224, 0.0224, 0.1224], (700, 3)),
                            np.ones((700, 1))], 1).astype(np.float32)
    pos = np.concatenate([lattice, clump])
    cfg.particle_count = len(pos)
    _, dt = cfg.substep_dt(1.0 / 24.0)
    p = cfg.params(dt)
    n = len(pos)
    keys, idx, start, end = oracle.build_grid(p, pos, "stable")
    counts = (end - start)[start >= 0]
    assert counts.max() > 512 and n > 1 << 17
    ref = oracle.lambda_(p, pos, idx[:n], start, end)
    g = GpuState(fx, p, pred=pos, sorted_key=keys[:n], sorted_idx=idx[:n])
    g.run("lambda")
    out = g.get("lam")
    assert np.array_equal(out[:, 0], ref[:, 0])
    scale = float(np.sqrt(np.mean(ref[:, 1].astype(np.float64) ** 2)))
    assert rel_err(out[:, 1], ref[:, 1], scale) <= TOL
    saved = _lib.get_tuning()
    res = []
    try:
        for walk in (0, 1, 2, 3, 4):
            _lib.set_tuning(walk=walk)
            s = fx.Solver(cfg, keep_scratch=True)
            s.reset(pos)
            s.substeps_n(dt, 2)
            torch.cuda.synchronize()
            res.append((s.pos.cpu().numpy(), s.vel.cpu().numpy(), s.lam.cpu().numpy()))
            del s
    finally:
        _lib.set_tuning(**saved)
    assert np.isfinite(res[0][0]).all()
    for w, r in enumerate(res[1:], 1):
        for name, a, b in zip(("pos", "vel", "lam"), res[0], r):
            assert np.array_equal(a, b), (w, name)


def test_splat_dense_cells(fx, oracle):
    """The splat over a lattice plus a 3000-particle clump spread over 2x2x2
    cells (the tiles around it stage more particles than their LDS holds and
    take the per-row path) and a 700-particle cell elsewhere (both paths stop
    each cell's scan after 512 slots): bit-exact against the oracle."""
    import ctypes

    import torch

    rng = np.random.default_rng(5)
    cfg, _, _ = _params(fx, 0)
    lattice = fx.seed_positions(cfg)
    clump = rng.uniform([-0.19, -0.19, 0.21], [-0.01, -0.01, 0.39], (3000, 3))
    cell = rng.uniform([0.21, 0.21, 0.61], [0.29, 0.29, 0.69], (700, 3))
    extra = np.concatenate([np.concatenate([clump, cell]), np.ones((3700, 1))], 1).astype(np.float32)
    pos = np.concatenate([lattice, extra])
    cfg, p, dt = _params(fx, len(pos))
    keys, idx, start, end = oracle.build_grid(p, pos, "stable")
    counts = (end - start)[start >= 0]
    assert counts.max() > 512
    n = len(pos)
    g = GpuState(fx, p, pos=pos, pred=pos)
    g.run("build_grid")
    for resolution in (48, 96):
        surf = fx.resolve(cfg, resolution, 0.5)
        ref = oracle.splat(p, surf.struct(), pos, idx[:n], start, end, None)
        field = torch.zeros(surf.sample_count, dtype=torch.float32, device="cuda")
        fx._lib.check(fx._lib.lib().fx_splat(ctypes.byref(p), ctypes.byref(surf.struct()), ctypes.byref(g.state),
                                             ctypes.c_void_p(field.data_ptr()), g.stream()), "fx_splat")
        out = field.cpu().numpy()
        print(f"dense splat {surf.dims}: max field {ref.max():.2f}, max |diff| {np.abs(out - ref).max():.3e}")
        assert np.array_equal(out, ref.astype(np.float32))


def test_out_of_domain_predictions_clamp_into_edge_cells(fx, oracle):
    rng = np.random.default_rng(12)
    n = 4000
    pos = np.concatenate([rng.uniform([-1.5, -1.5, -0.7], [1.5, 1.5, 1.8], (n, 3)), np.ones((n, 1))], 1)
    pos = pos.astype(np.float32)
    cfg, p, dt = _params(fx, n)
    keys, _ = oracle.grid_key(p, pos, n)
    g = GpuState(fx, p, pred=pos)
    g.run("grid_key")
    assert np.array_equal(g.get("key"), keys)
    okeys, idx, start, end = oracle.build_grid(p, pos, "stable")
    g.run("build_grid")
    assert np.array_equal(g.get("sorted_key"), okeys[:n])
    assert np.array_equal(g.get("sorted_idx"), idx[:n])
    assert np.array_equal(g.get("cell_start"), start) and np.array_equal(g.get("cell_end"), end)
    outside = ((pos[:, :3] < p.lo[:]) | (pos[:, :3] >= p.hi[:])).any(1)
    assert outside.sum() > 1000


def test_single_particle_free_fall_and_empty_scene(fx, oracle):
    cfg, p, dt = _params(fx, 1)
    pos = np.array([[0.05, -0.05, 0.6, 1.0]], np.float32)
    st = oracle.State(p, pos, precision="f32", sort_mode="stable")
    g = GpuState(fx, p, pos=pos)
    g.run("load_state")
    for _ in range(3):
        st.substep(None)
        g.run("substep")
    assert np.array_equal(g.get("pos"), st.pos) and np.array_equal(g.get("vel"), st.vel)
    assert g.get("vel")[0, 2] < 0  # falling
    cfg0, p0, _ = _params(fx, 0)
    g0 = GpuState(fx, p0)
    g0.run("load_state")
    g0.run("substep")  # nothing to do, and no error


@pytest.mark.parametrize("lattice", [(40, 40, 16), (70, 68, 20)])
def test_small_scene_pair_masks_bit_identical(fx, lattice):
    """Small and mid-size scenes (25,600 and 95,200 lattice particles: lambda
    records in quads, delta walks the recorded words in quads at the smaller
    size and with one lane per particle at the larger) plus a 150-particle clump
    in one cell, whose rows pass the 64 slots a word records: two substeps
    with the pair masks (walk 0) equal the walks without them (walk 4) bit
    for bit."""
    import torch

    from flowx_amd import _lib
    from flowx_amd.config import converging_update

    rng = np.random.default_rng(21)
    cfg = converging_update(fx.lattice_config((-1.0, -1.0, 0.0), (1.0, 1.0, 2.0), 0.025, lattice, iterations=4,
                                              viscosity=0.01, surface_tension=0.001))
    cfg.extra.update(colliders=[("box", (-0.2, -0.2, 0.0), (0.2, 0.2, 0.3))], collider_resolution=40)
    clump = np.concatenate([rng.uniform([0.0026, 0.0026, 0.1026], [0.0224, 0.0224, 0.1224], (150, 3)),
                            np.ones((150, 1))], 1).astype(np.float32)
    pos = np.concatenate([fx.seed_positions(cfg), clump])
    cfg.particle_count = len(pos)
    _, dt = cfg.substep_dt(1.0 / 24.0)
    saved = _lib.get_tuning()
    res = []
    try:
        for walk in (4, 0):
            _lib.set_tuning(walk=walk)
            s = fx.Solver(cfg, keep_scratch=True)
            s.reset(pos)
            s.substeps_n(dt, 2)
            torch.cuda.synchronize()
            res.append((s.pos.cpu().numpy(), s.vel.cpu().numpy(), s.lam.cpu().numpy()))
            del s
    finally:
        _lib.set_tuning(**saved)
    assert np.isfinite(res[0][0]).all()
    for name, a, b in zip(("pos", "vel", "lam"), res[0], res[1]):
        assert np.array_equal(a, b), name
