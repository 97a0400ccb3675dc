"""Batched scenes and HIP-graph replay (GPU).

The batched launch runs the same pass code as the plain solver with each
thread's scene parameters read from a table, and the counting sort over
scene-offset keys orders each scene exactly as its own sort would, so every
scene's state after K batched substeps must equal K plain substeps of that
scene bit for bit.  Graph replays must equal eager substeps bit for bit, with
eager work and synchronisation between replays.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes(fx):
    from flowx_amd.config import converging_update

    a = fx.dam_break_config()
    b = converging_update(fx.lattice_config((-0.4, -0.4, 0.0), (0.4, 0.4, 0.8), 0.06, (13, 13, 7), iterations=4))
    b.extra.update(colliders=[("box", (-0.1, -0.1, 0.0), (0.1, 0.1, 0.3))], collider_resolution=14)
    c = converging_update(fx.lattice_config((-0.3, -0.3, 0.0), (0.3, 0.3, 0.6), 0.04, (15, 15, 9), iterations=4,
                                            viscosity=0.05))
    return [a, b, c]


def test_batch_equals_plain_per_scene(fx):
    import torch

    cfgs = _scenes(fx)
    batch = fx.BatchSolver(cfgs)
    steps = 6
    for _ in range(steps):
        batch.substep()
    torch.cuda.synchronize()
    for i, cfg in enumerate(_scenes(fx)):
        s = fx.Solver(cfg, keep_scratch=False)
        s.reset()
        _, dt = cfg.substep_dt(1.0 / 24.0)
        assert dt == batch.dts[i]
        s.substeps_n(dt, steps)
        torch.cuda.synchronize()
        assert np.array_equal(batch.scene_positions(i), s.read_positions()), f"scene {i} positions"
        assert np.array_equal(batch.scene_velocities(i), s.read_velocities()), f"scene {i} velocities"


@pytest.mark.parametrize("persistent", [0, 8])
def test_batch_pair_masks_bit_identical(fx, persistent):
    """The batch's lambda records each particle's in-range candidates per
    stencil row and its delta walks only those (walk 0, flat and persistent
    launches): equal bit for bit to the batch without them (walk 4).  One
    scene holds a 120-particle clump in one cell, so some of its particles
    have a row past the 64 slots a word records and their delta walks the
    full stencil."""
    import torch

    from flowx_amd import _lib

    rng = np.random.default_rng(3)
    clump = np.concatenate([rng.uniform([0.021, 0.021, 0.121], [0.059, 0.059, 0.159], (120, 3)),
                            np.ones((120, 1))], 1).astype(np.float32)
    starts = [fx.seed_positions(c) for c in _scenes(fx)]
    starts[2] = np.concatenate([starts[2], clump])

    saved = _lib.get_tuning()
    res = []
    try:
        for walk in (4, 0):
            _lib.set_tuning(walk=walk, persistent=persistent)
            batch = fx.BatchSolver(_scenes(fx), positions=starts)
            for _ in range(4):
                batch.substep()
            torch.cuda.synchronize()
            res.append([(batch.scene_positions(i), batch.scene_velocities(i)) for i in range(3)])
            del batch
    finally:
        _lib.set_tuning(**saved)
    for i in range(3):
        assert np.isfinite(res[0][i][0]).all()
        assert np.array_equal(res[0][i][0], res[1][i][0]), f"scene {i} positions"
        assert np.array_equal(res[0][i][1], res[1][i][1]), f"scene {i} velocities"


def test_batch_with_per_scene_iteration_counts(fx):
    """Scenes with 4, 2, 0, 6 and 3 Jacobi iterations (and their own dt) in
    one batch: the Jacobi launches run six times, a scene past its own count
    carries its state through unchanged, and every scene equals its own
    plain run bit for bit."""
    import torch

    from flowx_amd.config import converging_update

    def scenes():
        out = _scenes(fx)
        d = converging_update(fx.lattice_config((-0.3, -0.3, 0.0), (0.3, 0.3, 0.6), 0.05, (11, 11, 8), iterations=6))
        e = converging_update(fx.lattice_config((-0.5, -0.5, 0.0), (0.5, 0.5, 1.0), 0.07, (13, 13, 6), iterations=3,
                                                viscosity=0.3))
        e.extra.update(colliders=[("box", (0.0, -0.1, 0.0), (0.2, 0.1, 0.2))], collider_resolution=12)
        out[1].iterations = 2
        out[2].iterations = 0
        return out + [d, e]

    cfgs = scenes()
    assert [c.iterations for c in cfgs] == [4, 2, 0, 6, 3]
    batch = fx.BatchSolver(cfgs)
    steps = 5
    for _ in range(steps):
        batch.substep()
    torch.cuda.synchronize()
    for i, cfg in enumerate(scenes()):
        s = fx.Solver(cfg, keep_scratch=False)
        s.reset()
        _, dt = cfg.substep_dt(1.0 / 24.0)
        s.substeps_n(dt, steps)
        torch.cuda.synchronize()
        assert np.array_equal(batch.scene_positions(i), s.read_positions()), f"scene {i} positions"
        assert np.array_equal(batch.scene_velocities(i), s.read_velocities()), f"scene {i} velocities"


@pytest.mark.parametrize("lattice", [(50, 50, 20), (70, 68, 20), (80, 80, 32)])
def test_walks_by_scene_size_are_bit_identical(fx, lattice):
    """A plain scene picks its neighbour walk by size (four lanes per particle
    below a pass's threshold, row by row up to 2^17, the stream walks above);
    the batched launch always streams one lane per particle.  A one-scene
    batch must therefore equal the plain run bit for bit: at 50,000 particles
    lambda and delta walk in quads and XSPH row by row, at 95,200 lambda in
    quads and delta and XSPH row by row, at 204,800 everything streams.
    Cohesion on (the XSPH pass also computes the normals) and a collider."""
    import torch

    from flowx_amd.config import converging_update

    def scene():
        c = converging_update(fx.lattice_config((-1.0, -1.0, 0.0), (1.0, 1.0, 2.0), 0.025, lattice, iterations=4,
                                                viscosity=0.01, surface_tension=0.001))
        c.extra.update(colliders=[("box", (-0.2, -0.2, 0.0), (0.2, 0.2, 0.3))], collider_resolution=40)
        return c

    steps = 5
    batch = fx.BatchSolver([scene()])
    for _ in range(steps):
        batch.substep()
    s = fx.Solver(scene(), keep_scratch=False)
    s.reset()
    s.substeps_n(batch.dts[0], steps)
    torch.cuda.synchronize()
    assert np.array_equal(batch.scene_positions(0), s.read_positions())
    assert np.array_equal(batch.scene_velocities(0), s.read_velocities())


@pytest.mark.parametrize("squeeze", [1.0, 0.5])
def test_lds_tile_walk_is_bit_identical(fx, squeeze):
    """Walk 3 (lambda and delta read their candidates from LDS tiles of the
    workgroup's stencil rows, global memory for runs outside the tile) against
    the default stream walks, bit for bit, on a 204,800-particle lattice with
    a collider.  As seeded, particles leave their build rows of cells between
    Jacobi iterations (runs found in shifted windows or read from global
    memory); squeezed to half its height (16 particles per cell) the nine
    windows of a workgroup overflow the tile and its later rows come from
    global memory."""
    import torch

    from flowx_amd import _lib
    from flowx_amd.config import converging_update

    cfg = converging_update(fx.lattice_config((-1.0, -1.0, 0.0), (1.0, 1.0, 2.0), 0.025, (80, 80, 32), iterations=4,
                                              viscosity=0.01, surface_tension=0.001))
    cfg.extra.update(colliders=[("box", (-0.2, -0.2, 0.0), (0.2, 0.2, 0.3))], collider_resolution=40)
    _, dt = cfg.substep_dt(1.0 / 24.0)
    seed = fx.seed_positions(cfg)
    seed[:, 2] = np.float32(cfg.lo[2]) + (seed[:, 2] - np.float32(cfg.lo[2])) * np.float32(squeeze)
    saved = _lib.get_tuning()
    out = []
    try:
        for walk in (0, 3):
            _lib.set_tuning(walk=walk)
            s = fx.Solver(cfg, keep_scratch=True)
            s.reset()
            s.load_state(seed, None)
            s.substeps_n(dt, 3)
            s.substeps_n(dt, 2)
            torch.cuda.synchronize()
            out.append([getattr(s, f).cpu().numpy() for f in ("pos", "vel", "lam", "pred")])
            del s
    finally:
        _lib.set_tuning(**saved)
    assert np.isfinite(out[0][0]).all()
    for f, a, b in zip(("pos", "vel", "lam", "pred"), *out):
        assert np.array_equal(a, b), f


def test_batch_from_generated_scene_set_runs(fx):
    import torch

    cfgs = fx.batch_configs(count=8, seed=3)
    batch = fx.BatchSolver(cfgs)
    for _ in range(4):
        batch.substep()
    torch.cuda.synchronize()
    pos = batch.pos.cpu().numpy()
    assert np.isfinite(pos).all()
    for i, c in enumerate(cfgs):
        p = batch.scene_positions(i)
        lo, hi = np.array(c.lo), np.array(c.hi)
        assert (p[:, :3] >= lo - 1e-6).all() and (p[:, :3] <= hi + 1e-6).all()


def test_plain_graph_replay_matches_eager(fx):
    import torch

    cfg = fx.dam_break_config()
    _, dt = cfg.substep_dt(1.0 / 24.0)
    eager = fx.Solver(cfg, keep_scratch=False)
    eager.reset()
    graphed = fx.Solver(cfg, keep_scratch=False)
    graphed.reset()
    g = graphed.capture(dt, 3)  # advances graphed by one warm-up substep
    eager.substep(dt)
    for _ in range(3):
        g.replay()
        eager.substeps_n(dt, 3)
        torch.cuda.synchronize()
        _ = eager.pos.sum().item()  # eager work between replays
        assert np.array_equal(graphed.read_positions(), eager.read_positions())
        assert np.array_equal(graphed.read_velocities(), eager.read_velocities())


def test_batch_graph_replay_matches_eager(fx):
    import torch

    eager = fx.BatchSolver(_scenes(fx))
    graphed = fx.BatchSolver(_scenes(fx))
    g = graphed.capture(2)
    eager.substep()
    for _ in range(3):
        g.replay()
        eager.substep()
        eager.substep()
        torch.cuda.synchronize()
        assert np.array_equal(graphed.pos.cpu().numpy(), eager.pos.cpu().numpy())


def test_autotune_sweep_keeps_results_identical(fx, tmp_path):
    """The autotuner's sweep on a 204,800-particle scene (block size x list
    skin x stencil walk, where every full-stencil pass streams or reads LDS
    tiles) and on a three-scene batch (flat and persistent launches): every setting
    reproduces the first setting's positions, and the fastest is written and
    applied."""
    from flowx_amd import _lib, autotune
    from flowx_amd.config import converging_update

    saved = _lib.get_tuning()
    try:
        cfg = converging_update(fx.lattice_config((-1.0, -1.0, 0.0), (1.0, 1.0, 2.0), 0.025, (80, 80, 32),
                                                  iterations=4, viscosity=0.01, surface_tension=0.001))
        cfg.extra.update(colliders=[("box", (-0.2, -0.2, 0.0), (0.2, 0.2, 0.3))], collider_resolution=40)
        out = autotune.tune(fx, cfg, blocks=(64, 256), skins=(0.0, 0.2), walks=(0, 1, 2, 3, 4), batch_cfgs=_scenes(fx),
                            persistents=(0, 2, 8), warmup=1, steps=2, path=tmp_path / "t.json")
        assert len(out["results"]) == 12 and all(r["same_result"] for r in out["results"])
        assert len(out["batch_results"]) == 3 and all(r["same_result"] for r in out["batch_results"])
        assert autotune.apply(tmp_path / "t.json") == out["best"]
        assert _lib.get_tuning()["walk"] == out["best"]["walk"]
    finally:
        _lib.set_tuning(**saved)
