"""The timeline driver on the GPU with the real solver, the FLWXCA01 cache,
an animated collider and the surface: playing a run, scrubbing back through
the cache and re-running from the seed frame reproduces every cached frame
bit for bit (the determinism the reference's cache relies on,
solver/cache.py:1-12)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_timeline_replays_cached_frames_exactly(fx, tmp_path):
    from flowx_amd.timeline import Timeline

    cfg = fx.dam_break_config()
    cfg.extra["colliders"] = [("sphere", (0.0, 0.0, 1.2), 0.25)]

    def colliders_at(frame):  # a sphere sinking 2 cm per frame
        return [("sphere", (0.0, 0.0, 1.2 - 0.02 * (frame - 1)), 0.25)]

    solver = fx.Solver(cfg, keep_scratch=False)
    surf = fx.Surface(solver, 32, 0.5, max_samples=1 << 30)
    digest = fx.config_hash(fx.DomainSettings(), 24.0, cfg.lo, cfg.hi, [])
    caches = []

    def factory(seed, n):
        caches.append(fx.FrameCache(tmp_path / f"run{len(caches)}.flowx_cache", n, seed, 24.0, digest))
        return caches[-1]

    t = Timeline(solver, frame_start=1, cache_factory=factory, colliders_at=colliders_at, surface=surf)
    t.start()
    t.on_frame(6)
    first = {f: caches[0].load(f) for f in range(2, 7)}
    assert t.warning is None and surf.triangles > 0
    t.on_frame(3)  # scrub back: load, no simulation, no warning
    assert t.warning is None and np.array_equal(solver.read_positions(), first[3][0])
    t.on_frame(1)  # re-seed, then replay
    t.on_frame(6)
    for f in range(2, 7):
        p, v = caches[-1].load(f)
        assert p.tobytes() == first[f][0].tobytes() and v.tobytes() == first[f][1].tobytes(), f
    t.stop()
