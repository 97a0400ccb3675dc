"""The FLWXCA01 cache fed from the GPU: asynchronous device-to-host frame
writes round-trip exactly, and a run resumed from a cached frame reproduces
the frames that followed it bit for bit (the reference smoke test's exact
round-trip check, scripts/smoke_test.py:187-338)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_async_frames_round_trip_and_resume_bit_exact(fx, tmp_path):
    cfg = fx.dam_break_config()
    assert cfg.surface_tension == 0.0  # then (pos, vel) is the whole state at a frame boundary
    s = fx.Solver(cfg, keep_scratch=False)
    s.reset()
    digest = fx.config_hash(fx.DomainSettings(), 24.0, cfg.lo, cfg.hi, [("Box", None, b"static")])
    cache = fx.FrameCache(tmp_path / "dam.flowx_cache", cfg.particle_count, seed_frame=1, fps=24.0,
                          config_digest=digest, colliders=["Box"])
    snaps = {}
    for frame in (2, 3, 4, 5):
        s.step(1 / 24.0)
        cache.write_frame(frame, s.pos, s.vel)  # queued; the solver keeps stepping
        snaps[frame] = (s.read_positions().copy(), s.read_velocities().copy())
    cache.flush()
    for frame, (p, v) in snaps.items():
        lp, lv = cache.load(frame)
        assert lp.tobytes() == p.tobytes() and lv.tobytes() == v.tobytes(), frame
    assert cache.header["last_frame"] == 5
    p3, v3 = cache.load(3)
    resumed = fx.Solver(fx.dam_break_config(), keep_scratch=False)
    resumed.reset()
    resumed.load_state(p3, v3)
    for frame in (4, 5):
        resumed.step(1 / 24.0)
        assert np.array_equal(resumed.read_positions(), snaps[frame][0]), frame
        assert np.array_equal(resumed.read_velocities(), snaps[frame][1]), frame
    cache.close()
