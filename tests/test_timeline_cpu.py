"""The timeline driver (flow-x_amd/timeline.py, restating solver/sph.py:773-855)
on the host, with a stand-in solver and cache: the reference smoke test's
playback contract (scripts/smoke_test.py:137-184) and its cache contract
(:187-338) - re-seed at or before the start frame, step forward, hold and
warn on a backward scrub without a cache, load from the cache with no
warning, clamp long jumps to 30 frames after picking up the cache's end."""

import numpy as np


class FakeSolver:
    class Cfg:
        particle_count = 4

    def __init__(self):
        self.cfg = self.Cfg()
        self.collider = None
        self.state = 0.0  # the frame the particle state represents
        self.resets = self.steps = self.loads = 0

    @property
    def pos(self):
        return np.full((4, 4), self.state, np.float32)

    @property
    def vel(self):
        return np.full((4, 4), -self.state, np.float32)

    def reset(self):
        self.resets += 1
        self.state = 0.0

    def step(self, frame_dt):
        self.steps += 1
        self.state += 1.0

    def load_state(self, p, v):
        self.loads += 1
        self.state = float(p[0, 0])

    def stats(self):
        return {}


class FakeCache:
    def __init__(self, seed):
        self.frames = {}
        self.header = {"seed_frame": seed, "last_frame": seed}

    def write_frame(self, frame, pos, vel):
        self.frames[frame] = (np.array(pos), np.array(vel))
        self.header["last_frame"] = max(self.header["last_frame"], frame)

    def load(self, frame):
        return self.frames.get(frame) if self.header["seed_frame"] < frame <= self.header["last_frame"] else None

    def flush(self):
        pass

    def close(self):
        pass


def _timeline(fx, cache=False, start=1):
    from flowx_amd.timeline import Timeline

    s = FakeSolver()
    caches = []

    def factory(seed, n):
        c = FakeCache(seed)
        caches.append(c)
        return c

    t = Timeline(s, frame_start=start, cache_factory=factory if cache else None)
    t.start()
    return t, s, caches


def test_playback_without_cache(fx):
    t, s, _ = _timeline(fx)
    t.on_frame(5)
    assert (t.last_frame, s.steps, t.warning) == (5, 4, None)
    t.on_frame(5)  # re-entering the current frame does nothing
    assert s.steps == 4
    t.on_frame(3)  # backward scrub: hold and say so
    assert t.last_frame == 5 and s.state == 4.0 and "behind the simulation" in t.warning
    t.on_frame(1)  # the start frame re-seeds and clears the warning
    assert (t.last_frame, s.resets, t.warning, s.state) == (1, 2, None, 0.0)
    t.on_frame(-3)  # anything before the start frame re-seeds too
    assert (t.last_frame, s.resets, t.seed_frame) == (-3, 3, 1)


def test_long_jump_is_clamped_to_the_catch_up_budget(fx):
    from flowx_amd.timeline import MAX_CATCHUP_FRAMES

    t, s, _ = _timeline(fx)
    t.on_frame(100)
    assert s.steps == MAX_CATCHUP_FRAMES == 30
    assert t.last_frame == 100 and "approximate" in t.warning
    t.on_frame(101)  # a normal step clears the warning
    assert t.warning is None and s.steps == 31


def test_cache_scrub_back_and_pick_up_of_its_end(fx):
    t, s, caches = _timeline(fx, cache=True)
    t.on_frame(30)
    t.on_frame(41)
    c = caches[-1]
    assert sorted(c.frames) == list(range(2, 42)) and s.steps == 40
    t.on_frame(10)  # backward scrub loads from the cache, no warning
    assert (t.last_frame, s.state, s.loads, t.warning) == (10, 9.0, 1, None)
    p1 = c.load(10)[0].copy()
    t.on_frame(11)  # cached too: loads, no simulation
    t.on_frame(10)
    assert s.steps == 40 and np.array_equal(c.load(10)[0], p1)  # loading twice round-trips
    # from frame 10, a jump to 60 (50 frames) first picks up the cache's end
    # (frame 41), then steps the remaining 19 frames: no approximation warning
    t.on_frame(60)
    assert t.last_frame == 60 and t.warning is None and s.steps == 40 + 19
    assert s.state == 59.0  # frame f holds state f - 1 (the seed frame is state 0)
    # re-seeding re-opens the cache
    t.on_frame(1)
    assert len(caches) == 2 and s.resets == 2
