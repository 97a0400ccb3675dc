"""The add-on's timeline driver without Blender.

solver/sph.py:773-855 `_on_frame_change` decides, for every frame the
timeline lands on, whether the run re-seeds, loads a cached frame, steps
forward (at most MAX_CATCHUP_FRAMES frames, picking up the cache's end first
when a jump outruns that budget), or holds with a warning.  `Timeline`
restates that state machine around a `Solver` and an optional `FrameCache`
so a host other than Blender - or a test - drives the same semantics:

* a frame at or before the seed frame re-seeds (`_reset_clock`,
  solver/sph.py:613-619: seed frame = start frame, last frame = the frame);
* a frame the cache holds loads with no simulation (`_apply_cached`,
  :700-733: pred = pos, lambda = 0, grid rebuilt);
* a forward jump steps frame by frame, each step writing the cache
  (`_step`, :517-545);
* a backward jump with no usable cache holds and warns.

Collider motion is the caller's `colliders_at(frame)` (the animated-collider
rebuild of collision/__init__.py:191-222), applied before any of the above.
"""

MAX_CATCHUP_FRAMES = 30  # solver/sph.py:101


class Timeline:
    def __init__(self, solver, frame_start, fps=24.0, cache_factory=None, colliders_at=None, surface=None):
        """solver: a Solver; cache_factory(seed_frame, particle_count) returns
        a FrameCache (or None for no cache) and is called at every (re)seed,
        as cache.open is (solver/sph.py:621-660); surface: an optional
        Surface updated after every step and load."""
        self.solver = solver
        self.frame_start = int(frame_start)
        self.fps = float(fps)
        self.cache_factory = cache_factory
        self.colliders_at = colliders_at
        self.surface = surface
        self.cache = None
        self.seed_frame = self.last_frame = self.frame_start
        self.warning = None
        self.running = False

    # ---- run control (solver/sph.py:613-690) -------------------------------
    def start(self, frame_current=None):
        self.running = True
        self.reseed(self.frame_start if frame_current is None else frame_current)

    def reseed(self, frame_current):
        if self.cache is not None:
            self.cache.close()
        self._apply_colliders(frame_current)
        self.solver.reset()
        self.seed_frame = self.frame_start
        self.last_frame = int(frame_current)
        self.warning = None
        self.cache = self.cache_factory(int(frame_current), self.solver.cfg.particle_count) if self.cache_factory else None
        self._update_surface()

    def stop(self):
        if self.cache is not None:
            self.cache.close()
            self.cache = None
        self.running = False
        self.warning = None

    # ---- per frame (solver/sph.py:773-855) ----------------------------------
    def on_frame(self, frame):
        if not self.running:
            return
        frame = int(frame)
        if frame == self.last_frame:
            return
        self._apply_colliders(frame)
        if frame <= self.seed_frame:
            self.reseed(frame)
            return
        loaded = self._load_cached(frame)
        if loaded is not None:
            self._apply_cached(*loaded, frame)
            return
        pending = frame - self.last_frame
        if pending < 0:
            self.warning = (f"Frame {frame} is behind the simulation, which is at frame {self.last_frame}: "
                            "no cached frame to scrub back through; return to the seed frame to re-run.")
            return
        if pending > MAX_CATCHUP_FRAMES:
            if self._pick_up_cache_end(frame):
                pending = frame - self.last_frame
            if pending > MAX_CATCHUP_FRAMES:
                self.warning = (f"Jumped {frame - self.last_frame} frames; only {MAX_CATCHUP_FRAMES} were "
                                f"simulated, so this frame is approximate.  Return to frame {self.seed_frame} "
                                "and play forward for a correct result.")
                pending = MAX_CATCHUP_FRAMES
        else:
            self.warning = None
        target = self.last_frame + pending
        while self.last_frame < target:
            self._step(self.last_frame + 1)
        if frame != target:
            self.last_frame = frame

    # ---- pieces ---------------------------------------------------------------
    def _apply_colliders(self, frame):
        if self.colliders_at is not None and self.solver.collider is not None:
            shapes = self.colliders_at(frame)
            if shapes is not None:
                self.solver.collider.rebuild(shapes)

    def _step(self, frame):
        """solver/sph.py:517-545: the frame's substeps, the cache write and the
        surface."""
        self.solver.step(1.0 / self.fps)
        if self.cache is not None:
            self.cache.write_frame(frame, self.solver.pos, self.solver.vel)
        self._update_surface()
        self.last_frame = frame

    def _load_cached(self, frame):
        """solver/sph.py:736-750."""
        if self.cache is None:
            return None
        return self.cache.load(frame)

    def _apply_cached(self, positions, velocities, frame):
        self.solver.load_state(positions, velocities)
        self.last_frame = frame
        self.warning = None
        self._update_surface()

    def _pick_up_cache_end(self, frame):
        """solver/sph.py:753-770."""
        if self.cache is None:
            return False
        self.cache.flush()
        end = self.cache.header["last_frame"]
        if not self.last_frame < end < frame:
            return False
        loaded = self._load_cached(end)
        if loaded is None:
            return False
        self._apply_cached(*loaded, end)
        return True

    def _update_surface(self):
        if self.surface is not None:
            self.surface.update()

    def stats(self):
        out = self.solver.stats()
        out.update(frame=self.last_frame, seed_frame=self.seed_frame, warning=self.warning)
        return out
