"""Autotuner for the neighbour-pass launch parameters.

No reference counterpart.  Sweeps the tuning knobs of include/flowx.h
(FxTuning): the block size of the neighbour passes, the neighbour-list
skin, the stencil walk of a large scene's neighbour passes (how the
stencil's cell rows are staged into batches of loads in flight, or into LDS
tiles per workgroup, and whether delta walks the pair masks lambda
records) on a plain scene, and the workgroups per CU of the batched
passes' persistent launch (0 = flat launch) on a batch of scenes.
Each setting is timed with HIP events after a warm-up - the plain scene
over a few frames, each one call of the frame's substeps (the first enters
from the canonical state, the rest in sorted order, as playback runs them;
single-substep calls favour other block sizes), the batch over a few
batched substeps - and every setting must give the same positions as the
first (the knobs change launch shape and candidate iteration, never
results).  The fastest setting is written to a JSON file
that ``apply()`` loads - by default ``tuning.json`` next to the library,
read by ``Solver`` and ``BatchSolver`` at construction when present.
``python bench.py --autotune`` tunes on the 1.64M-particle scene and the
64-scene batch set.
"""

import json
import os
from pathlib import Path

from . import _lib

DEFAULT_PATH = Path(__file__).resolve().parent / "tuning.json"
BLOCKS = (64, 128, 256)
SKINS = (0.0, 0.1, 0.2, 0.3)
WALKS = (0, 1, 2, 3, 4)
PERSISTENT = (0, 8, 16, 24, 32)
KNOBS = ("block", "list_skin", "walk", "persistent")


def tuning_path():
    return Path(os.environ.get("FLOWX_TUNING", DEFAULT_PATH))


def apply(path=None):
    """Load a tuning file written by ``tune`` (if any) into the library.
    Returns the applied settings, or None when there is no file."""
    p = Path(path) if path else tuning_path()
    if not p.exists():
        return None
    best = json.loads(p.read_text())["best"]
    _lib.set_tuning(**{k: (float(v) if k == "list_skin" else int(v)) for k, v in best.items() if k in KNOBS})
    return best


def _timed(run, setting, warmup, steps):
    import torch

    _lib.set_tuning(**setting)
    state, step, per_step = run()
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / (steps * per_step)
    pos = state().clone()
    torch.cuda.empty_cache()
    return ms, pos


def _sweep(run, settings, warmup, steps):
    import torch

    results, ref = [], None
    for setting in settings:
        ms, pos = _timed(run, setting, warmup, steps)
        if ref is None:
            ref = pos
        same = bool(torch.equal(pos, ref))
        results.append({**setting, "ms_per_substep": round(ms, 4), "same_result": same})
        if not same:
            raise RuntimeError(f"tuning {setting} changed the result")
    return results


def tune(fx, cfg=None, blocks=BLOCKS, skins=SKINS, walks=WALKS, batch_cfgs=None, persistents=PERSISTENT,
         warmup=4, steps=12, path=None):
    """Time every (block, skin, walk) on the plain scene `cfg` (default: the
    1.64M scene), then every persistent setting on the batch `batch_cfgs`
    (default: the 64-scene set; None skips) with the best plain block, and
    write the results and the fastest settings to `path`."""
    cfg = cfg or fx.scale_config()
    saved = _lib.get_tuning()
    frame, dt = cfg.substep_dt(1.0 / cfg.extra.get("fps", 24.0))

    def plain():
        # a frame's substeps per call, as a playback frame runs them: the
        # first enters from the canonical state, the rest in sorted order
        solver = fx.Solver(cfg, keep_scratch=False)
        solver.reset()
        return (lambda: solver.pos), (lambda: solver.substeps_n(dt, frame)), frame

    try:
        base = {"persistent": 0}
        plain_settings = [{**base, "block": b, "list_skin": s, "walk": w} for b in blocks for s in skins for w in walks
                          if s == 0.0 or w == 0]  # the walk knob applies to the full-stencil passes
        results = _sweep(plain, plain_settings, warmup, steps)
        best = dict(min(results, key=lambda r: r["ms_per_substep"]))
        batch_results = []
        if batch_cfgs is not False:
            cfgs = batch_cfgs if batch_cfgs is not None else fx.batch_configs(count=64, seed=0)

            def batched():
                b = fx.BatchSolver(cfgs)
                return (lambda: b.pos), b.substep, 1

            batch_settings = [{"block": best["block"], "list_skin": best["list_skin"], "walk": best["walk"],
                               "persistent": k} for k in persistents]
            batch_results = _sweep(batched, batch_settings, warmup, steps)
            best["persistent"] = min(batch_results, key=lambda r: r["ms_per_substep"])["persistent"]
    finally:
        _lib.set_tuning(**saved)
    out = {"scene": f"{cfg.particle_count} particles", "particles": cfg.particle_count, "results": results,
           "batch_results": batch_results, "best": {k: best[k] for k in KNOBS}}
    p = Path(path) if path else tuning_path()
    p.write_text(json.dumps(out, indent=1))
    return out
