"""Autotuner for the neighbour-pass launch parameters.

No reference counterpart.  Sweeps the tuning knobs of include/flowx.h
(FxTuning: block size of the neighbour passes, neighbour-list skin) on a
scene, times each setting with HIP events over a few substeps after a
warm-up, checks that every setting gives the same positions as the first
(the knobs change launch shape and candidate iteration, never results), and
writes the fastest to a JSON file that ``apply()`` loads - by default
``tuning.json`` next to the library, read by ``Solver`` at construction when
present.  ``python bench.py --autotune`` tunes on the 1.64M-particle scene.
"""

import json
import os
from pathlib import Path

from . import _lib

DEFAULT_PATH = Path(__file__).resolve().parent / "tuning.json"
BLOCKS = (64, 128, 256)
SKINS = (0.0, 0.1, 0.2, 0.3)


def tuning_path():
    return Path(os.environ.get("FLOWX_TUNING", DEFAULT_PATH))


def apply(path=None):
    """Load a tuning file written by ``tune`` (if any) into the library.
    Returns the applied settings, or None when there is no file."""
    p = Path(path) if path else tuning_path()
    if not p.exists():
        return None
    best = json.loads(p.read_text())["best"]
    _lib.set_tuning(list_skin=float(best["list_skin"]), block=int(best["block"]))
    return best


def _time_setting(fx, cfg, block, skin, warmup, steps):
    import torch

    _lib.set_tuning(list_skin=skin, block=block)
    solver = fx.Solver(cfg, keep_scratch=False)
    solver.reset()
    _, dt = cfg.substep_dt(1.0 / cfg.extra.get("fps", 24.0))
    solver.substeps_n(dt, warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    solver.substeps_n(dt, steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    pos = solver.pos.clone()
    del solver
    torch.cuda.empty_cache()
    return ms, pos


def tune(fx, cfg=None, blocks=BLOCKS, skins=SKINS, warmup=4, steps=12, path=None):
    """Time every (block, skin) pair on `cfg` (default: the 1.64M scene) and
    write the results and the fastest pair to `path`."""
    import torch

    cfg = cfg or fx.scale_config()
    saved = _lib.get_tuning()
    results, ref = [], None
    try:
        for block in blocks:
            for skin in skins:
                ms, pos = _time_setting(fx, cfg, block, skin, warmup, steps)
                if ref is None:
                    ref = pos
                same = bool(torch.equal(pos, ref))
                results.append({"block": block, "list_skin": skin, "ms_per_substep": round(ms, 4),
                                "same_result": same})
                if not same:
                    raise RuntimeError(f"tuning block={block} skin={skin} changed the result")
    finally:
        _lib.set_tuning(list_skin=saved["list_skin"], block=saved["block"])
    best = min(results, key=lambda r: r["ms_per_substep"])
    out = {"scene": cfg.extra.get("update", ""), "particles": cfg.particle_count, "results": results,
           "best": {"block": best["block"], "list_skin": best["list_skin"]}}
    p = Path(path) if path else tuning_path()
    p.write_text(json.dumps(out, indent=1))
    return out

