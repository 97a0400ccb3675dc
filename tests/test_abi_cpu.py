"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/flowx.h declares, its ctypes structs match the header's layout, and
argument validation errors come back as codes (no GPU calls are made)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    text = (ROOT / "include" / "flowx.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(fx):
    lib = fx._lib.lib()
    missing = [name for name in _declared() if not hasattr(lib, name)]
    assert not missing, f"declared in flowx.h but not exported: {missing}"
    assert lib.fx_abi_version() == 2


def test_every_declared_symbol_has_a_ctypes_signature(fx):
    missing = [name for name in _declared() if name not in fx._lib.SIGNATURES]
    assert not missing


def test_struct_layouts_match_between_product_and_oracle(fx, oracle):
    for a, b in ((fx._lib.FxParams, oracle.FxParams), (fx._lib.FxSurface, oracle.FxSurface)):
        assert ctypes.sizeof(a) == ctypes.sizeof(b)
        assert [f[0] for f in a._fields_] == [f[0] for f in b._fields_]
    assert ctypes.sizeof(fx._lib.FxParams) == 30 * 4  # thirty 4-byte lanes


def test_workspace_size_grows_with_particles(fx):
    cfg = fx.dam_break_config()
    p = cfg.params(0.01)
    a = fx._lib.lib().fx_workspace_bytes(ctypes.byref(p))
    p.n = p.n * 2
    b = fx._lib.lib().fx_workspace_bytes(ctypes.byref(p))
    assert b > a > 16 * cfg.particle_count


def test_invalid_arguments_return_codes_without_gpu(fx):
    L = fx._lib
    cfg = fx.dam_break_config()
    p = cfg.params(0.01)
    s = L.FxState()
    # no workspace -> FX_EINVAL before any device work
    rc = L.lib().fx_substep(ctypes.byref(p), ctypes.byref(s), None)
    assert rc == -1
    assert b"workspace" in L.lib().fx_last_error()
    p.cells[0] = 0
    rc = L.lib().fx_build_grid(ctypes.byref(p), ctypes.byref(s), None)
    assert rc == -1 and b"cell dims" in L.lib().fx_last_error()
    with pytest.raises(RuntimeError):
        L.check(rc, "fx_build_grid")
    # slots are addressed by 32-bit byte offsets: 2^28 particles are refused
    p = cfg.params(0.01)
    p.n = 1 << 28
    rc = L.lib().fx_build_grid(ctypes.byref(p), ctypes.byref(s), None)
    assert rc == -4 and b"particle count" in L.lib().fx_last_error()


def test_mc_table_of_library_matches_reference_golden(fx):
    import numpy as np

    g = np.load(ROOT / "tests" / "golden" / "mc_golden.npz")
    tri = np.zeros((256, 16), np.int8)
    mask = np.zeros(256, np.uint16)
    fx._lib.check(fx._lib.lib().fx_mc_tables(ctypes.c_void_p(tri.ctypes.data), ctypes.c_void_p(mask.ctypes.data)),
                  "fx_mc_tables")
    assert np.array_equal(tri, g["tri_table"])
    assert np.array_equal(mask, g["edge_table"])


def test_autotune_file_is_applied(fx, tmp_path):
    """A tuning file written by the autotuner is loaded into the library."""
    import json

    from flowx_amd import _lib, autotune

    saved = _lib.get_tuning()
    try:
        f = tmp_path / "tuning.json"
        f.write_text(json.dumps({"best": {"block": 128, "list_skin": 0.2}, "results": []}))
        assert autotune.apply(f) == {"block": 128, "list_skin": 0.2}
        t = _lib.get_tuning()
        assert t["block"] == 128 and abs(t["list_skin"] - 0.2) < 1e-7
        assert autotune.apply(tmp_path / "missing.json") is None
    finally:
        _lib.set_tuning(**saved)


def test_tuning_round_trip_and_rejects_bad_values(fx):
    """fx_set_tuning stores every FxTuning field and refuses out-of-range
    values without changing the stored tuning."""
    from flowx_amd import _lib

    saved = _lib.get_tuning()
    try:
        _lib.set_tuning(list_skin=0.1, block=256)
        t = _lib.get_tuning()
        assert t["block"] == 256 and abs(t["list_skin"] - 0.1) < 1e-7
        _lib.set_tuning(persistent=4, walk=3)
        t = _lib.get_tuning()
        assert t["persistent"] == 4 and t["walk"] == 3
        for bad in ({"block": 96}, {"list_skin": -0.5}, {"list_skin": 2.0}, {"persistent": -1}, {"walk": 5},
                    {"walk": -1}):
            with pytest.raises(RuntimeError):
                _lib.set_tuning(**bad)
            assert _lib.get_tuning() == t
    finally:
        _lib.set_tuning(**saved)
