"""FLWXCA01 v2 cache file (flow-x_amd/cache.py) on the host: the byte layout
of solver/cache.py:12-34 / :74 / :203-266, exact frame round trips, the torn
tail rule (:339-373) and reuse of a matching file (:293-305)."""

import struct

import numpy as np
import pytest


@pytest.fixture()
def cache_mod(fx):
    from flowx_amd import cache

    return cache


def _header(cache_mod, **kw):
    h = {"format_version": 2, "flowx_version": "0.3.1", "particle_count": 5, "seed_frame": 1,
         "last_frame": 1, "fps": 24.0, "colliders": ["Box", "Sphéré"], "config_hash": bytes(range(32))}
    h.update(kw)
    return h


def test_fixed_header_is_48_bytes_little_endian(cache_mod):
    raw = cache_mod.pack_header(_header(cache_mod))
    assert cache_mod.FIXED_HEADER.size == 48
    assert raw[0:8] == b"FLWXCA01"
    assert int.from_bytes(raw[8:12], "little") == 2
    assert raw[12:28] == b"0.3.1".ljust(16, b"\x00")
    assert int.from_bytes(raw[28:32], "little") == 5
    assert int.from_bytes(raw[32:36], "little", signed=True) == 1
    assert int.from_bytes(raw[36:40], "little", signed=True) == 1
    assert np.frombuffer(raw[40:44], "<f4")[0] == 24.0
    assert int.from_bytes(raw[44:48], "little") == 2
    # (u32 length + UTF-8 name) per collider, then the 32-byte hash
    assert int.from_bytes(raw[48:52], "little") == 3 and raw[52:55] == b"Box"
    name = "Sphéré".encode("utf-8")
    assert int.from_bytes(raw[55:59], "little") == len(name) and raw[59:59 + len(name)] == name
    assert raw[59 + len(name):] == bytes(range(32))
    assert len(raw) == 59 + len(name) + 32


def test_header_round_trip_and_rejections(cache_mod):
    h = _header(cache_mod, seed_frame=-3, last_frame=7)
    raw = cache_mod.pack_header(h)
    u = cache_mod.unpack_header(raw + b"trailing frame bytes")
    assert u["colliders"] == h["colliders"] and u["config_hash"] == h["config_hash"]
    assert (u["seed_frame"], u["last_frame"], u["particle_count"]) == (-3, 7, 5)
    assert u["header_size"] == len(raw) and u["frame_size"] == 8 * 4 * 5
    assert cache_mod.unpack_header(b"FLWXCA02" + raw[8:]) is None
    assert cache_mod.unpack_header(raw[:8] + struct.pack("<I", 1) + raw[12:]) is None
    assert cache_mod.unpack_header(raw[:40]) is None
    assert cache_mod.unpack_header(raw[:-1]) is None  # hash cut short


def _frames(n, count, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        p = rng.standard_normal((n, 4)).astype(np.float32)
        p[:, 3] = 1
        v = rng.standard_normal((n, 4)).astype(np.float32)
        v[:, 3] = 0
        out.append((p, v))
    out[0][0][0, 0] = np.float32(-0.0)
    out[0][1][1, 1] = np.float32(np.inf)
    return out


def test_frames_round_trip_bit_exact_at_arithmetic_offsets(cache_mod, tmp_path):
    n = 37
    frames = _frames(n, 4)
    path = tmp_path / "run.flowx_cache"
    c = cache_mod.FrameCache(path, n, seed_frame=10, fps=24.0, config_digest=b"\x07" * 32, colliders=["A"])
    for k in (3, 1, 4, 2):  # any order
        c.write_frame_host(10 + k, *frames[k - 1])
    for k in range(1, 5):
        p, v = c.load(10 + k)
        assert p.tobytes() == frames[k - 1][0].tobytes() and v.tobytes() == frames[k - 1][1].tobytes()
    assert c.load(10) is None and c.load(15) is None  # the seed frame is never stored
    c.close()
    raw = path.read_bytes()
    h = cache_mod.unpack_header(raw)
    assert h["last_frame"] == 14
    off = h["header_size"] + (12 - 10 - 1) * 8 * 4 * n
    assert raw[off: off + 16 * n] == frames[1][0].tobytes()
    assert raw[off + 16 * n: off + 32 * n] == frames[1][1].tobytes()
    assert len(raw) == h["header_size"] + 4 * 32 * n


def test_torn_tail_is_ignored(cache_mod, tmp_path):
    n = 8
    frames = _frames(n, 3, seed=1)
    path = tmp_path / "torn.flowx_cache"
    c = cache_mod.FrameCache(path, n, 0, 24.0, b"\x01" * 32)
    for k in (1, 2):
        c.write_frame_host(k, *frames[k - 1])
    c.close()
    # a crash after the frame-3 data but before its header rewrite
    h = cache_mod.read_header(path)
    with open(path, "r+b") as f:
        f.seek(h["header_size"] + 2 * h["frame_size"])
        f.write(frames[2][0].tobytes() + frames[2][1].tobytes()[:-7])
    c2 = cache_mod.FrameCache(path, n, 0, 24.0, b"\x01" * 32)
    assert c2.header["last_frame"] == 2
    assert c2.load(3) is None
    assert c2.load(2)[0].tobytes() == frames[1][0].tobytes()
    c2.close()


def test_matching_file_is_reused_and_mismatch_rewritten(cache_mod, tmp_path):
    n = 6
    frames = _frames(n, 2, seed=2)
    path = tmp_path / "reuse.flowx_cache"
    c = cache_mod.FrameCache(path, n, 5, 24.0, b"\x02" * 32)
    c.write_frame_host(6, *frames[0])
    c.close()
    c = cache_mod.FrameCache(path, n, 5, 24.0, b"\x02" * 32)  # same run: append
    assert c.load(6)[1].tobytes() == frames[0][1].tobytes()
    c.write_frame_host(7, *frames[1])
    c.close()
    for kw in ({"config_digest": b"\x03" * 32}, {"particle_count": n + 1}, {"seed_frame": 4}):
        args = {"particle_count": n, "seed_frame": 5, "fps": 24.0, "config_digest": b"\x02" * 32}
        args.update(kw)
        c = cache_mod.FrameCache(path, **args)
        assert c.header["last_frame"] == c.header["seed_frame"] and c.load(6) is None
        c.close()
        # restore the matching file for the next case
        c = cache_mod.FrameCache(path, n, 5, 24.0, b"\x02" * 32)
        c.write_frame_host(6, *frames[0])
        c.close()


def test_config_hash_covers_every_physics_input(cache_mod, fx):
    s = fx.DomainSettings()
    base = dict(settings=s, fps=24.0, lo=(-1.0, -1.0, 0.0), hi=(1.0, 1.0, 2.0),
                colliders=[("Box", b"mesh", b"motion")])
    d0 = cache_mod.config_hash(**base)
    assert d0 == cache_mod.config_hash(**base) and len(d0) == 32
    seen = {d0}
    for field in ("resolution", "fluid_level", "smoothing_radius", "rest_density", "pbf_relaxation",
                  "pbf_scorr_k", "viscosity", "surface_tension", "max_substeps", "pbf_iterations"):
        t = fx.DomainSettings(**{**s.__dict__, field: getattr(s, field) + 1})
        seen.add(cache_mod.config_hash(**{**base, "settings": t}))
    seen.add(cache_mod.config_hash(**{**base, "fps": 25.0}))
    seen.add(cache_mod.config_hash(**{**base, "hi": (1.0, 1.0, 2.5)}))
    seen.add(cache_mod.config_hash(**{**base, "colliders": [("Box", b"mesh2", b"motion")]}))
    seen.add(cache_mod.config_hash(**{**base, "colliders": [("Box", b"mesh", b"motion2")]}))
    seen.add(cache_mod.config_hash(**{**base, "colliders": []}))
    assert len(seen) == 16
