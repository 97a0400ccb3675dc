"""The FLWXCA01 v2 per-frame disk cache, without Blender.

Byte-for-byte the reference's file format (solver/cache.py:12-34, :74-87,
:203-266): a little-endian header

    magic "FLWXCA01" (8s), format_version (I) = 2, flowx_version (16s),
    particle_count (I), seed_frame (i), last_frame (i), fps (f),
    collider_count (I), per collider: name length (I) + UTF-8 name,
    config_hash (32s, sha256)

followed by fixed-size frames seed_frame + 1 ... last_frame, each N x 4
float32 positions then N x 4 float32 velocities (the solver's AoS-4 device
layout, w = 1 / 0).  last_frame is rewritten after every frame, so a torn
tail is ignored on load (:339-373).

The device side: positions and velocities are copied from the GPU into
pinned host buffers asynchronously on a side stream, and a writer thread
waits for that copy and writes the frame and the header, so the solver keeps
stepping while the previous frame goes to disk.  ``flush()`` waits for every
queued frame.

The config hash (solver/cache.py:153-190) covers the same fields in the same
byte order; Blender-only inputs (a collider's mesh and motion fingerprints)
are passed in as bytes by the caller.
"""

import hashlib
import queue
import struct
import threading
from pathlib import Path

import numpy as np

MAGIC = b"FLWXCA01"
FORMAT_VERSION = 2
FIXED_HEADER = struct.Struct("<8sI16sIiifI")  # solver/cache.py:74
HEADER_READ_BYTES = 8192  # solver/cache.py:78


def config_hash(settings, fps, lo, hi, colliders=()):
    """sha256 over the physics inputs (solver/cache.py:153-190).

    settings: a DomainSettings-like object; lo, hi: world bounds; colliders:
    (name, mesh_fingerprint bytes or None, motion_fingerprint bytes) in the
    cache's canonical (sorted-name) order."""
    digest = hashlib.sha256()
    digest.update(struct.pack("<f3f3fI", fps, *lo, *hi, int(settings.resolution)))
    digest.update(struct.pack(
        "<fffffffII", settings.fluid_level, settings.smoothing_radius, settings.rest_density,
        settings.pbf_relaxation, settings.pbf_scorr_k, settings.viscosity, settings.surface_tension,
        int(settings.max_substeps), int(settings.pbf_iterations)))
    for name, mesh_fp, motion_fp in colliders:
        digest.update(name.encode("utf-8"))
        digest.update(mesh_fp if mesh_fp is not None else b"")
        digest.update(motion_fp)
    return digest.digest()


def pack_header(header):
    """solver/cache.py:203-222 _pack_header."""
    parts = [FIXED_HEADER.pack(
        MAGIC, header["format_version"], header["flowx_version"].encode("utf-8")[:16].ljust(16, b"\x00"),
        header["particle_count"], header["seed_frame"], header["last_frame"], header["fps"],
        len(header["colliders"]))]
    for name in header["colliders"]:
        raw = name.encode("utf-8")
        parts.append(struct.pack("<I", len(raw)))
        parts.append(raw)
    parts.append(header["config_hash"])
    return b"".join(parts)


def unpack_header(data):
    """solver/cache.py:225-266 _unpack_header: a header dict, or None when the
    bytes are not a version-2 FLWXCA01 header."""
    if len(data) < FIXED_HEADER.size:
        return None
    try:
        magic, version, flowx, count, seed, last, fps, ncol = FIXED_HEADER.unpack_from(data, 0)
        if magic != MAGIC or version != FORMAT_VERSION:
            return None
        offset = FIXED_HEADER.size
        names = []
        for _ in range(ncol):
            (length,) = struct.unpack_from("<I", data, offset)
            raw = data[offset + 4: offset + 4 + length]
            if len(raw) < length:
                return None
            names.append(raw.decode("utf-8"))
            offset += 4 + length
        digest = data[offset: offset + 32]
        if len(digest) < 32:
            return None
    except (struct.error, UnicodeDecodeError):
        return None
    return {
        "format_version": version, "flowx_version": flowx.decode("utf-8", "replace").rstrip("\x00"),
        "particle_count": count, "seed_frame": seed, "last_frame": last, "fps": fps, "colliders": names,
        "config_hash": digest, "header_size": offset + 32, "frame_size": 8 * 4 * count,
    }


def read_header(path):
    try:
        with Path(path).open("rb") as handle:
            return unpack_header(handle.read(HEADER_READ_BYTES))
    except OSError:
        return None


class FrameCache:
    """One run's cache file (solver/cache.py open / write_frame / try_load).

    An existing file is reused for appending only while its config hash,
    particle count and seed frame match the run (solver/cache.py:293-305);
    otherwise it is rewritten with a fresh header."""

    def __init__(self, path, particle_count, seed_frame, fps, config_digest, colliders=(),
                 flowx_version="flowx-hip"):
        self.path = Path(path)
        self.path.parent.mkdir(parents=True, exist_ok=True)
        existing = read_header(self.path)
        if (existing is not None and existing["config_hash"] == config_digest
                and existing["particle_count"] == particle_count and existing["seed_frame"] == seed_frame):
            self.file = self.path.open("r+b")
            self.header = existing
        else:
            self.file = self.path.open("w+b")
            self.header = {
                "format_version": FORMAT_VERSION, "flowx_version": flowx_version,
                "particle_count": int(particle_count), "seed_frame": int(seed_frame),
                "last_frame": int(seed_frame), "fps": float(fps), "colliders": list(colliders),
                "config_hash": config_digest, "header_size": 0, "frame_size": 8 * 4 * int(particle_count),
            }
            packed = pack_header(self.header)
            self.header["header_size"] = len(packed)
            self.file.write(packed)
            self.file.flush()
        self._lock = threading.Lock()
        self._queue = queue.Queue()
        self._writer = None
        self._pinned = []
        self._error = None

    # ---- synchronous frame IO (host arrays) -------------------------------
    def _offset(self, frame):
        return self.header["header_size"] + (frame - self.header["seed_frame"] - 1) * self.header["frame_size"]

    def write_frame_host(self, frame, positions, velocities):
        """Write one frame from N x 4 float32 host arrays (the seed frame and
        earlier frames are not stored, solver/cache.py:325-326)."""
        n = self.header["particle_count"]
        pos = np.ascontiguousarray(positions, dtype="<f4").reshape(n, 4)
        vel = np.ascontiguousarray(velocities, dtype="<f4").reshape(n, 4)
        if frame <= self.header["seed_frame"]:
            return
        with self._lock:
            self.file.seek(self._offset(frame))
            self.file.write(pos.tobytes())
            self.file.write(vel.tobytes())
            self.header["last_frame"] = max(self.header["last_frame"], int(frame))
            self.file.seek(0)
            self.file.write(pack_header(self.header))
            self.file.flush()

    def load(self, frame):
        """(positions, velocities) as N x 4 float32 arrays for a frame inside
        (seed_frame, last_frame], else None (solver/cache.py:339-373)."""
        self.flush()
        h = self.header
        if not h["seed_frame"] < frame <= h["last_frame"]:
            return None
        n4 = h["particle_count"] * 4
        with self._lock:
            self.file.seek(self._offset(frame))
            raw = self.file.read(8 * n4)
        if len(raw) < 8 * n4:
            return None
        data = np.frombuffer(raw, dtype="<f4").astype(np.float32)
        return data[:n4].reshape(-1, 4), data[n4:].reshape(-1, 4)

    # ---- asynchronous frame writes from the GPU ---------------------------
    def write_frame(self, frame, pos, vel):
        """Queue one frame from device tensors (N x 4 float32): the D2H copy
        runs on a side stream into a pinned buffer and the writer thread
        stores it; the caller's stream only waits for the copy to start."""
        import torch

        if frame <= self.header["seed_frame"]:
            return
        if self._error is not None:
            raise self._error
        if self._writer is None:
            self._writer = threading.Thread(target=self._drain, daemon=True)
            self._writer.start()
        buf = self._take_buffer(pos, torch)
        side = torch.cuda.Stream(pos.device)
        side.wait_stream(torch.cuda.current_stream(pos.device))
        with torch.cuda.stream(side):
            buf[0].copy_(pos, non_blocking=True)
            buf[1].copy_(vel, non_blocking=True)
            done = torch.cuda.Event()
            done.record(side)
        pos.record_stream(side)
        vel.record_stream(side)
        self._queue.put((int(frame), buf, done))

    def _take_buffer(self, pos, torch):
        # reuse a pinned pair once the writer has released it
        for i, (busy, pair) in enumerate(self._pinned):
            if not busy.is_set():
                busy.set()
                return pair + (busy,)
        pair = (torch.empty(pos.shape, dtype=torch.float32, pin_memory=True),
                torch.empty(pos.shape, dtype=torch.float32, pin_memory=True))
        busy = threading.Event()
        busy.set()
        self._pinned.append((busy, pair))
        return pair + (busy,)

    def _drain(self):
        while True:
            item = self._queue.get()
            if item is None:
                self._queue.task_done()
                return
            frame, (pos, vel, busy), done = item
            try:
                done.synchronize()
                self.write_frame_host(frame, pos.numpy(), vel.numpy())
            except Exception as exc:  # surfaced by the next write_frame / flush
                self._error = exc
            finally:
                busy.clear()
                self._queue.task_done()

    def flush(self):
        """Wait until every queued frame is on disk."""
        if self._writer is not None:
            self._queue.join()
        if self._error is not None:
            raise self._error

    def close(self):
        self.flush()
        if self._writer is not None:
            self._queue.put(None)
            self._writer.join()
            self._writer = None
        self.file.close()
