"""Write the marching-cubes golden vectors from the reference module.

The exact per-case triangulation of solver/marching_cubes.py (TRI_TABLE,
built in dict-insertion order at :122-162) and the vertex / triangle order of
extract() (:191-280) are pinned by no closed form, so this script loads that
one stdlib-only module by path - as scripts/test_marching_cubes.py:19-25 does
- and records its tables and its output on small lattices.  It runs only in
the container that holds the reference; the GPU tests read the .npz it wrote.

    python tests/golden/make_mc_golden.py /path/to/reference
"""

import importlib.util
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _load(reference_root):
    source = Path(reference_root) / "solver" / "marching_cubes.py"
    spec = importlib.util.spec_from_file_location("flowx_reference_mc", source)
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


def _sample(dims, origin, spacing, fn):
    nx, ny, nz = dims
    out = []
    for k in range(nz):
        z = origin[2] + k * spacing
        for j in range(ny):
            y = origin[1] + j * spacing
            for i in range(nx):
                out.append(fn(origin[0] + i * spacing, y, z))
    # the reference reads an R32F texture back, so fields are float32 values
    return np.asarray(out, dtype=np.float32)


def _noise_field(dims, spacing):
    # the splitmix-style hash of scripts/test_marching_cubes.py:162-177
    nx, ny, nz = dims
    mask = (1 << 64) - 1

    def noise(x, y, z):
        i = int(round(x / spacing))
        j = int(round(y / spacing))
        k = int(round(z / spacing))
        if i in (0, nx - 1) or j in (0, ny - 1) or k in (0, nz - 1):
            return 0.0
        h = (i * 0x9E3779B97F4A7C15 + j * 0xBF58476D1CE4E5B9 + k * 0x94D049BB133111EB) & mask
        h ^= h >> 30
        h = (h * 0xBF58476D1CE4E5B9) & mask
        h ^= h >> 27
        h = (h * 0x94D049BB133111EB) & mask
        h ^= h >> 31
        return h / mask

    return _sample(dims, (0.0, 0.0, 0.0), spacing, noise)


def main(reference_root):
    mc = _load(reference_root)
    tri = np.full((256, 16), -1, dtype=np.int8)
    for case, triangles in enumerate(mc.TRI_TABLE):
        flat = [e for t in triangles for e in t]
        tri[case, : len(flat)] = flat
    edge_table = np.asarray(mc.EDGE_TABLE, dtype=np.uint16)
    walk = np.asarray(mc._EDGE_WALK, dtype=np.int32)

    cases = {}
    # sphere of scripts/test_marching_cubes.py:118-139
    sphere = _sample(
        (41, 41, 41), (-1.0, -1.0, -1.0), 0.05, lambda x, y, z: 0.7 - math.sqrt(x * x + y * y + z * z)
    )
    cases["sphere"] = (sphere, (41, 41, 41), (-1.0, -1.0, -1.0), 0.05, 0.0)
    cases["noise"] = (_noise_field((18, 18, 18), 0.25), (18, 18, 18), (0.0, 0.0, 0.0), 0.25, 0.5)
    rng = np.random.Generator(np.random.PCG64(7))
    small = rng.random(9 * 7 * 6).astype(np.float32)
    small[0] = small[1]  # an edge with equal end values (t = 0.5 branch needs equal values
    small[2] = 0.5       # across iso: put one sample exactly at iso)
    cases["random"] = (small, (9, 7, 6), (-0.3, 0.2, 1.5), 0.125, 0.5)

    payload = {"tri_table": tri, "edge_table": edge_table, "edge_walk": walk}
    for name, (field, dims, origin, spacing, iso) in cases.items():
        verts, tris = mc.extract([float(v) for v in field], dims, origin, spacing, iso)
        payload[f"{name}_field"] = field
        payload[f"{name}_dims"] = np.asarray(dims, dtype=np.int32)
        payload[f"{name}_origin"] = np.asarray(origin, dtype=np.float64)
        payload[f"{name}_spacing"] = np.float64(spacing)
        payload[f"{name}_iso"] = np.float64(iso)
        payload[f"{name}_verts"] = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
        payload[f"{name}_tris"] = np.asarray(tris, dtype=np.int32).reshape(-1, 3)
    np.savez_compressed(HERE / "mc_golden.npz", **payload)
    print("wrote", HERE / "mc_golden.npz", {k: v.shape for k, v in payload.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
