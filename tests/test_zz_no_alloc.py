"""fx_mc_extract_static - marching cubes with its scratch in the library's
fixed device array instead of a caller workspace - against fx_mc_extract on
the same fields, through the C ABI: two calls in a row, the second on a
smaller lattice (the array keeps the first call's contents), and a lattice
above the array's limit refused before any launch."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(n, seed):
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1)[..., ::-1].astype(np.float64)
    f = np.zeros((n, n, n))
    for _ in range(4):
        c = rng.uniform(0.25 * n, 0.75 * n, 3)
        f = np.maximum(f, 1 - np.linalg.norm(g - c, axis=-1) / rng.uniform(0.15 * n, 0.3 * n))
    f += rng.normal(0, 0.02, f.shape)  # ragged surface: many cases
    return f.astype(np.float32).reshape(-1)


def _extract(fx, field, n, static):
    import torch

    L = fx._lib
    d = (ctypes.c_int32 * 3)(n, n, n)
    o = (ctypes.c_double * 3)(-1.0, -0.5, 0.25)
    f = torch.as_tensor(field).cuda()
    cap = 6 * n * n * n // 4 + 64
    verts = torch.full((cap, 3), -7.0, dtype=torch.float32, device="cuda")
    tris = torch.full((cap, 3), -7, dtype=torch.int32, device="cuda")
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    cases = torch.zeros((n - 1) ** 3, dtype=torch.uint8, device="cuda")
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    args = [P(f), d, o, ctypes.c_double(0.01), ctypes.c_float(0.5), P(verts), ctypes.c_int64(cap), P(tris),
            ctypes.c_int64(cap), P(counts), P(cases)]
    if static:
        rc = L.lib().fx_mc_extract_static(*args, stream)
    else:
        work = torch.empty(int(L.lib().fx_mc_workspace_bytes(d)), dtype=torch.uint8, device="cuda")
        rc = L.lib().fx_mc_extract(*args, P(work), ctypes.c_size_t(work.numel()), stream)
    L.check(rc, "marching cubes")
    torch.cuda.synchronize()
    nv, nt = (int(c) for c in counts.cpu())
    assert nv <= cap and nt <= cap
    return verts[:nv].cpu().numpy(), tris[:nt].cpu().numpy(), cases.cpu().numpy()


def test_static_scratch_equals_workspace_path(fx):
    for n, seed in ((120, 1), (57, 2)):  # 1.73M samples, then a smaller lattice
        field = _field(n, seed)
        a = _extract(fx, field, n, static=False)
        b = _extract(fx, field, n, static=True)
        print(f"{n}^3: {len(a[0])} vertices, {len(a[1])} triangles")
        assert len(a[1]) > 1000
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_static_scratch_refuses_a_lattice_above_its_limit(fx):
    L = fx._lib
    d = (ctypes.c_int32 * 3)(129, 128, 128)  # 2^21 + 16384 samples
    o = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    counts = ctypes.c_void_p(1)  # never dereferenced: the size check comes first
    rc = L.lib().fx_mc_extract_static(ctypes.c_void_p(1), d, o, ctypes.c_double(1.0), ctypes.c_float(0.5), None,
                                      ctypes.c_int64(0), None, ctypes.c_int64(0), counts, None, None)
    assert rc == -4 and b"fixed scratch" in L.lib().fx_last_error()
