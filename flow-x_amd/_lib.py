"""ctypes binding of libflowx_hip.so (include/flowx.h).

This is the boundary that replaces the reference's Blender GPU plumbing
(solver/gpu_util.py: make_texture / bind_push_constants / bind_image /
dispatch_1d / read_texture): device buffers are torch tensors on the ROCm
device, passed by data_ptr(); kernels are launched on torch's current stream.
Errors come back as return codes and become RuntimeError here, the way a
failed start becomes a reported warning in solver/sph.py:926-933.

The library is built in-tree (see build()); importing the package does not
build anything, and every GPU entry point raises if the library is missing.
"""

import ctypes
import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
LIB_PATH = PKG_DIR / "libflowx_hip.so"
CSRC = PKG_DIR / "csrc"
INCLUDE = REPO_DIR / "include"

HIP_SOURCES = ("fx_scan.hip", "fx_solver.hip", "fx_surface.hip", "fx_adjoint.hip")

HIPCC_FLAGS = (
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep every FP32 op scalar: no packed v_pk_*_f32 forms
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # IEEE float32 division and sqrt (the exact passes rely on it)
    "-fhip-fp32-correctly-rounded-divide-sqrt",
)

FX_OK = 0
PASS_IDS = {
    "normal": 0, "predict": 1, "grid_key": 2, "sort": 3, "cell_clear": 4, "cell_range": 5,
    "lambda": 6, "delta": 7, "apply_delta": 8, "velocity": 9, "xsph": 10, "finalize": 11,
}


class FxParams(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("cells", ctypes.c_int32 * 3),
        ("n_cells", ctypes.c_int32),
        ("lo", ctypes.c_float * 3),
        ("cell_size", ctypes.c_float),
        ("hi", ctypes.c_float * 3),
        ("particle_radius", ctypes.c_float),
        ("h", ctypes.c_float),
        ("mass", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("relaxation", ctypes.c_float),
        ("viscosity", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("boundary_damping", ctypes.c_float),
        ("scorr_k", ctypes.c_float),
        ("surface_tension", ctypes.c_float),
        ("collider_dims", ctypes.c_int32 * 3),
        ("collider_voxel", ctypes.c_float),
        ("iterations", ctypes.c_int32),
        ("max_cell_scan", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


class FxSurface(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32 * 3),
        ("lo", ctypes.c_float * 3),
        ("spacing", ctypes.c_float),
        ("kernel_radius", ctypes.c_float),
    ]


class FxTuning(ctypes.Structure):
    _fields_ = [("list_skin", ctypes.c_float), ("block", ctypes.c_int32), ("persistent", ctypes.c_int32),
                ("walk", ctypes.c_int32)]


class FxScene(ctypes.Structure):
    _fields_ = [("params", FxParams), ("collider_offset", ctypes.c_int64)]


class FxState(ctypes.Structure):
    _fields_ = [
        ("pos", ctypes.c_void_p),
        ("vel", ctypes.c_void_p),
        ("pred", ctypes.c_void_p),
        ("lambda_", ctypes.c_void_p),
        ("delta", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("key", ctypes.c_void_p),
        ("sorted_key", ctypes.c_void_p),
        ("sorted_idx", ctypes.c_void_p),
        ("cell_start", ctypes.c_void_p),
        ("cell_end", ctypes.c_void_p),
        ("collider", ctypes.c_void_p),
        ("work", ctypes.c_void_p),
        ("work_bytes", ctypes.c_size_t),
    ]


# entry point -> (restype, argtypes); every symbol include/flowx.h declares
_VP = ctypes.c_void_p
_I = ctypes.c_int
_PP = ctypes.POINTER(FxParams)
_SP = ctypes.POINTER(FxState)
SIGNATURES = {
    "fx_last_error": (ctypes.c_char_p, []),
    "fx_abi_version": (_I, []),
    "fx_profile": (_I, [_I]),
    "fx_profile_report": (_I, [ctypes.c_char_p, ctypes.c_size_t]),
    "fx_set_tuning": (_I, [ctypes.POINTER(FxTuning)]),
    "fx_get_tuning": (_I, [ctypes.POINTER(FxTuning)]),
    "fx_workspace_bytes": (ctypes.c_size_t, [_PP]),
    "fx_pass": (_I, [_I, _PP, _SP, _VP]),
    "fx_build_grid": (_I, [_PP, _SP, _VP]),
    "fx_substep": (_I, [_PP, _SP, _VP]),
    "fx_step": (_I, [_PP, _SP, _I, _VP]),
    "fx_load_state": (_I, [_PP, _SP, _VP]),
    "fx_splat": (_I, [_PP, ctypes.POINTER(FxSurface), _SP, _VP, _VP]),
    "fx_mc_workspace_bytes": (ctypes.c_size_t, [_VP]),
    "fx_mc_extract": (
        _I,
        [_VP, _VP, _VP, ctypes.c_double, ctypes.c_float, _VP, ctypes.c_int64, _VP, ctypes.c_int64, _VP, _VP,
         _VP, ctypes.c_size_t, _VP],
    ),
    "fx_mc_extract_static": (
        _I,
        [_VP, _VP, _VP, ctypes.c_double, ctypes.c_float, _VP, ctypes.c_int64, _VP, ctypes.c_int64, _VP, _VP, _VP],
    ),
    "fx_mc_tables": (_I, [_VP, _VP]),
    "fx_voxelize_boxes": (_I, [_VP, _VP, ctypes.c_float, _VP, _I, _VP, _I, _VP]),
    "fx_voxelize_sphere": (_I, [_VP, _VP, ctypes.c_float, _VP, ctypes.c_float, _VP, _I, _VP]),
    "fx_voxelize_cylinder": (
        _I, [_VP, _VP, ctypes.c_float, _VP, ctypes.c_float, ctypes.c_float, _VP, _I, _VP]
    ),
    "fx_mc_dedup_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int64]),
    "fx_mc_dedup": (_I, [_VP, ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, ctypes.c_size_t, _VP]),
    "fx_voxelize_mesh": (
        _I, [_VP, _VP, ctypes.c_float, _VP, _I, _VP, _I, _VP, _VP, ctypes.c_float, _VP, _I, _VP, _VP]
    ),
    "fx_batch_workspace_bytes": (ctypes.c_size_t, [_VP, _I]),
    "fx_batch_substep": (_I, [_VP, _I, _VP, _VP, _VP, _VP, ctypes.c_size_t, _VP]),
    "fx_batch_load_state": (_I, [_VP, _I, _VP, _VP, _VP, _VP, ctypes.c_size_t, _VP]),
    "fx_adjoint_workspace_bytes": (ctypes.c_size_t, [_PP, _I]),
    "fx_adjoint_record_bytes": (ctypes.c_size_t, [_PP]),
    "fx_adjoint_forward": (_I, [_PP, _I, ctypes.c_double, _VP, ctypes.c_double, _VP, _VP, _I, _VP, _VP, _VP, _VP,
                                _VP, ctypes.c_size_t, _VP]),
    "fx_adjoint_backward": (_I, [_PP, _I, ctypes.c_double, _VP, ctypes.c_double, _VP, _VP, _VP, _VP, _VP, _VP,
                                 _VP, _VP, _VP, _VP, ctypes.c_size_t, _VP]),
    "fx_adjoint_overflow": (_I, [_PP, _I, _VP, _VP, _VP]),
    "fx_adjoint_splat": (_I, [_PP, ctypes.POINTER(FxSurface), _I, _VP, _VP, _VP, _VP, ctypes.c_size_t, _VP]),
    "fx_adjoint_splat_backward": (_I, [_PP, ctypes.POINTER(FxSurface), _I, _VP, _VP, _VP, _VP, _VP]),
}

_lib = None


def build(verbose=False, out=None, defines=()):
    """Compile libflowx_hip.so for gfx950 in-tree (hipcc, no GPU needed)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    sources = [str(CSRC / s) for s in HIP_SOURCES if (CSRC / s).exists()]
    target = Path(out) if out else LIB_PATH
    tmp = str(target) + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", *[f"-D{d}" for d in defines], "-o", tmp, *sources]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-6000:])
    if verbose and res.stderr:
        print(res.stderr[-2000:])
    os.replace(tmp, target)
    return target


def lib():
    """The loaded library.  Raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("FLOWX_LIB", LIB_PATH))
        if not path.exists():
            raise RuntimeError(
                f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
            )
        L = ctypes.CDLL(str(path))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return [name for name in SIGNATURES if hasattr(L, name)]


def check(rc, what):
    if rc != FX_OK:
        msg = lib().fx_last_error()
        msg = msg.decode("utf-8", "replace") if msg else ""
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def set_tuning(list_skin=None, block=None, persistent=None, walk=None):
    t = FxTuning()
    check(lib().fx_get_tuning(ctypes.byref(t)), "fx_get_tuning")
    if list_skin is not None:
        t.list_skin = list_skin
    if block is not None:
        t.block = block
    if persistent is not None:
        t.persistent = persistent
    if walk is not None:
        t.walk = walk
    check(lib().fx_set_tuning(ctypes.byref(t)), "fx_set_tuning")


def get_tuning():
    t = FxTuning()
    check(lib().fx_get_tuning(ctypes.byref(t)), "fx_get_tuning")
    return {"list_skin": t.list_skin, "block": t.block, "persistent": t.persistent, "walk": t.walk}


def profile(enable):
    check(lib().fx_profile(1 if enable else 0), "fx_profile")


def profile_report():
    """{kernel: (launches, total_ms)} since the last report."""
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().fx_profile_report(buf, len(buf)), "fx_profile_report")
    out = {}
    for line in buf.value.decode().splitlines():
        name, count, ms = line.rsplit(" ", 2)
        out[name] = (int(count), float(ms))
    return out
