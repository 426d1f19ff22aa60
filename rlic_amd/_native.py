"""Loader of the native library ``librlic_hip.so`` (HIP kernels + C ABI).

The library is built in-tree (``rlic_amd/_build.py``) and loaded from this
package directory.  There is no fallback: if the library is missing or does
not load, importing the compute entry points fails with an ImportError that
says how to build it.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_NAME = "librlic_hip.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

# Entry points declared in include/rlic_hip.h, with their ctypes signatures.
_i64 = ctypes.c_int64
_int = ctypes.c_int
_vp = ctypes.c_void_p

_CONVOLVE_ARGTYPES = [
    _vp, _i64, _i64,  # texture, sy, sx
    _vp, _i64, _i64,  # u
    _vp, _i64, _i64,  # v
    _i64, _i64,  # ny, nx
    _vp, _i64,  # kernel, kernel_len
    _int,  # uv_mode
    _int, _int, _int, _int,  # x_left, x_right, y_left, y_right
    _i64,  # iterations
    _vp,  # out
    _int, _int, _vp,  # location, profile, stream
]

def _field_struct(name: str, ctype) -> type:
    """ctypes mirror of rlic_field_f32 / rlic_field_f64 (include/rlic_hip.h)."""
    ptr = ctypes.POINTER(ctype)
    return type(name, (ctypes.Structure,), {"_fields_": [
        ("texture", ptr), ("tex_sy", _i64), ("tex_sx", _i64),
        ("u", ptr), ("u_sy", _i64), ("u_sx", _i64),
        ("v", ptr), ("v_sy", _i64), ("v_sx", _i64),
        ("ny", _i64), ("nx", _i64),
        ("kernel", ptr), ("kernel_len", _i64),
        ("uv_mode", _int), ("x_left", _int), ("x_right", _int), ("y_left", _int),
        ("y_right", _int),
        ("iterations", _i64),
        ("out", ptr),
    ]})


FieldF32 = _field_struct("FieldF32", ctypes.c_float)
FieldF64 = _field_struct("FieldF64", ctypes.c_double)

# The same layout as a NumPy structured dtype (pointers as addresses), for
# filling many descriptors at once.
FIELD_DTYPE = np.dtype([
    ("texture", "<u8"), ("tex_sy", "<i8"), ("tex_sx", "<i8"),
    ("u", "<u8"), ("u_sy", "<i8"), ("u_sx", "<i8"),
    ("v", "<u8"), ("v_sy", "<i8"), ("v_sx", "<i8"),
    ("ny", "<i8"), ("nx", "<i8"),
    ("kernel", "<u8"), ("kernel_len", "<i8"),
    ("uv_mode", "<i4"), ("x_left", "<i4"), ("x_right", "<i4"), ("y_left", "<i4"),
    ("y_right", "<i4"),
    ("iterations", "<i8"),
    ("out", "<u8"),
], align=True)
assert FIELD_DTYPE.itemsize == ctypes.sizeof(FieldF32) == ctypes.sizeof(FieldF64)
assert all(FIELD_DTYPE.fields[name][1] == getattr(FieldF32, name).offset
           for name, _ in FieldF32._fields_)

_i64p = ctypes.POINTER(ctypes.c_int64)
_intp = ctypes.POINTER(ctypes.c_int)

SIGNATURES: dict[str, tuple[list, object]] = {
    "rlic_server_create": ([_i64, _int, _vp], _vp),
    "rlic_server_destroy": ([_vp], None),
    "rlic_server_admit_f32": ([_vp, ctypes.POINTER(FieldF32), _int, _i64p], _int),
    "rlic_server_admit_f64": ([_vp, ctypes.POINTER(FieldF64), _int, _i64p], _int),
    "rlic_server_field_bytes_f32": ([_i64, _i64, _i64, _int, _i64], _i64),
    "rlic_server_field_bytes_f64": ([_i64, _i64, _i64, _int, _i64], _i64),
    "rlic_server_step": ([_vp, _i64p, _i64, _i64p], _int),
    "rlic_server_extend": ([_vp, _i64, _i64], _int),
    "rlic_server_read_f32": ([_vp, _i64, _vp], _int),
    "rlic_server_read_f64": ([_vp, _i64, _vp], _int),
    "rlic_server_release": ([_vp, _i64], _int),
    "rlic_server_usage": ([_vp, _i64p, _i64p, _i64p], _int),
    "rlic_convolve_batch_f32": ([ctypes.POINTER(FieldF32), _i64, _int, _vp], _int),
    "rlic_convolve_batch_f64": ([ctypes.POINTER(FieldF64), _i64, _int, _vp], _int),
    "rlic_batch_check_values_f32": ([ctypes.POINTER(FieldF32), _i64, _i64p, _intp, _vp], _int),
    "rlic_batch_check_values_f64": ([ctypes.POINTER(FieldF64), _i64, _i64p, _intp, _vp], _int),
    "rlic_check_values_f32": ([_vp, _i64, _i64, _i64, _i64, _vp, _i64, _intp, _vp], _int),
    "rlic_check_values_f64": ([_vp, _i64, _i64, _i64, _i64, _vp, _i64, _intp, _vp], _int),
    "rlic_convolve_f32": (_CONVOLVE_ARGTYPES, _int),
    "rlic_convolve_f64": (_CONVOLVE_ARGTYPES, _int),
    # the allocation-free variants take no `location` (device arrays only)
    "rlic_convolve_noalloc_f32": (_CONVOLVE_ARGTYPES[:20] + _CONVOLVE_ARGTYPES[21:], _int),
    "rlic_convolve_noalloc_f64": (_CONVOLVE_ARGTYPES[:20] + _CONVOLVE_ARGTYPES[21:], _int),
    "rlic_equalize_histogram_f32": ([_vp, _i64, _i64, _i64, _i64, _i64, _vp, _int, _vp], _int),
    "rlic_equalize_histogram_f64": ([_vp, _i64, _i64, _i64, _i64, _i64, _vp, _int, _vp], _int),
    "rlic_equalize_histogram_sliding_tile_f32":
        ([_vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _int, _vp], _int),
    "rlic_equalize_histogram_sliding_tile_f64":
        ([_vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _int, _vp], _int),
    "rlic_equalize_histogram_tile_interpolation_f32":
        ([_vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _int, _vp], _int),
    "rlic_equalize_histogram_tile_interpolation_f64":
        ([_vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _int, _vp], _int),
    "rlic_last_error": ([], ctypes.c_char_p),
    "rlic_trace_stats_read": ([_i64p, _i64p], _int),
    "rlic_autotune_read": ([ctypes.c_char_p, _i64], _i64),
    "rlic_autotune_clear": ([], _int),
    "rlic_autotune_load": ([ctypes.c_char_p], _i64),
    "rlic_div_check_f32": ([_vp, _vp, _vp, _i64, _vp], _int),
    "rlic_scratch_release": ([], _int),
    "rlic_profiler_enable": ([_int], _int),
    "rlic_profiler_read": ([ctypes.c_char_p, _i64], _i64),
    "rlic_fill_noise_f32": ([_vp, _i64, _i64, ctypes.c_uint64, _vp], _int),
    "rlic_fill_noise_f64": ([_vp, _i64, _i64, ctypes.c_uint64, _vp], _int),
    "rlic_fill_modes_f32": ([_vp, _i64, _i64, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp], _int),
    "rlic_fill_modes_f64": ([_vp, _i64, _i64, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp], _int),
    "rlic_version": ([], ctypes.c_char_p),
}

# Error codes (include/rlic_hip.h)
RLIC_OK = 0
RLIC_ERR_INVALID_ARGUMENT = -1
RLIC_ERR_OUT_OF_MEMORY = -2
RLIC_ERR_HIP = -3
RLIC_ERR_EMPTY_KERNEL = -4

RLIC_HOST = 0
RLIC_DEVICE = 1

_lib: ctypes.CDLL | None = None


def _share_torch_hip_runtime() -> None:
    """Make this library and PyTorch-ROCm use one HIP runtime in the process.

    PyTorch ships its own ``libamdhip64.so`` (soname ``libamdhip64.so.7``) and
    loads it by file name from its ``lib`` directory.  If the system copy were
    loaded first (as this library's dependency), PyTorch would then load a
    second HIP runtime next to it and find no device.  Loading PyTorch's copy
    first (when PyTorch is installed) lets this library's dependency resolve
    to it by soname, and PyTorch later finds its own file already loaded.
    ``RLIC_AMD_SYSTEM_HIP=1`` keeps the system runtime instead.
    """
    if os.environ.get("RLIC_AMD_SYSTEM_HIP") == "1":
        return
    import importlib.util
    import sys

    if "torch" in sys.modules:
        return  # PyTorch is loaded: its runtime is the process's runtime already
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    candidate = Path(spec.origin).parent / "lib" / "libamdhip64.so"
    if candidate.is_file():
        try:
            ctypes.CDLL(str(candidate), mode=ctypes.RTLD_GLOBAL)
        except OSError:  # pragma: no cover - broken PyTorch install
            pass


def load() -> ctypes.CDLL:
    """Load (once) and return the native library, with typed signatures."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("RLIC_AMD_LIB", LIB_PATH))
    if not path.is_file():
        raise ImportError(
            f"{path} is missing: the HIP extension has not been built. "
            "Build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m rlic_amd._build`."
        )
    _share_torch_hip_runtime()
    try:
        lib = ctypes.CDLL(str(path))
    except OSError as exc:  # pragma: no cover - depends on the machine
        raise ImportError(f"cannot load {path}: {exc}") from exc
    for name, (argtypes, restype) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().rlic_last_error()
    return msg.decode() if msg else ""


def profiler_enable(on: bool = True) -> None:
    """Record hipEvents around every kernel launch of the library."""
    load().rlic_profiler_enable(1 if on else 0)


def profiler_read() -> dict:
    """Per-kernel launch count and total milliseconds since the last read."""
    import json

    lib = load()
    need = lib.rlic_profiler_read(None, 0)
    buf = ctypes.create_string_buffer(int(need) + 256)
    lib.rlic_profiler_read(buf, len(buf))
    return json.loads(buf.value.decode())


def autotune_table() -> list:
    """The replay autotuner's table (one entry per tuned key)."""
    import json

    lib = load()
    need = lib.rlic_autotune_read(None, 0)
    buf = ctypes.create_string_buffer(int(need) + 256)
    lib.rlic_autotune_read(buf, len(buf))
    return json.loads(buf.value.decode())


def autotune_clear() -> None:
    load().rlic_autotune_clear()


def autotune_load(path) -> int:
    """Add the plan lines of a text file to the autotuner's table."""
    n = load().rlic_autotune_load(str(path).encode())
    if n < 0:
        raise_for_status(int(n))
    return int(n)


def trace_stats_read() -> dict:
    """Live and executed lane-steps of the tracing passes run with
    RLIC_TRACE_STATS=1 since the last read (then reset)."""
    live, executed = ctypes.c_int64(), ctypes.c_int64()
    load().rlic_trace_stats_read(ctypes.byref(live), ctypes.byref(executed))
    eff = live.value / executed.value if executed.value else None
    return {"live_lane_steps": live.value, "executed_lane_steps": executed.value,
            "lane_efficiency": eff}


def raise_for_status(code: int) -> None:
    """Map a negative C-ABI status to a Python exception."""
    if code == RLIC_OK:
        return
    msg = last_error()
    if code == RLIC_ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    if code in (RLIC_ERR_INVALID_ARGUMENT, RLIC_ERR_EMPTY_KERNEL):
        raise ValueError(msg)
    raise RuntimeError(f"HIP failure in librlic_hip: {msg} (status {code})")
