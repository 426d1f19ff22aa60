"""ctypes wrapper of the C++ oracle ``liblic_oracle.so`` (test infrastructure).

Functions restated (``lic_oracle.cpp``), with what pins each one:

* ``time_to_next_pixel`` (reference ``src/lib.rs:157-180``): the
  reference's known-answer tests ``src/lib.rs:187-205``, hand-derived closed
  forms for +-0 / NaN / inf velocities per profile, and the exact-arithmetic
  restatement in :mod:`oracle.pyref` on random inputs (tests/test_oracle.py).
* ``advance`` (``src/lib.rs:235-273``, with ``update_state`` :208-233 and the
  boundary rule :83-115): the reference's known-answer test
  ``src/lib.rs:279-296``, hand-derived single steps at every edge, and
  :mod:`oracle.pyref`.
* ``convolve`` (``src/lib.rs:304-443``): closed forms (uniform rows/columns,
  zero field, NaN field, unit half-plane fields walked by an exact geometric
  stepper), the invariants of the reference's ``tests/test_convolution.py``,
  and :mod:`oracle.pyref` on every dtype x profile x mode x boundary combination.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liblic_oracle.so"

PROFILES = {"none": 0, "fma": 1, "branchless": 2, "fma+branchless": 3}
_BOUND = {"closed": 0, "periodic": 1}
_MODE = {"velocity": 0, "polarization": 1}

_lib = None


def build() -> Path:
    """Compile the oracle with its Makefile (always rebuilds)."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE), "liblic_oracle.so"], check=True)
    return LIB


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.is_file():
        build()
    lib = ctypes.CDLL(str(LIB))
    i64, vp, c_int = ctypes.c_int64, ctypes.c_void_p, ctypes.c_int
    conv = [vp, i64, i64, vp, i64, i64, vp, i64, i64, i64, i64, vp, i64, c_int, vp, i64,
            c_int, vp, vp, i64, i64, c_int]
    for name in ("lic_oracle_convolve_f64", "lic_oracle_convolve_f32"):
        fn = getattr(lib, name)
        fn.argtypes = conv
        fn.restype = c_int
    lib.lic_oracle_time_to_next_pixel_f64.argtypes = [ctypes.c_double, ctypes.c_double, c_int]
    lib.lic_oracle_time_to_next_pixel_f64.restype = ctypes.c_double
    lib.lic_oracle_time_to_next_pixel_f32.argtypes = [ctypes.c_float, ctypes.c_float, c_int]
    lib.lic_oracle_time_to_next_pixel_f32.restype = ctypes.c_float
    for name, ft in (("lic_oracle_advance_f64", ctypes.c_double),
                     ("lic_oracle_advance_f32", ctypes.c_float)):
        fn = getattr(lib, name)
        fn.argtypes = [ft, ft, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(ft),
                       ctypes.POINTER(ft), i64, i64, vp, c_int]
        fn.restype = None
    lib.lic_oracle_max_threads.restype = c_int
    _lib = lib
    return lib


def _profile(profile: int | str) -> int:
    return PROFILES[profile] if isinstance(profile, str) else int(profile)


def _bounds(boundaries) -> np.ndarray:
    if isinstance(boundaries, str):
        boundaries = {"x": boundaries, "y": boundaries}
    if isinstance(boundaries, dict):
        bx, by = boundaries["x"], boundaries["y"]
        bx = (bx, bx) if isinstance(bx, str) else tuple(bx)
        by = (by, by) if isinstance(by, str) else tuple(by)
    else:
        bx, by = boundaries
    return np.array([_BOUND[bx[0]], _BOUND[bx[1]], _BOUND[by[0]], _BOUND[by[1]]], dtype=np.intc)


def convolve(
    texture: np.ndarray,
    u: np.ndarray,
    v: np.ndarray,
    *,
    kernel: np.ndarray,
    uv_mode: str = "velocity",
    boundaries="closed",
    iterations: int = 1,
    profile: int | str = "fma+branchless",
    rows: tuple[int, int] | None = None,
    threads: int = 1,
) -> np.ndarray:
    """Oracle convolution (no input validation beyond dtype/shape).

    ``rows=(r0, r1)`` computes only those rows of the last pass (the others
    are left as zeros); used to time bounded CPU samples.
    """
    dtype = texture.dtype
    assert dtype in (np.float32, np.float64)
    assert u.dtype == dtype and v.dtype == dtype and kernel.dtype == dtype
    assert texture.shape == u.shape == v.shape and texture.ndim == 2
    lib = load()
    fn = lib.lic_oracle_convolve_f64 if dtype == np.float64 else lib.lic_oracle_convolve_f32
    ny, nx = texture.shape
    item = dtype.itemsize
    k = np.ascontiguousarray(kernel)
    out = np.zeros((ny, nx), dtype=dtype)
    work = np.empty((ny, nx), dtype=dtype)
    b = _bounds(boundaries)
    r0, r1 = rows if rows is not None else (0, ny)
    status = fn(
        texture.ctypes.data, texture.strides[0] // item, texture.strides[1] // item,
        u.ctypes.data, u.strides[0] // item, u.strides[1] // item,
        v.ctypes.data, v.strides[0] // item, v.strides[1] // item,
        ny, nx, k.ctypes.data, k.shape[0], _MODE[uv_mode], b.ctypes.data, iterations,
        _profile(profile), out.ctypes.data, work.ctypes.data, r0, r1, threads,
    )
    if status == -2:
        raise ValueError("empty kernel")
    if status != 0:
        raise ValueError(f"oracle status {status}")
    return out


def time_to_next_pixel(velocity, frac, profile: int | str = "fma+branchless", dtype=np.float64):
    lib = load()
    if np.dtype(dtype) == np.float32:
        return np.float32(lib.lic_oracle_time_to_next_pixel_f32(float(velocity), float(frac),
                                                                _profile(profile)))
    return np.float64(lib.lic_oracle_time_to_next_pixel_f64(float(velocity), float(frac),
                                                            _profile(profile)))


def advance(u, v, i, j, fx, fy, *, shape, boundaries="closed",
            profile: int | str = "fma+branchless", dtype=np.float64):
    """One streamline step; returns (i, j, fx, fy).  Coordinates are ints."""
    lib = load()
    f32 = np.dtype(dtype) == np.float32
    ft = ctypes.c_float if f32 else ctypes.c_double
    fn = lib.lic_oracle_advance_f32 if f32 else lib.lic_oracle_advance_f64
    ci, cj = ctypes.c_int64(i), ctypes.c_int64(j)
    cfx, cfy = ft(fx), ft(fy)
    b = _bounds(boundaries)
    fn(float(u), float(v), ctypes.byref(ci), ctypes.byref(cj), ctypes.byref(cfx),
       ctypes.byref(cfy), shape[0], shape[1], b.ctypes.data, _profile(profile))
    cast = np.float32 if f32 else np.float64
    return ci.value, cj.value, cast(cfx.value), cast(cfy.value)


def max_threads() -> int:
    return int(load().lic_oracle_max_threads())
