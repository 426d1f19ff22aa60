"""Native entry points with the reference's FFI signature.

``convolve_f32`` / ``convolve_f64`` take the same arguments as the
reference's PyO3 functions (``src/lib.rs:450-482``, stub
``src/rlic/_core.pyi:8-29``)::

    convolve_f32(texture, (u, v, uv_mode), kernel, ((xl, xr), (yl, yr)), iterations)

and call ``rlic_convolve_f32`` / ``rlic_convolve_f64`` of ``librlic_hip.so``
through ctypes (which releases the GIL for the duration of the call).

Arrays may be NumPy arrays (any strides; uploaded, computed on the GPU and
downloaded; a new NumPy array is returned) or PyTorch tensors on a ROCm
device (computed in place on the tensors' device and the current stream;
a new tensor is returned without synchronising, as torch operations do).
CPU PyTorch tensors are treated as NumPy arrays and a CPU tensor returned.

Argument errors raise TypeError/OverflowError/ValueError as the reference's
argument extraction does.  An empty kernel with a non-empty image raises
ValueError (the reference aborts the process there, src/lib.rs:382).
"""

from __future__ import annotations

__all__ = ["ARITHMETIC_PROFILES", "convolve_f32", "convolve_f64", "equalize_histogram_f32",
           "equalize_histogram_f64", "equalize_histogram_sliding_tile_f32",
           "equalize_histogram_sliding_tile_f64", "equalize_histogram_tile_interpolation_f32",
           "equalize_histogram_tile_interpolation_f64"]

import ctypes
from typing import Any

import numpy as np

from rlic_amd import _native

_UV_MODES = {"velocity": 0, "polarization": 1}
_BOUNDS = {"closed": 0, "periodic": 1}

#: reference arithmetic profiles (Cargo features, Cargo.toml:27-30)
ARITHMETIC_PROFILES = {
    "none": 0,
    "fma": 1,
    "branchless": 2,
    "fma+branchless": 3,  # Cargo default (Cargo.toml:28)
}
DEFAULT_PROFILE = ARITHMETIC_PROFILES["fma+branchless"]

_I64_MIN, _I64_MAX = -(2**63), 2**63 - 1


def _is_torch(x: Any) -> bool:
    return type(x).__module__.split(".", 1)[0] == "torch"


def _check_iterations(iterations: Any) -> int:
    try:
        it = int(iterations.__index__())
    except AttributeError:
        raise TypeError(
            f"argument 'iterations': {type(iterations).__name__!r} object cannot be "
            "interpreted as an integer"
        ) from None
    if not (_I64_MIN <= it <= _I64_MAX):
        raise OverflowError("argument 'iterations': int too big to convert")
    return it


def _mode_code(uv_mode: Any) -> int:
    if not isinstance(uv_mode, str):
        raise TypeError(f"argument 'uv': uv_mode must be str, not {type(uv_mode).__name__}")
    if uv_mode not in _UV_MODES:
        raise ValueError(f"unknown uv_mode {uv_mode!r}")
    return _UV_MODES[uv_mode]


def _bound_codes(boundaries: Any) -> tuple[int, int, int, int]:
    try:
        (xl, xr), (yl, yr) = boundaries
    except (TypeError, ValueError):
        raise TypeError(
            "argument 'boundaries': expected ((x_left, x_right), (y_left, y_right))"
        ) from None
    codes = []
    for name in (xl, xr, yl, yr):
        if not isinstance(name, str):
            raise TypeError(f"argument 'boundaries': expected str, got {type(name).__name__}")
        if name not in _BOUNDS:
            raise ValueError(f"unknown boundary {name!r}")
        codes.append(_BOUNDS[name])
    return codes[0], codes[1], codes[2], codes[3]


def _numpy_view(arr: Any, name: str, dtype: np.dtype, ndim: int) -> tuple[np.ndarray, list[int]]:
    """Validate a host array; return it (aligned) with its element strides."""
    if not isinstance(arr, np.ndarray):
        raise TypeError(f"argument {name!r}: expected a numpy.ndarray, got {type(arr).__name__}")
    if arr.dtype != dtype:
        raise TypeError(f"argument {name!r}: type mismatch: expected {dtype}, got {arr.dtype}")
    if arr.ndim != ndim:
        raise TypeError(f"argument {name!r}: dimensionality mismatch: expected {ndim}, got {arr.ndim}")
    item = arr.itemsize
    if arr.ctypes.data % item or any(s % item for s in arr.strides):
        arr = np.ascontiguousarray(arr)  # unaligned view: re-lay it out
    return arr, [s // item for s in arr.strides]


def _convolve(
    dtype: np.dtype,
    entry: str,
    texture: Any,
    uv: Any,
    kernel: Any,
    boundaries: Any,
    iterations: Any,
    profile: int | str | None,
) -> Any:
    try:
        u, v, uv_mode = uv
    except (TypeError, ValueError):
        raise TypeError("argument 'uv': expected a (u, v, uv_mode) tuple") from None
    mode = _mode_code(uv_mode)
    xl, xr, yl, yr = _bound_codes(boundaries)
    it = _check_iterations(iterations)
    if profile is None:
        prof = DEFAULT_PROFILE
    elif isinstance(profile, str):
        prof = ARITHMETIC_PROFILES[profile]
    else:
        prof = int(profile)

    lib = _native.load()
    fn = getattr(lib, entry)
    arrays = (texture, u, v, kernel)
    if any(_is_torch(a) for a in arrays):
        return _convolve_torch(fn, dtype, arrays, mode, (xl, xr, yl, yr), it, prof)

    return _convolve_host_numpy(fn, dtype, arrays, mode, (xl, xr, yl, yr), it, prof)


def _convolve_torch(fn, dtype, arrays, mode, bounds, it, prof):
    import torch

    tdtype = {np.dtype("float32"): torch.float32, np.dtype("float64"): torch.float64}[dtype]
    if not all(_is_torch(a) for a in arrays):
        raise TypeError("texture, u, v and kernel must all be torch tensors, or none of them")
    texture, u, v, kernel = arrays
    names = ("texture", "u", "v", "kernel")
    for name, a, nd in zip(names, arrays, (2, 2, 2, 1)):
        if a.dtype != tdtype:
            raise TypeError(f"argument {name!r}: type mismatch: expected {tdtype}, got {a.dtype}")
        if a.dim() != nd:
            raise TypeError(f"argument {name!r}: dimensionality mismatch: expected {nd}, got {a.dim()}")
    devices = {a.device for a in arrays}
    if len(devices) != 1:
        raise ValueError(f"all tensors must be on one device, got {sorted(map(str, devices))}")
    device = texture.device
    if u.shape != texture.shape or v.shape != texture.shape:
        raise ValueError(
            "texture, u and v must share one shape; got "
            f"{tuple(texture.shape)}, {tuple(u.shape)}, {tuple(v.shape)}"
        )
    if device.type == "cpu":
        out = _convolve_host_numpy(fn, dtype, [a.numpy() for a in arrays], mode, bounds, it, prof)
        return torch.from_numpy(out)
    kernel = kernel.contiguous()
    ny, nx = texture.shape
    with torch.cuda.device(device):
        out = torch.empty((ny, nx), dtype=tdtype, device=device)
        stream = torch.cuda.current_stream(device).cuda_stream
        status = fn(
            texture.data_ptr(), texture.stride(0), texture.stride(1),
            u.data_ptr(), u.stride(0), u.stride(1),
            v.data_ptr(), v.stride(0), v.stride(1),
            ny, nx,
            kernel.data_ptr(), kernel.shape[0],
            mode, *bounds,
            it,
            out.data_ptr(),
            _native.RLIC_DEVICE, prof, ctypes.c_void_p(stream),
        )
    _native.raise_for_status(status)
    return out


def _convolve_host_numpy(fn, dtype, arrays, mode, bounds, it, prof):
    tex, ts = _numpy_view(arrays[0], "texture", dtype, 2)
    uu, us = _numpy_view(arrays[1], "u", dtype, 2)
    vv, vs = _numpy_view(arrays[2], "v", dtype, 2)
    k = np.ascontiguousarray(_numpy_view(arrays[3], "kernel", dtype, 1)[0])
    if uu.shape != tex.shape or vv.shape != tex.shape:
        raise ValueError(
            f"texture, u and v must share one shape; got {tex.shape}, {uu.shape}, {vv.shape}"
        )
    ny, nx = tex.shape
    out = np.empty((ny, nx), dtype=dtype)
    status = fn(
        tex.ctypes.data, ts[0], ts[1], uu.ctypes.data, us[0], us[1],
        vv.ctypes.data, vs[0], vs[1], ny, nx, k.ctypes.data, k.shape[0],
        mode, *bounds, it, out.ctypes.data, _native.RLIC_HOST, prof, None,
    )
    _native.raise_for_status(status)
    return out


def convolve_f32(texture, uv, kernel, boundaries, iterations=1, *, profile=None):
    """float32 line integral convolution (reference ``src/lib.rs:450-465``)."""
    return _convolve(
        np.dtype("float32"), "rlic_convolve_f32", texture, uv, kernel, boundaries, iterations, profile
    )


def convolve_f64(texture, uv, kernel, boundaries, iterations=1, *, profile=None):
    """float64 line integral convolution (reference ``src/lib.rs:467-482``)."""
    return _convolve(
        np.dtype("float64"), "rlic_convolve_f64", texture, uv, kernel, boundaries, iterations, profile
    )


def _check_nbins(nbins: Any) -> int:
    try:
        n = int(nbins.__index__())
    except AttributeError:
        raise TypeError(
            f"argument 'nbins': {type(nbins).__name__!r} object cannot be interpreted as an "
            "integer"
        ) from None
    if n < 1:
        raise ValueError(f"nbins must be a strictly positive integer, got {n}")
    if n > 1 << 24:
        raise ValueError(f"nbins must be at most {1 << 24}, got {n}")
    return n


def _check_tile_shape(tile_shape: Any) -> tuple[int, int]:
    try:
        th, tw = tile_shape
        th, tw = int(th.__index__()), int(tw.__index__())
    except (TypeError, ValueError, AttributeError):
        raise TypeError(
            f"argument 'tile_shape': expected a pair of integers, got {tile_shape!r}"
        ) from None
    if th < 1 or tw < 1:
        raise ValueError(f"tile_shape must be strictly positive, got {(th, tw)}")
    return th, tw


def _equalize(dtype: np.dtype, entry: str, image: Any, nbins: Any, tile_shape: Any = None) -> Any:
    nb = _check_nbins(nbins)
    tile = () if tile_shape is None else _check_tile_shape(tile_shape)
    fn = getattr(_native.load(), entry)
    if _is_torch(image):
        import torch

        tdtype = {np.dtype("float32"): torch.float32, np.dtype("float64"): torch.float64}[dtype]
        if image.dtype != tdtype:
            raise TypeError(f"argument 'image': type mismatch: expected {tdtype}, got {image.dtype}")
        if image.dim() != 2:
            raise TypeError(f"argument 'image': dimensionality mismatch: expected 2, got {image.dim()}")
        if image.device.type == "cpu":
            return torch.from_numpy(_equalize(dtype, entry, image.numpy(), nb, tile_shape))
        ny, nx = image.shape
        with torch.cuda.device(image.device):
            out = torch.empty((ny, nx), dtype=tdtype, device=image.device)
            stream = torch.cuda.current_stream(image.device).cuda_stream
            status = fn(image.data_ptr(), image.stride(0), image.stride(1), ny, nx, nb, *tile,
                        out.data_ptr(), _native.RLIC_DEVICE, ctypes.c_void_p(stream))
        _native.raise_for_status(status)
        return out
    img, st = _numpy_view(image, "image", dtype, 2)
    ny, nx = img.shape
    out = np.empty((ny, nx), dtype=dtype)
    status = fn(img.ctypes.data, st[0], st[1], ny, nx, nb, *tile, out.ctypes.data,
                _native.RLIC_HOST, None)
    _native.raise_for_status(status)
    return out


def equalize_histogram_f32(image, nbins):
    """float32 global histogram equalization; the reference declares this
    function (``src/rlic/_core.pyi:30-33``) but implements it nowhere."""
    return _equalize(np.dtype("float32"), "rlic_equalize_histogram_f32", image, nbins)


def equalize_histogram_f64(image, nbins):
    """float64 global histogram equalization (declared, unimplemented, in
    the reference: ``src/rlic/_core.pyi:34-37``)."""
    return _equalize(np.dtype("float64"), "rlic_equalize_histogram_f64", image, nbins)


def equalize_histogram_sliding_tile_f32(image, nbins, tile_shape):
    """float32 sliding-window adaptive equalization (declared, unimplemented,
    in the reference: ``src/rlic/_core.pyi:38-42``)."""
    return _equalize(np.dtype("float32"), "rlic_equalize_histogram_sliding_tile_f32", image,
                     nbins, tile_shape)


def equalize_histogram_sliding_tile_f64(image, nbins, tile_shape):
    """float64 sliding-window adaptive equalization (``src/rlic/_core.pyi:43-47``)."""
    return _equalize(np.dtype("float64"), "rlic_equalize_histogram_sliding_tile_f64", image,
                     nbins, tile_shape)


def equalize_histogram_tile_interpolation_f32(image, nbins, tile_shape):
    """float32 tile-interpolated adaptive equalization (declared,
    unimplemented, in the reference: ``src/rlic/_core.pyi:48-52``)."""
    return _equalize(np.dtype("float32"), "rlic_equalize_histogram_tile_interpolation_f32", image,
                     nbins, tile_shape)


def equalize_histogram_tile_interpolation_f64(image, nbins, tile_shape):
    """float64 tile-interpolated adaptive equalization (``src/rlic/_core.pyi:53-57``)."""
    return _equalize(np.dtype("float64"), "rlic_equalize_histogram_tile_interpolation_f64", image,
                     nbins, tile_shape)
