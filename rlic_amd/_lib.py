"""Public API: ``convolve`` (drop-in for ``rlic.convolve``).

Signature, defaults, validation order, error types and messages follow the
reference's ``src/rlic/_lib.py:37-235`` (pinned by its
``tests/test_exceptions.py``):

* every validator runs before anything returns (``_lib.py:135-200``);
* one problem is raised bare, several as an ExceptionGroup
  "Invalid inputs were received." (``_lib.py:202-205``);
* ``iterations == 0`` returns a copy of ``texture`` (``_lib.py:208-209``);
* otherwise the call dispatches on dtype to the native
  ``convolve_f32`` / ``convolve_f64`` (``_lib.py:228-235``).

Beyond the reference, the arrays may also be PyTorch tensors on a ROCm
device: the same checks then run on the device and the result is a device
tensor.  Arrays of other libraries that export DLPack (``__dlpack__``) are
viewed in place as PyTorch tensors (no copy) and handled the same way.
"""

from __future__ import annotations

__all__ = ["convolve", "validate"]

import sys
from typing import TYPE_CHECKING, Any

import numpy as np

from rlic_amd._boundaries import BoundarySet
from rlic_amd._core import convolve_f32, convolve_f64

if sys.version_info >= (3, 11):
    _ExceptionGroup = ExceptionGroup  # noqa: F821
else:  # pragma: no cover - exercised on the 3.10 interpreter
    from exceptiongroup import ExceptionGroup as _ExceptionGroup

if TYPE_CHECKING:
    from rlic_amd._boundaries import BoundarySpec
    from rlic_amd._typing import UVMode

_KNOWN_UV_MODES = ["velocity", "polarization"]
_SUPPORTED_DTYPES: list[np.dtype] = [np.dtype("float32"), np.dtype("float64")]


def _is_torch(x: Any) -> bool:
    return type(x).__module__.split(".", 1)[0] == "torch"


def _from_dlpack(arr: Any) -> Any:
    """A zero-copy PyTorch view of a DLPack-exporting array of another
    library; NumPy arrays, tensors and everything else pass unchanged."""
    if isinstance(arr, np.ndarray) or _is_torch(arr) or not hasattr(arr, "__dlpack__"):
        return arr
    import torch

    return torch.from_dlpack(arr)


_TORCH_DTYPES: dict = {}


def _np_dtype(arr: Any) -> Any:
    """The array's dtype as a NumPy dtype (torch dtypes mapped when possible)."""
    if _is_torch(arr):
        dt = arr.dtype
        if dt not in _TORCH_DTYPES:
            import torch

            try:
                _TORCH_DTYPES[dt] = np.dtype(torch.empty((), dtype=dt).numpy().dtype)
            except TypeError:  # e.g. bfloat16: no NumPy equivalent
                _TORCH_DTYPES[dt] = dt
        return _TORCH_DTYPES[dt]
    return arr.dtype


def _shape(arr: Any) -> tuple[int, ...]:
    return tuple(int(s) for s in arr.shape)


def _any_negative(arr: Any) -> bool:
    if _is_torch(arr):
        if arr.is_complex():
            return False
        return bool((arr < 0).any())
    return bool(np.any(arr < 0))


def _any_non_finite(arr: Any) -> bool:
    if _is_torch(arr):
        import torch

        return bool((~torch.isfinite(arr)).any())
    return bool(np.any(~np.isfinite(arr)))


_CHECK_ENTRIES: dict = {}


def _device_value_flags(texture: Any, kernel: Any) -> tuple[bool, bool] | None:
    """The library's one-launch check (``rlic_check_values_*``) for a 2-D
    float32/float64 texture and a 1-D kernel of its dtype on one ROCm
    device, or None when the arrays are not of that form."""
    if not _CHECK_ENTRIES:
        import torch

        from rlic_amd import _native

        lib = _native.load()
        _CHECK_ENTRIES.update({torch.float32: lib.rlic_check_values_f32,
                               torch.float64: lib.rlic_check_values_f64})
    entry = _CHECK_ENTRIES.get(texture.dtype)
    if entry is None or kernel.dtype != texture.dtype or texture.dim() != 2 \
            or kernel.dim() != 1:
        return None
    import ctypes

    import torch

    from rlic_amd import _native

    kern = kernel.contiguous()
    what = ctypes.c_int(0)
    dev = texture.device

    def run():
        stream = torch.cuda.current_stream(dev).cuda_stream
        _native.raise_for_status(entry(texture.data_ptr(), texture.shape[0], texture.shape[1],
                                       texture.stride(0), texture.stride(1), kern.data_ptr(),
                                       kern.shape[0], ctypes.byref(what), stream))

    if dev.index is None or dev.index == torch.cuda.current_device():
        run()
    else:
        with torch.cuda.device(dev):
            run()
    return bool(what.value & 1), bool(what.value & 2)


def _value_flags(texture: Any, kernel: Any) -> tuple[bool, bool]:
    """(texture has a negative element, kernel has a non-finite weight);
    float32/float64 device tensors are checked by the library in one
    launch, other device tensors reduced by torch and read back together."""
    if _is_torch(texture) and _is_torch(kernel) and texture.device == kernel.device \
            and texture.device.type != "cpu" and not texture.is_complex() \
            and not kernel.is_complex():
        flags = _device_value_flags(texture, kernel)
        if flags is not None:
            return flags
        import torch

        flags = torch.stack([(texture < 0).any(), (~torch.isfinite(kernel)).any()])
        neg, nonfinite = flags.tolist()
        return bool(neg), bool(nonfinite)
    return _any_negative(texture), _any_non_finite(kernel)


def _copy(arr: Any) -> Any:
    return arr.clone() if _is_torch(arr) else arr.copy()


def convolve(
    texture: Any,
    /,
    u: Any,
    v: Any,
    *,
    kernel: Any,
    uv_mode: UVMode = "velocity",
    boundaries: BoundarySpec = "closed",
    iterations: int = 1,
) -> Any:
    """2-dimensional line integral convolution.

    Apply Line Integral Convolution to a texture array, against a 2D flow
    (u, v) and via a 1D kernel, on an AMD Instinct GPU.

    Arguments
    ---------
    texture: 2D array, positional-only
      The tracer image to smear along the flow (random noise in general).
      Elements must not be negative.

    u, v: 2D arrays
      Horizontal (along axis 1, x) and vertical (along axis 0, y) components
      of the vector field, same shape as ``texture``.

    kernel: 1D array, keyword-only
      Weights along a field line.  ``kernel[len(kernel) // 2]`` weighs the
      starting pixel, later elements the downstream ("future") part of the
      line and earlier elements the upstream ("past") part.  Values must be
      finite; negative values are allowed.

    uv_mode: 'velocity' (default) or 'polarization', keyword-only
      In polarization mode the sign of (u, v) is ignored: along a line each
      vector is flipped when it points against the previous one.

    boundaries: 'closed' (default), 'periodic', or a dict with keys 'x' and
      'y' whose values are one of these strings or (left, right) pairs of
      them.  A 'periodic' side must face a 'periodic' side.

    iterations: int >= 0 (default: 1), keyword-only
      Number of passes; each pass convolves the previous pass's output.

    Returns
    -------
    A new 2D array of the inputs' dtype (float32 or float64).  NumPy inputs
    give a NumPy array; PyTorch tensors on a ROCm device give a tensor on
    that device.  Other DLPack-exporting arrays are read in place and give a
    PyTorch tensor.  ``iterations=0`` returns a copy of ``texture``.

    Raises
    ------
    TypeError
      Unsupported or mismatched dtypes, or an invalid boundary structure.
    ValueError
      Invalid values (see the messages).
    ExceptionGroup
      When more than one problem is found; all are reported.
    """
    texture, u, v, kernel = (_from_dlpack(a) for a in (texture, u, v, kernel))
    bs, dtype = validate(
        texture, u, v, kernel=kernel, uv_mode=uv_mode, boundaries=boundaries, iterations=iterations
    )
    if iterations == 0:
        return _copy(texture)

    native = convolve_f32 if dtype == np.dtype("float32") else convolve_f64
    return native(texture, (u, v, uv_mode), kernel, (bs.x, bs.y), iterations)


def validate(texture, u, v, *, kernel, uv_mode, boundaries, iterations, check_values=True):
    """Run every input check of ``convolve`` (src/rlic/_lib.py:135-205).

    Raises the single problem found, or all of them as an ExceptionGroup;
    returns the normalised BoundarySet and the common dtype otherwise.
    ``check_values=False`` skips the two checks that read array elements
    (negative texture values, non-finite kernel weights), for callers that
    run those checks themselves.
    """
    problems: list[Exception] = []
    neg_texture = nonfinite_kernel = False
    if check_values:
        neg_texture, nonfinite_kernel = _value_flags(texture, kernel)
    if iterations < 0:
        problems.append(
            ValueError(
                f"Invalid number of iterations: {iterations}\n"
                "Expected a strictly positive integer."
            )
        )
    if uv_mode not in _KNOWN_UV_MODES:
        problems.append(
            ValueError(f"Invalid uv_mode {uv_mode!r}. Expected one of {_KNOWN_UV_MODES}")
        )

    dtypes = [_np_dtype(a) for a in (texture, u, v, kernel)]
    distinct = list(dict.fromkeys(dtypes))
    unsupported = [dt for dt in distinct if dt not in _SUPPORTED_DTYPES]
    if unsupported or len(distinct) != 1:
        expectation = (
            "Expected texture, u, v and kernel with identical dtype, from "
            f"{_SUPPORTED_DTYPES}. Got texture.dtype={dtypes[0]!r}, u.dtype={dtypes[1]!r}, "
            f"v.dtype={dtypes[2]!r}, kernel.dtype={dtypes[3]!r}"
        )
        if unsupported:
            problems.append(
                TypeError(f"Found unsupported data type(s): {unsupported}. {expectation}")
            )
        if len(distinct) != 1:
            problems.append(TypeError(f"Data types mismatch. {expectation}"))

    if texture.ndim != 2:
        problems.append(
            ValueError(f"Expected a texture with exactly two dimensions. Got texture.ndim={texture.ndim}")
        )
    if neg_texture:
        problems.append(
            ValueError("Found invalid texture element(s). Expected only positive values.")
        )
    t_shape, u_shape, v_shape = _shape(texture), _shape(u), _shape(v)
    if u_shape != t_shape or v_shape != t_shape:
        problems.append(
            ValueError(
                "Shape mismatch: expected texture, u and v with identical shapes. "
                f"Got texture.shape={t_shape}, u.shape={u_shape}, v.shape={v_shape}"
            )
        )
    if kernel.ndim != 1:
        problems.append(
            ValueError(f"Expected a kernel with exactly one dimension. Got kernel.ndim={kernel.ndim}")
        )
    if nonfinite_kernel:
        problems.append(ValueError("Found non-finite value(s) in kernel."))

    bs = BoundarySet.from_spec(boundaries)
    if bs is None:
        problems.append(TypeError(f"Invalid boundary specification {boundaries}"))
    else:
        problems.extend(bs.collect_exceptions())

    if len(problems) == 1:
        raise problems[0]
    if problems:
        raise _ExceptionGroup("Invalid inputs were received.", problems)
    assert bs is not None
    return bs, dtypes[0]
