"""Many convolutions in one persistent GPU launch.

``convolve_batch(fields)`` runs one ``convolve`` per field — each field with
its own size, kernel, boundaries, uv_mode and iteration count — through
``rlic_convolve_batch_*`` of ``librlic_hip.so``: fields of at least 2^18
pixels run as single-field launches, all others together in one persistent
launch fed by a device-side work queue, instead of one launch sequence per
field.  Each result is bit-identical to the corresponding
``rlic_amd.convolve`` call.

Fields may hold NumPy arrays (uploaded to ``device`` and returned as NumPy)
or ROCm PyTorch tensors (returned as tensors on their device).  Every field
goes through the same validation as ``convolve``; float32 and float64 fields
may be mixed (they run as two launches).
"""

from __future__ import annotations

__all__ = ["Field", "convolve_batch"]

import ctypes
from dataclasses import dataclass, field
from typing import Any, Sequence

import numpy as np

from rlic_amd import _core, _native
from rlic_amd._boundaries import BoundarySet
from rlic_amd._lib import _KNOWN_UV_MODES, _value_flags, validate

_MODE = {"velocity": 0, "polarization": 1}
_BOUND = {"closed": 0, "periodic": 1}


@dataclass
class Field:
    """Arguments of one ``convolve`` call."""

    texture: Any
    u: Any
    v: Any
    kernel: Any
    uv_mode: str = "velocity"
    boundaries: Any = "closed"
    iterations: int = 1
    extra: dict = field(default_factory=dict)


def _as_field(f) -> Field:
    if isinstance(f, Field):
        return f
    if isinstance(f, dict):
        return Field(**f)
    texture, u, v = f[:3]
    return Field(texture, u, v, **(f[3] if len(f) > 3 else {}))


_TORCH_FLOAT_DTYPES: dict = {}
_BOUNDARY_CACHE: dict = {}


def _spec_key(spec):
    """A hashable form of a boundary specification, or None."""
    if isinstance(spec, str):
        return spec
    if isinstance(spec, dict):
        try:
            return tuple(sorted((k, tuple(v) if isinstance(v, (list, tuple)) else v)
                                for k, v in spec.items()))
        except TypeError:
            return None
    return None


def _quick_structure_check(f: Field):
    """``validate(..., check_values=False)``'s result for a field of ROCm
    tensors that passes every structural check (float32/float64 throughout,
    2-D texture, u, v of its shape, 1-D kernel, iterations >= 0, a known
    uv_mode, a valid boundary specification), or None to run ``validate``."""
    if not _TORCH_FLOAT_DTYPES:
        import torch

        _TORCH_FLOAT_DTYPES.update({torch.float32: np.dtype("float32"),
                                    torch.float64: np.dtype("float64")})
    t, u, v, k = f.texture, f.u, f.v, f.kernel
    if not all(_core._is_torch(a) for a in (t, u, v, k)):
        return None
    dtype = _TORCH_FLOAT_DTYPES.get(t.dtype)
    if dtype is None or u.dtype != t.dtype or v.dtype != t.dtype or k.dtype != t.dtype:
        return None
    if t.dim() != 2 or k.dim() != 1 or u.shape != t.shape or v.shape != t.shape:
        return None
    if not isinstance(f.iterations, int) or f.iterations < 0 or f.uv_mode not in _KNOWN_UV_MODES:
        return None
    key = _spec_key(f.boundaries)
    if key is None:
        return None
    bs = _BOUNDARY_CACHE.get(key)
    if bs is None:
        bs = BoundarySet.from_spec(f.boundaries)
        if bs is None or bs.collect_exceptions():
            return None
        _BOUNDARY_CACHE[key] = bs
    return bs, dtype


def _first_value_problem(items: list[Field], host: bool) -> int | None:
    """Index of the first field with a negative texture element or a
    non-finite kernel weight, or None, checking one field at a time: the
    host batch's check, and a device batch's for the fields ahead of a
    structurally invalid one.  A valid device batch is checked by the
    library instead (``rlic_batch_check_values_*``: every field of a dtype
    in one launch, over the field table the batched call takes)."""
    for k, f in enumerate(items):
        if host:
            if np.any(f.texture < 0) or not np.all(np.isfinite(f.kernel)):
                return k
        elif any(_value_flags(f.texture, f.kernel)):
            return k
    return None


def _raise_for(f: Field):
    """Raise what ``convolve`` raises for the invalid field f."""
    validate(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
             boundaries=f.boundaries, iterations=f.iterations)
    raise AssertionError("unreachable: the field failed validation above")


def convolve_batch(fields: Sequence, *, profile: int | str | None = None, device=None) -> list:
    """Convolve every field; returns the outputs in order.

    Every field is validated as ``convolve`` validates its arguments before
    anything runs; the first invalid field raises what ``convolve`` would
    raise for it.
    """
    import torch

    items = [_as_field(f) for f in fields]
    if not items:
        return []
    prof = _core.DEFAULT_PROFILE if profile is None else (
        _core.ARITHMETIC_PROFILES[profile] if isinstance(profile, str) else int(profile))
    host = not _core._is_torch(items[0].texture)
    checked = []
    first_bad = None
    for idx, f in enumerate(items):  # structure first, element values below
        arrays = (f.texture, f.u, f.v, f.kernel)
        if any(_core._is_torch(a) for a in arrays) == host:
            raise TypeError("a batch holds NumPy arrays or torch tensors, not both")
        quick = None if host else _quick_structure_check(f)
        if quick is not None:
            checked.append(quick)
            continue
        try:
            checked.append(validate(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                                    boundaries=f.boundaries, iterations=f.iterations,
                                    check_values=False))
        except Exception:
            first_bad = idx
            break
    if first_bad is not None:  # an earlier field's element values take precedence
        value_bad = _first_value_problem(items[:first_bad], host)
        _raise_for(items[first_bad if value_bad is None else value_bad])
    if host:
        value_bad = _first_value_problem(items, host)
        if value_bad is not None:
            _raise_for(items[value_bad])
    dev = torch.device(device) if device is not None else (
        torch.device("cuda", torch.cuda.current_device()) if host else items[0].texture.device)

    prepared = []  # (index, dtype, tensors, bs, mode, iterations)
    for idx, (f, (bs, dtype)) in enumerate(zip(items, checked)):
        arrays = (f.texture, f.u, f.v, f.kernel)
        if host:
            tensors = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrays]
        else:
            tensors = list(arrays)
            if any(t.device != dev for t in tensors):
                raise ValueError("all tensors of a batch must be on one device")
        tensors[3] = tensors[3].contiguous()
        prepared.append((idx, np.dtype(dtype), tensors, bs, _MODE[f.uv_mode], int(f.iterations)))

    outputs: list = [None] * len(items)
    lib = _native.load()
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev).cuda_stream
        tables = []  # (field table, struct, batched entry, value check, batch indices)
        for dtype, struct, entry, check, tdtype in (
            (np.dtype("float32"), _native.FieldF32, lib.rlic_convolve_batch_f32,
             lib.rlic_batch_check_values_f32, torch.float32),
            (np.dtype("float64"), _native.FieldF64, lib.rlic_convolve_batch_f64,
             lib.rlic_batch_check_values_f64, torch.float64),
        ):
            group = [p for p in prepared if p[1] == dtype]
            if not group:
                continue
            rows = []
            for idx, _, (tex, u, v, k), bs, mode, iters in group:
                ny, nx = tex.shape
                out = torch.empty((ny, nx), dtype=tdtype, device=dev)
                outputs[idx] = out
                rows.append((tex.data_ptr(), tex.stride(0), tex.stride(1),
                             u.data_ptr(), u.stride(0), u.stride(1),
                             v.data_ptr(), v.stride(0), v.stride(1),
                             ny, nx, k.data_ptr(), k.shape[0], mode,
                             _BOUND[bs.x[0]], _BOUND[bs.x[1]], _BOUND[bs.y[0]], _BOUND[bs.y[1]],
                             iters, out.data_ptr()))
            tables.append((np.array(rows, dtype=_native.FIELD_DTYPE), struct, entry, check,
                           [p[0] for p in group]))
        if not host:  # element values of every field, before any compute
            value_bad = None
            for arr, struct, _, check, index in tables:
                bad = ctypes.c_int64(-1)
                _native.raise_for_status(check(arr.ctypes.data_as(ctypes.POINTER(struct)),
                                               len(index), ctypes.byref(bad), None,
                                               ctypes.c_void_p(stream)))
                if bad.value >= 0 and (value_bad is None or index[bad.value] < value_bad):
                    value_bad = index[bad.value]
            if value_bad is not None:
                _raise_for(items[value_bad])
        for arr, struct, entry, _, index in tables:
            status = entry(arr.ctypes.data_as(ctypes.POINTER(struct)), len(index), prof,
                           ctypes.c_void_p(stream))
            _native.raise_for_status(status)
    if host:
        torch.cuda.synchronize(dev)
        return [o.cpu().numpy() for o in outputs]
    return outputs
