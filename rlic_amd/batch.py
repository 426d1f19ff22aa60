"""Many convolutions in one persistent GPU launch.

``convolve_batch(fields)`` runs one ``convolve`` per field — each field with
its own size, kernel, boundaries, uv_mode and iteration count — in a single
persistent launch fed by a device-side work queue (``rlic_convolve_batch_*``
of ``librlic_hip.so``), instead of one launch sequence per field.  Each result
is bit-identical to the corresponding ``rlic_amd.convolve`` call.

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
from rlic_amd._lib import validate

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


def convolve_batch(fields: Sequence, *, profile: int | str | None = None, device=None) -> list:
    """Convolve every field; returns the outputs in order."""
    import torch

    items = [_as_field(f) for f in fields]
    if not items:
        return []
    prof = _core.DEFAULT_PROFILE if profile is None else (
        _core.ARITHMETIC_PROFILES[profile] if isinstance(profile, str) else int(profile))
    host = not _core._is_torch(items[0].texture)
    checked = []
    for f in items:  # every field is validated before anything touches the device
        arrays = (f.texture, f.u, f.v, f.kernel)
        if any(_core._is_torch(a) for a in arrays) == host:
            raise TypeError("a batch holds NumPy arrays or torch tensors, not both")
        checked.append(validate(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                                boundaries=f.boundaries, iterations=f.iterations))
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
        for dtype, struct, entry, tdtype in (
            (np.dtype("float32"), _native.FieldF32, lib.rlic_convolve_batch_f32, torch.float32),
            (np.dtype("float64"), _native.FieldF64, lib.rlic_convolve_batch_f64, torch.float64),
        ):
            group = [p for p in prepared if p[1] == dtype]
            if not group:
                continue
            arr = (struct * len(group))()
            keep = []
            ctype = ctypes.c_float if dtype == np.float32 else ctypes.c_double
            ptr = lambda t: ctypes.cast(ctypes.c_void_p(t.data_ptr()),  # noqa: E731
                                        ctypes.POINTER(ctype))
            for slot, (idx, _, (tex, u, v, k), bs, mode, iters) in zip(arr, group):
                ny, nx = tex.shape
                out = torch.empty((ny, nx), dtype=tdtype, device=dev)
                keep.append(out)
                outputs[idx] = out
                slot.texture, slot.tex_sy, slot.tex_sx = ptr(tex), tex.stride(0), tex.stride(1)
                slot.u, slot.u_sy, slot.u_sx = ptr(u), u.stride(0), u.stride(1)
                slot.v, slot.v_sy, slot.v_sx = ptr(v), v.stride(0), v.stride(1)
                slot.ny, slot.nx = ny, nx
                slot.kernel, slot.kernel_len = ptr(k), k.shape[0]
                slot.uv_mode = mode
                slot.x_left, slot.x_right = _BOUND[bs.x[0]], _BOUND[bs.x[1]]
                slot.y_left, slot.y_right = _BOUND[bs.y[0]], _BOUND[bs.y[1]]
                slot.iterations = iters
                slot.out = ptr(out)
            status = entry(arr, len(group), prof, ctypes.c_void_p(stream))
            _native.raise_for_status(status)
    if host:
        torch.cuda.synchronize(dev)
        return [o.cpu().numpy() for o in outputs]
    return outputs
