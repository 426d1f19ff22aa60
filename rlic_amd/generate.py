"""On-device synthetic inputs: noise textures and Fourier-mode vector fields.

HIP kernels write the arrays straight into device memory, so large
benchmark and serving inputs (a 16384² float32 field is 1 GiB per array)
need neither host generation nor a host-to-device copy.  This module has no
counterpart in the reference, whose inputs are NumPy arrays made by the
caller.

* ``noise(shape, seed=...)``: uniform values in [0, 1) from a counter-based
  hash (splitmix64) of the seed and the element index, so a seed gives the
  same array on any device and at any launch shape;
* ``fourier_modes(shape, amplitude=..., kx=..., ky=..., phase=...)``: the
  sum over modes of ``a * cos(2 pi (kx j / nx + ky i / ny) + phase)``,
  evaluated in double precision as ``cos(ay) cos(ax) - sin(ay) sin(ax)``
  with ``ax = 2 pi kx (j / nx) + phase`` and ``ay = 2 pi ky (i / ny)``;
* ``random_fourier_modes(rng, shape)``: draws the mode parameters from a
  NumPy generator in the order ``rlic_amd.workloads.fourier_modes`` does,
  so both give the same field up to the last-place rounding of the
  device's and the host's cos and sin.

Results are PyTorch tensors on the given ROCm device, written on its
current stream.
"""

from __future__ import annotations

__all__ = ["MAX_MODES", "fourier_modes", "noise", "random_fourier_modes"]

import ctypes
from typing import Any

import numpy as np

from rlic_amd import _native

MAX_MODES = 32


def _torch_dtype(dtype: Any):
    import torch

    dt = np.dtype(dtype)
    if dt == np.float32:
        return torch.float32
    if dt == np.float64:
        return torch.float64
    raise TypeError(f"dtype must be float32 or float64, got {dt}")


def _shape(shape: Any) -> tuple[int, int]:
    ny, nx = (int(n) for n in shape)
    if ny < 0 or nx < 0:
        raise ValueError(f"negative image size {shape}")
    return ny, nx


def _stream(device) -> ctypes.c_void_p:
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def noise(shape: Any, *, seed: int = 0, dtype: Any = np.float32, device: Any = "cuda"):
    """Uniform noise in [0, 1) of the given 2-D shape on a ROCm device."""
    import torch

    ny, nx = _shape(shape)
    tdtype = _torch_dtype(dtype)
    device = torch.device(device)
    lib = _native.load()
    fn = lib.rlic_fill_noise_f32 if tdtype == torch.float32 else lib.rlic_fill_noise_f64
    with torch.cuda.device(device):
        out = torch.empty((ny, nx), dtype=tdtype, device=device)
        status = fn(out.data_ptr(), ny, nx, int(seed) % 2**64, _stream(device))
    _native.raise_for_status(status)
    return out


def fourier_modes(shape: Any, *, amplitude: Any, kx: Any, ky: Any, phase: Any,
                  dtype: Any = np.float32, device: Any = "cuda"):
    """Sum of plane waves (integer wavenumbers kx, ky; at most 32 modes) on
    a ROCm device, evaluated in double precision and rounded to ``dtype``."""
    import torch

    ny, nx = _shape(shape)
    tdtype = _torch_dtype(dtype)
    amp = np.ascontiguousarray(amplitude, dtype=np.float64).ravel()
    ph = np.ascontiguousarray(phase, dtype=np.float64).ravel()
    kxs = np.ascontiguousarray(kx, dtype=np.int32).ravel()
    kys = np.ascontiguousarray(ky, dtype=np.int32).ravel()
    n = amp.size
    if not (ph.size == kxs.size == kys.size == n):
        raise ValueError("amplitude, kx, ky and phase must have one length")
    if n > MAX_MODES:
        raise ValueError(f"at most {MAX_MODES} modes, got {n}")
    device = torch.device(device)
    lib = _native.load()
    fn = lib.rlic_fill_modes_f32 if tdtype == torch.float32 else lib.rlic_fill_modes_f64
    with torch.cuda.device(device):
        out = torch.empty((ny, nx), dtype=tdtype, device=device)
        status = fn(out.data_ptr(), ny, nx, n, amp.ctypes.data, kxs.ctypes.data,
                    kys.ctypes.data, ph.ctypes.data, _stream(device))
    _native.raise_for_status(status)
    return out


def random_fourier_modes(rng: np.random.Generator, shape: Any, n_modes: int = 8,
                         kmax: int = 16, *, dtype: Any = np.float32, device: Any = "cuda"):
    """``rlic_amd.workloads.fourier_modes(rng, shape, n_modes, kmax)``
    computed on the device: the same draws from ``rng``, in the same order."""
    amp, kx, ky, ph = [], [], [], []
    for _ in range(n_modes):
        kx_m, ky_m = rng.integers(1, kmax + 1, size=2)
        kx.append(kx_m)
        ky.append(ky_m)
        amp.append(rng.normal())
        ph.append(rng.uniform(0, 2 * np.pi))
    return fourier_modes(shape, amplitude=amp, kx=kx, ky=ky, phase=ph, dtype=dtype,
                         device=device)
