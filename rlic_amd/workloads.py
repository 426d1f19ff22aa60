"""Seeded synthetic inputs for the benchmark configurations.

* ``readme_example`` — 256x256 float64, the reference README's example
  (``README.md:50-56, 78-91``): periodic, velocity, L=65 triangular kernel.
* ``fourier_modes`` — a field that is a sum of random Fourier modes
  ``a * cos(2*pi*(kx*j/W + ky*i/H) + phi)``, wavenumbers 1..16,
  ``a ~ N(0, 1)``, ``phi ~ U[0, 2*pi)``.
* ``plain_at_scale`` — 4096x4096 float32, texture U[0,1), u and v of 8 modes
  each, Gaussian kernel of 129 taps (sigma 32), x periodic / y closed,
  5 iterations.
* ``polarization_field`` — 16384x16384 float32 (or 8192x8192 float64),
  U = -1 left half / +1 right half plus 0.3 x modes, V = modes, triangular
  kernel of 257 taps, closed, polarization, 1 iteration.

Fields are built in float64 with NumPy (separable: cos(A + B) expanded into
outer products, so a 16384^2 field costs a few seconds) and cast to the
requested dtype, so inputs are identical wherever they are generated.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np


@dataclass
class Workload:
    name: str
    texture: np.ndarray
    u: np.ndarray
    v: np.ndarray
    kernel: np.ndarray
    boundaries: Any
    uv_mode: str
    iterations: int
    meta: dict = field(default_factory=dict)

    @property
    def shape(self) -> tuple[int, int]:
        return tuple(self.texture.shape)

    @property
    def pixel_steps(self) -> int:
        ny, nx = self.shape
        return ny * nx * (self.kernel.size - 1) * self.iterations

    def algorithmic_bytes_per_pass(self) -> int:
        """Reference gathers per pass: u, v, texture per step + centre read + write."""
        ny, nx = self.shape
        return ny * nx * (3 * (self.kernel.size - 1) + 2) * self.texture.itemsize

    def kwargs(self) -> dict:
        return dict(kernel=self.kernel, uv_mode=self.uv_mode, boundaries=self.boundaries,
                    iterations=self.iterations)


def fourier_modes(rng: np.random.Generator, shape: tuple[int, int], n_modes: int = 8,
                  kmax: int = 16) -> np.ndarray:
    """Sum of ``n_modes`` random plane waves on an H x W grid (float64)."""
    ny, nx = shape
    jj = np.arange(nx) / nx
    ii = np.arange(ny) / ny
    out = np.zeros(shape)
    for _ in range(n_modes):
        kx, ky = rng.integers(1, kmax + 1, size=2)
        a = rng.normal()
        phi = rng.uniform(0, 2 * np.pi)
        ax = 2 * np.pi * kx * jj + phi
        ay = 2 * np.pi * ky * ii
        # cos(ax + ay) = cos(ax)cos(ay) - sin(ax)sin(ay)
        out += a * (np.outer(np.cos(ay), np.cos(ax)) - np.outer(np.sin(ay), np.sin(ax)))
    return out


def readme_example() -> Workload:
    shape = (256, 256)
    texture = np.random.default_rng(0).random(shape)
    x = np.linspace(0, np.pi, 256)
    u = np.ascontiguousarray(np.broadcast_to(np.cos(2 * x), shape))
    v = np.ascontiguousarray(np.broadcast_to(np.sin(x).T, shape))
    kernel = 1 - np.abs(np.linspace(-1, 1, 65))
    return Workload("readme_example_256x256_f64", texture, u, v, kernel, "periodic",
                    "velocity", 1)


def gaussian_kernel(n: int, sigma: float, dtype=np.float32) -> np.ndarray:
    c = (n - 1) / 2
    return np.exp(-0.5 * ((np.arange(n) - c) / sigma) ** 2).astype(dtype)


def plain_at_scale(n: int = 4096, dtype=np.float32, iterations: int = 5,
                   seed: int = 2) -> Workload:
    rng = np.random.default_rng(seed)
    texture = rng.random((n, n), dtype=np.float64).astype(dtype)
    u = fourier_modes(rng, (n, n)).astype(dtype)
    v = fourier_modes(rng, (n, n)).astype(dtype)
    kernel = gaussian_kernel(129, 32.0, dtype)
    return Workload(f"plain_{n}x{n}_{np.dtype(dtype).name}_L129_it{iterations}", texture, u, v,
                    kernel, {"x": "periodic", "y": "closed"}, "velocity", iterations)


def polarization_field(n: int = 16384, dtype=np.float32, seed: int = 3) -> Workload:
    rng = np.random.default_rng(seed)
    texture = rng.random((n, n), dtype=np.float64).astype(dtype)
    half = np.where(np.arange(n) < n / 2, -1.0, 1.0)
    u = (half[None, :] + 0.3 * fourier_modes(rng, (n, n))).astype(dtype)
    v = fourier_modes(rng, (n, n)).astype(dtype)
    kernel = (1 - np.abs(np.linspace(-1, 1, 257))).astype(dtype)
    return Workload(f"polarization_{n}x{n}_{np.dtype(dtype).name}_L257", texture, u, v, kernel,
                    "closed", "polarization", 1)


WORKLOADS = {
    "readme": readme_example,
    "plain4096": plain_at_scale,
    "pol16384": lambda: polarization_field(16384, np.float32),
    "pol8192f64": lambda: polarization_field(8192, np.float64),
}
