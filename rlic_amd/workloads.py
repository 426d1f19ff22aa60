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
        return tuple(int(n) for n in self.texture.shape)

    @property
    def kernel_len(self) -> int:
        return int(self.kernel.shape[0])

    @property
    def itemsize(self) -> int:
        t = self.texture
        return int(t.element_size()) if hasattr(t, "element_size") else int(t.itemsize)

    @property
    def pixel_steps(self) -> int:
        ny, nx = self.shape
        return ny * nx * (self.kernel_len - 1) * self.iterations

    def algorithmic_bytes_per_pass(self) -> int:
        """Reference gathers per pass: u, v, texture per step + centre read + write."""
        ny, nx = self.shape
        return ny * nx * (3 * (self.kernel_len - 1) + 2) * self.itemsize

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


def masked_at_scale(n: int = 4096, dtype=np.float32, iterations: int = 5, seed: int = 6,
                    n_holes: int = 24) -> Workload:
    """The plain 4096^2 workload with NaN holes: u and v are NaN inside
    `n_holes` random disks (radius 40..240 px, about 10% of the pixels), so
    streamlines end at the holes' edges (src/lib.rs:336-338) after uneven
    numbers of steps: the case that makes lanes of a wave finish unevenly."""
    wl = plain_at_scale(n, dtype, iterations)
    rng = np.random.default_rng(seed)
    ii = np.arange(n)[:, None]
    jj = np.arange(n)[None, :]
    mask = np.zeros((n, n), dtype=bool)
    for _ in range(n_holes):
        ci, cj = rng.integers(0, n, size=2)
        r = rng.uniform(40, 240)
        mask |= (ii - ci) ** 2 + (jj - cj) ** 2 < r * r
    u = wl.u.copy()
    v = wl.v.copy()
    u[mask] = np.nan
    v[mask] = np.nan
    return Workload(f"masked_{n}x{n}_{np.dtype(dtype).name}_L129_it{iterations}", wl.texture, u, v,
                    wl.kernel, wl.boundaries, wl.uv_mode, iterations,
                    meta={"nan_fraction": float(mask.mean())})


def polarization_field(n: int = 16384, dtype=np.float32, seed: int = 3) -> Workload:
    rng = np.random.default_rng(seed)
    texture = rng.random((n, n), dtype=np.float64).astype(dtype)
    half = np.where(np.arange(n) < n / 2, -1.0, 1.0)
    u = (half[None, :] + 0.3 * fourier_modes(rng, (n, n))).astype(dtype)
    v = fourier_modes(rng, (n, n)).astype(dtype)
    kernel = (1 - np.abs(np.linspace(-1, 1, 257))).astype(dtype)
    return Workload(f"polarization_{n}x{n}_{np.dtype(dtype).name}_L257", texture, u, v, kernel,
                    "closed", "polarization", 1)


def batch_uneven(n_fields: int = 512, seed: int = 4, device=None, dtype=np.float32,
                 max_side: int = 2048, min_side: int = 128) -> list[dict]:
    """The uneven batch: n_fields independent convolve calls.

    Sides log-uniform in [min_side, max_side] (redrawn until the aspect
    ratio is within 1:4..4:1), kernel lengths uniform over odd 17..257
    (triangular or boxcar), each axis periodic or closed, uv_mode velocity or
    polarization, iterations uniform in 1..10, u and v sums of 8 Fourier
    modes, texture U[0,1).  The parameters come from a NumPy generator seeded
    with ``seed``; with ``device`` the arrays are generated there with torch
    (same distributions, different samples), otherwise with NumPy.
    Returns a list of dicts with keys texture, u, v, kernel, uv_mode,
    boundaries, iterations.
    """
    rng = np.random.default_rng(seed)
    fields = []
    for k in range(n_fields):
        while True:
            h, w = np.exp(rng.uniform(np.log(min_side), np.log(max_side), size=2)).round()
            if 0.25 <= h / w <= 4.0:
                break
        h, w = int(h), int(w)
        klen = int(rng.choice(np.arange(17, 258, 2)))
        if rng.random() < 0.5:
            kernel = 1 - np.abs(np.linspace(-1, 1, klen))
        else:
            kernel = np.ones(klen)
        bx = "periodic" if rng.random() < 0.5 else "closed"
        by = "periodic" if rng.random() < 0.5 else "closed"
        mode = "polarization" if rng.random() < 0.5 else "velocity"
        iterations = int(rng.integers(1, 11))
        fseed = int(rng.integers(0, 2**31))
        if device is not None:
            import torch

            gen = torch.Generator(device=device).manual_seed(fseed)
            tdt = torch.float32 if np.dtype(dtype) == np.float32 else torch.float64
            tex = torch.rand((h, w), generator=gen, device=device, dtype=torch.float64).to(tdt)
            u = _torch_modes(gen, h, w, device).to(tdt)
            v = _torch_modes(gen, h, w, device).to(tdt)
            kern = torch.from_numpy(kernel).to(device=device, dtype=tdt)
        else:
            frng = np.random.default_rng(fseed)
            tex = frng.random((h, w)).astype(dtype)
            u = fourier_modes(frng, (h, w)).astype(dtype)
            v = fourier_modes(frng, (h, w)).astype(dtype)
            kern = kernel.astype(dtype)
        fields.append(dict(texture=tex, u=u, v=v, kernel=kern, uv_mode=mode,
                           boundaries={"x": bx, "y": by}, iterations=iterations))
    return fields


def _torch_modes(gen, h: int, w: int, device, n_modes: int = 8, kmax: int = 16):
    import torch

    jj = torch.arange(w, device=device, dtype=torch.float64) / w
    ii = torch.arange(h, device=device, dtype=torch.float64) / h
    out = torch.zeros((h, w), device=device, dtype=torch.float64)
    for _ in range(n_modes):
        kx, ky = torch.randint(1, kmax + 1, (2,), generator=gen, device=device).tolist()
        a = torch.randn((), generator=gen, device=device, dtype=torch.float64)
        phi = torch.rand((), generator=gen, device=device, dtype=torch.float64) * 2 * np.pi
        out += a * torch.cos(2 * np.pi * (kx * jj[None, :] + ky * ii[:, None]) + phi)
    return out


def pixel_steps(fields: list[dict]) -> int:
    total = 0
    for f in fields:
        h, w = f["texture"].shape
        total += int(h) * int(w) * (int(f["kernel"].shape[0]) - 1) * int(f["iterations"])
    return total


def serving_pool(n_fields: int = 96, seed: int = 5, device=None, max_side: int = 2048,
                 min_side: int = 128) -> list[dict]:
    """Input fields for the serving workload, generated on `device` with
    torch: sides log-uniform in [min_side, max_side] with aspect ratio
    within 1:4..4:1, texture U[0,1), u and v sums of 8 Fourier modes;
    alternately float32 and float64.  Returns dicts with texture, u, v."""
    import torch

    rng = np.random.default_rng(seed)
    pool = []
    for k in range(n_fields):
        while True:
            h, w = np.exp(rng.uniform(np.log(min_side), np.log(max_side), size=2)).round()
            if 0.25 <= h / w <= 4.0:
                break
        h, w = int(h), int(w)
        tdt = torch.float32 if k % 2 == 0 else torch.float64
        gen = torch.Generator(device=device).manual_seed(int(rng.integers(0, 2**31)))
        tex = torch.rand((h, w), generator=gen, device=device, dtype=torch.float64).to(tdt)
        u = _torch_modes(gen, h, w, device).to(tdt)
        v = _torch_modes(gen, h, w, device).to(tdt)
        pool.append(dict(texture=tex, u=u, v=v))
    return pool


def serving_request(rng: np.random.Generator, pool: list[dict], max_iterations: int = 20) -> dict:
    """One request of the serving workload: a pool field with a kernel of
    odd length uniform over 17..257 (triangular or boxcar), each axis
    periodic or closed, uv_mode velocity or polarization, iterations
    uniform in 1..max_iterations."""
    import torch

    f = pool[int(rng.integers(len(pool)))]
    klen = int(rng.choice(np.arange(17, 258, 2)))
    kernel = 1 - np.abs(np.linspace(-1, 1, klen)) if rng.random() < 0.5 else np.ones(klen)
    bx = "periodic" if rng.random() < 0.5 else "closed"
    by = "periodic" if rng.random() < 0.5 else "closed"
    mode = "polarization" if rng.random() < 0.5 else "velocity"
    iterations = int(rng.integers(1, max_iterations + 1))
    kern = torch.from_numpy(kernel).to(device=f["texture"].device, dtype=f["texture"].dtype)
    return dict(texture=f["texture"], u=f["u"], v=f["v"], kernel=kern, uv_mode=mode,
                boundaries={"x": bx, "y": by}, iterations=iterations)


WORKLOADS = {
    "readme": readme_example,
    "plain4096": plain_at_scale,
    "masked4096": masked_at_scale,
    "pol16384": lambda: polarization_field(16384, np.float32),
    "pol8192f64": lambda: polarization_field(8192, np.float64),
}


def polarization_field_device(n: int = 16384, dtype=np.float32, seed: int = 3,
                              device="cuda") -> Workload:
    """``polarization_field`` generated on a ROCm device (no host arrays):
    texture from the counter-based noise of ``rlic_amd.generate.noise``
    (seeded), U = -1 left half / +1 right half plus 0.3 x modes, V = modes,
    the modes drawn from ``default_rng(seed)`` as ``polarization_field``
    draws them; closed, polarization, triangular kernel of 257 taps, one
    pass.  The arrays are torch tensors on ``device``."""
    import torch

    from rlic_amd import generate

    rng = np.random.default_rng(seed)
    dev = torch.device(device)
    texture = generate.noise((n, n), seed=seed, dtype=dtype, device=dev)
    u = generate.random_fourier_modes(rng, (n, n), dtype=dtype, device=dev)
    v = generate.random_fourier_modes(rng, (n, n), dtype=dtype, device=dev)
    tdt = texture.dtype
    half = torch.where(torch.arange(n, device=dev) < n / 2, -1.0, 1.0).to(tdt)
    u.mul_(0.3).add_(half[None, :])
    kernel = torch.from_numpy((1 - np.abs(np.linspace(-1, 1, 257))).astype(dtype)).to(dev)
    return Workload(f"polarization_{n}x{n}_{np.dtype(dtype).name}_L257", texture, u, v, kernel,
                    "closed", "polarization", 1, meta={"generated_on": "device"})
