"""Oracle results for GPU parity tests: the CPU restatement of the
reference (``oracle/``) on NumPy copies of NumPy or torch inputs."""

from __future__ import annotations

import numpy as np

from oracle import cpu


def to_numpy(a) -> np.ndarray:
    if hasattr(a, "detach"):  # a torch tensor, on any device
        return np.ascontiguousarray(a.detach().cpu().numpy())
    return np.ascontiguousarray(a)


def threads() -> int:
    return max(1, min(16, cpu.max_threads()))


def oracle_convolve(texture, u, v, *, kernel, uv_mode="velocity", boundaries="closed",
                    iterations=1, rows=None, profile="fma+branchless") -> np.ndarray:
    """The oracle's output of one convolve call (``iterations == 0``: a copy
    of the texture, src/rlic/_lib.py:208-209).  With ``rows=(r0, r1)`` only
    those rows of the last pass are computed (the others are zeros)."""
    tex = to_numpy(texture)
    if iterations == 0:
        return tex.copy()
    return cpu.convolve(tex, to_numpy(u), to_numpy(v), kernel=to_numpy(kernel), uv_mode=uv_mode,
                        boundaries=boundaries, iterations=int(iterations), profile=profile,
                        rows=rows, threads=threads())


def assert_matches_oracle(got, texture, u, v, *, kernel, uv_mode, boundaries, iterations,
                          max_whole_pixels=256 * 256, band_rows=2, what=""):
    """Compares `got` with the oracle, bit for bit (NaN where NaN).  Fields of
    at most `max_whole_pixels` pixels are compared whole; larger fields with
    one pass on a band of `band_rows` rows in the middle (the oracle computes
    only those).  Returns "whole", "band" or None (too large to check: more
    than one pass)."""
    ny, nx = int(texture.shape[0]), int(texture.shape[1])
    g = to_numpy(got)
    kw = dict(kernel=kernel, uv_mode=uv_mode, boundaries=boundaries, iterations=iterations)
    if ny * nx <= max_whole_pixels or iterations == 0:
        exp = oracle_convolve(texture, u, v, **kw)
        assert np.array_equal(g, exp, equal_nan=True), f"{what}: differs from the oracle"
        return "whole"
    if iterations == 1:
        r0 = ny // 2
        r1 = min(ny, r0 + band_rows)
        exp = oracle_convolve(texture, u, v, rows=(r0, r1), **kw)
        assert np.array_equal(g[r0:r1], exp[r0:r1], equal_nan=True), \
            f"{what}: rows [{r0}, {r1}) differ from the oracle"
        return "band"
    return None


def same_bits(a, b) -> bool:
    """Bit-for-bit equality of two arrays or tensors of one dtype and shape,
    any NaN matching any NaN (float32 fields iterated many times overflow
    to inf, and a zero end weight then makes 0 * inf = NaN)."""
    import torch

    if hasattr(a, "detach") and hasattr(b, "detach") and a.device == b.device:
        if a.shape != b.shape or a.dtype != b.dtype:
            return False
        ib = torch.int32 if a.dtype == torch.float32 else torch.int64
        same = (a.view(ib) == b.view(ib)) | (torch.isnan(a) & torch.isnan(b))
        return bool(same.all())
    x, y = to_numpy(a), to_numpy(b)
    if x.shape != y.shape or x.dtype != y.dtype:
        return False
    ib = np.int32 if x.dtype == np.float32 else np.int64
    return bool(((x.view(ib) == y.view(ib)) | (np.isnan(x) & np.isnan(y))).all())
