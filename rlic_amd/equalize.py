"""Histogram equalization of LIC images, on the GPU.

``equalize_histogram(image, nbins)`` maps every finite pixel to the
empirical cumulative distribution of the image at its bin,
``out = cdf[bin] / N``, with ``nbins`` equal-width bins over the finite
range (``librlic_hip.so``: ``rlic_equalize_histogram_f32/_f64``).  NaN and
infinite pixels are copied unchanged.  A ROCm tensor, e.g. the output of
``rlic_amd.convolve`` on device tensors, is equalized on its device and
stream without leaving HBM; NumPy arrays are uploaded and the result
downloaded.

The reference declares ``equalize_histogram_f32/_f64(image, nbins)`` in its
native-module stub (``src/rlic/_core.pyi:30-37``) but has no implementation
(post-processing is out of its scope, ``README.md:19-23``); these are the
same names and arguments.
"""

from __future__ import annotations

__all__ = ["equalize_histogram"]

from typing import Any

import numpy as np

from rlic_amd import _core


def _dtype_of(image: Any):
    if _core._is_torch(image):
        import torch

        return {torch.float32: np.dtype("float32"),
                torch.float64: np.dtype("float64")}.get(image.dtype)
    if isinstance(image, np.ndarray):
        return image.dtype if image.dtype in (np.float32, np.float64) else None
    return None


def equalize_histogram(image: Any, nbins: int = 256) -> Any:
    """Global histogram equalization of a 2-D float32 or float64 image."""
    dtype = _dtype_of(image)
    if dtype is None:
        found = getattr(image, "dtype", type(image).__name__)
        raise TypeError(f"image must be a float32 or float64 array or tensor, got {found}")
    fn = _core.equalize_histogram_f32 if dtype == np.float32 else _core.equalize_histogram_f64
    return fn(image, nbins)
