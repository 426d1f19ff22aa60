"""Histogram equalization of LIC images, on the GPU.

``equalize_histogram(image, nbins)`` maps every finite pixel to the
empirical cumulative distribution of the image at its bin,
``out = cdf[bin] / N``, with ``nbins`` equal-width bins over the finite
range (``librlic_hip.so``: ``rlic_equalize_histogram_f32/_f64``).  NaN and
infinite pixels are copied unchanged.  A ROCm tensor, e.g. the output of
``rlic_amd.convolve`` on device tensors, is equalized on its device and
stream without leaving HBM; NumPy arrays are uploaded and the result
downloaded.

``equalize_histogram_sliding_tile(image, nbins, tile_shape)`` equalizes
each pixel against its own window of ``tile_shape`` pixels (centred, clipped
to the image), and ``equalize_histogram_tile_interpolation(image, nbins,
tile_shape)`` against the mappings of the tiles of ``tile_shape`` around
it, interpolated bilinearly between tile centres; both bin over the whole
image's finite range.

The reference declares ``equalize_histogram_f32/_f64(image, nbins)`` and
the ``_sliding_tile`` / ``_tile_interpolation`` variants with a
``tile_shape`` in its native-module stub (``src/rlic/_core.pyi:30-57``) but
has no implementation (post-processing is out of its scope,
``README.md:19-23``); these are the same names and arguments.
"""

from __future__ import annotations

__all__ = ["equalize_histogram", "equalize_histogram_sliding_tile",
           "equalize_histogram_tile_interpolation"]

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


def _dtype_or_raise(image: Any):
    dtype = _dtype_of(image)
    if dtype is None:
        found = getattr(image, "dtype", type(image).__name__)
        raise TypeError(f"image must be a float32 or float64 array or tensor, got {found}")
    return dtype


def equalize_histogram(image: Any, nbins: int = 256) -> Any:
    """Global histogram equalization of a 2-D float32 or float64 image."""
    dtype = _dtype_or_raise(image)
    fn = _core.equalize_histogram_f32 if dtype == np.float32 else _core.equalize_histogram_f64
    return fn(image, nbins)


def equalize_histogram_sliding_tile(image: Any, nbins: int, tile_shape) -> Any:
    """Adaptive equalization: each pixel against its (th, tw) window."""
    dtype = _dtype_or_raise(image)
    fn = (_core.equalize_histogram_sliding_tile_f32 if dtype == np.float32
          else _core.equalize_histogram_sliding_tile_f64)
    return fn(image, nbins, tile_shape)


def equalize_histogram_tile_interpolation(image: Any, nbins: int, tile_shape) -> Any:
    """Adaptive equalization: per-tile mappings interpolated between tile
    centres."""
    dtype = _dtype_or_raise(image)
    fn = (_core.equalize_histogram_tile_interpolation_f32 if dtype == np.float32
          else _core.equalize_histogram_tile_interpolation_f64)
    return fn(image, nbins, tile_shape)
