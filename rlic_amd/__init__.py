"""Line Integral Convolution on AMD Instinct MI355X (gfx950).

Drop-in for ``rlic.convolve`` (reference ``src/rlic/__init__.py:1-7``):
``import rlic_amd as rlic`` and call ``rlic.convolve(...)`` unchanged.
The public surface is the reference's (``__all__``); beyond it,
``rlic_amd.equalize_histogram`` (and its sliding-tile and
tile-interpolation variants) post-processes a LIC image where it lies
(device tensors stay on the device).
"""

__all__ = [
    "convolve",
]

from rlic_amd._lib import convolve
from rlic_amd.equalize import (  # noqa: F401
    equalize_histogram,
    equalize_histogram_sliding_tile,
    equalize_histogram_tile_interpolation,
)
