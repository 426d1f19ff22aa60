"""Line Integral Convolution on AMD Instinct MI355X (gfx950).

Drop-in for ``rlic.convolve`` (reference ``src/rlic/__init__.py:1-7``):
``import rlic_amd as rlic`` and call ``rlic.convolve(...)`` unchanged.
"""

__all__ = [
    "convolve",
]

from rlic_amd._lib import convolve
