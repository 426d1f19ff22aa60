"""The reference's import name, ``rlic``, for the MI355X build.

``import rlic`` gives ``rlic.convolve`` from ``rlic_amd`` (same signature,
validation and results; reference ``src/rlic/__init__.py:1-7``), and the
submodules the reference's tests import (``rlic._typing``,
``rlic._boundaries``, ``rlic._core``) resolve to their ``rlic_amd``
counterparts, so code written against the reference runs unchanged.
"""

__all__ = [
    "convolve",
]

from rlic_amd import convolve
