"""Type aliases of the public API (same names as the reference's
``src/rlic/_typing.py:1-37``, which its tests import)."""

from __future__ import annotations

__all__ = [
    "D1",
    "D2",
    "F",
    "FArray1D",
    "FArray2D",
    "Pair",
    "PairSpec",
    "UNSET",
    "UVMode",
    "UnsetType",
    "f32",
    "f64",
]

import enum
from typing import Literal, Tuple, TypeVar, Union

import numpy as np
from numpy import float32 as f32
from numpy import float64 as f64

T = TypeVar("T")

#: a (left, right) or (x, y) couple of identical types
Pair = Tuple[T, T]
#: a bare value standing for both members of a pair, or the pair itself
PairSpec = Union[T, Tuple[T, T]]

UVMode = Literal["velocity", "polarization"]


class UnsetType(enum.Enum):
    UNSET = enum.auto()


UNSET = UnsetType.UNSET

F = TypeVar("F", f32, f64)
D1 = Tuple[int]
D2 = Tuple[int, int]
FArray1D = np.ndarray[D1, np.dtype[F]]
FArray2D = np.ndarray[D2, np.dtype[F]]
