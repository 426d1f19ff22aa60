"""Boundary specification: normalisation and validation.

Behaviour (accepted spellings, error types and messages) follows the
reference's ``src/rlic/_boundaries.py:38-114`` and is pinned by its
``tests/test_boundaries.py``:

* ``"closed"`` / ``"periodic"`` (any str) applies to all four sides;
* a mapping with exactly the keys ``"x"`` and ``"y"``, each a str or a
  2-sequence (tuple or list) of str ``(left, right)``;
* anything else is not a specification (``from_spec`` returns ``None``);
* unknown names are reported per side, and a ``"periodic"`` side must be
  paired with ``"periodic"`` on the opposite side of the same axis.
"""

from __future__ import annotations

__all__ = [
    "COMBO_ALLOWED_BOUNDS",
    "COMBO_DISALLOWED_BOUNDS",
    "SUPPORTED_BOUNDS",
    "BoundarySet",
]

import sys
from collections.abc import Mapping, Sequence
from dataclasses import dataclass
from typing import TYPE_CHECKING, Literal, Union

if sys.version_info >= (3, 11):
    _ExceptionGroup = ExceptionGroup  # noqa: F821
else:  # pragma: no cover - exercised on the 3.10 interpreter
    from exceptiongroup import ExceptionGroup as _ExceptionGroup

if TYPE_CHECKING:
    from typing import TypedDict

    from rlic_amd._typing import Pair, PairSpec

    class BoundaryDictSpec(TypedDict):
        x: PairSpec[BoundaryStr]
        y: PairSpec[BoundaryStr]


BoundaryStr = Literal["closed", "periodic"]
BoundarySpec = Union[BoundaryStr, "BoundaryDictSpec"]

#: sides that may face a different boundary type across the axis
COMBO_ALLOWED_BOUNDS: frozenset[str] = frozenset({"closed"})
#: sides that must face the same boundary type across the axis
COMBO_DISALLOWED_BOUNDS: frozenset[str] = frozenset({"periodic"})
SUPPORTED_BOUNDS: frozenset[str] = COMBO_ALLOWED_BOUNDS | COMBO_DISALLOWED_BOUNDS

_UNKNOWN = "Unknown {side} {ax} boundary {name!r}"
_BAD_COMBO = (
    "{side} {ax} boundary {name!r} cannot be combined with a different boundary ({other!r})"
)


def _as_pair(spec: object) -> tuple[str, str] | None:
    """A str, or a 2-sequence of str, as a (left, right) tuple; else None."""
    if isinstance(spec, str):
        return (spec, spec)
    if (
        isinstance(spec, Sequence)
        and not isinstance(spec, (bytes, bytearray))
        and len(spec) == 2
        and isinstance(spec[0], str)
        and isinstance(spec[1], str)
    ):
        return (spec[0], spec[1])
    return None


@dataclass(frozen=True, slots=True, kw_only=True)
class BoundarySet:
    """Per-axis (left, right) boundary names; x pairs with ``u``, y with ``v``."""

    x: Pair[BoundaryStr]
    y: Pair[BoundaryStr]

    @staticmethod
    def from_spec(spec: BoundarySpec, /) -> BoundarySet | None:
        """Normalise a user specification; ``None`` flags an invalid structure.

        Names are not checked here (see :meth:`collect_exceptions`).
        """
        if isinstance(spec, str):
            pair = (spec, spec)
            return BoundarySet(x=pair, y=pair)
        if isinstance(spec, Mapping) and len(spec) == 2 and "x" in spec and "y" in spec:
            px = _as_pair(spec["x"])
            py = _as_pair(spec["y"])
            if px is not None and py is not None:
                return BoundarySet(x=px, y=py)
        return None

    def collect_exceptions(self) -> list[Exception]:
        """All problems with the boundary names, as a list of ValueError."""
        found: list[Exception] = []
        for ax, sides in (("x", self.x), ("y", self.y)):
            left, right = sides
            names = {left, right}
            if names <= COMBO_ALLOWED_BOUNDS or (names <= SUPPORTED_BOUNDS and left == right):
                continue
            if names <= SUPPORTED_BOUNDS:
                # both names are known: only the pairing can be wrong
                for side, name, other in (("left", left, right), ("right", right, left)):
                    if name in COMBO_DISALLOWED_BOUNDS:
                        found.append(
                            ValueError(_BAD_COMBO.format(side=side, ax=ax, name=name, other=other))
                        )
                continue
            # a typo: report unknown names only, not the pairing
            for side, name in (("left", left), ("right", right)):
                if name not in SUPPORTED_BOUNDS:
                    found.append(ValueError(_UNKNOWN.format(side=side, ax=ax, name=name)))
        return found

    def validate(self) -> None:
        """Raise the single problem found, or all of them as a group."""
        found = self.collect_exceptions()
        if len(found) == 1:
            raise found[0]
        if found:
            raise _ExceptionGroup("Found multiple issues with boundary specifications", found)
