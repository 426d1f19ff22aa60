"""``rlic._boundaries``: ``rlic_amd._boundaries`` (reference
``src/rlic/_boundaries.py``)."""

from rlic_amd._boundaries import *  # noqa: F401,F403
from rlic_amd._boundaries import (  # noqa: F401
    COMBO_ALLOWED_BOUNDS,
    COMBO_DISALLOWED_BOUNDS,
    SUPPORTED_BOUNDS,
    BoundarySet,
    BoundarySpec,
)
