"""``rlic._typing``: the type aliases of ``rlic_amd._typing`` (reference
``src/rlic/_typing.py``)."""

from rlic_amd._typing import *  # noqa: F401,F403
from rlic_amd._typing import __all__  # noqa: F401
