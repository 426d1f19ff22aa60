"""The public API contract (CPU only: every check fails before native code).

Re-expresses the reference's ``tests/test_exceptions.py`` and
``tests/test_boundaries.py`` against ``rlic_amd``: identical inputs, error
types and message patterns (the reference's ``src/rlic/_lib.py:135-205`` and
``src/rlic/_boundaries.py:38-114``).
"""

from __future__ import annotations

from dataclasses import replace
from itertools import permutations, product

import numpy as np
import pytest
from pytest import RaisesExc, RaisesGroup

import rlic_amd as rlic
from rlic_amd._boundaries import (
    COMBO_ALLOWED_BOUNDS,
    COMBO_DISALLOWED_BOUNDS,
    SUPPORTED_BOUNDS,
    BoundarySet,
)

img = np.eye(64)
u = img.copy()
v = img.copy()
kernel = np.linspace(0, 1, 10, dtype="float64")


# ---------------------------------------------------------------------------
# tests/test_exceptions.py
# ---------------------------------------------------------------------------
def test_invalid_iterations():
    with pytest.raises(
        ValueError,
        match=r"^Invalid number of iterations: -1\nExpected a strictly positive integer\.$",
    ):
        rlic.convolve(img, u, v, kernel=kernel, iterations=-1)


def test_invalid_uv_mode():
    with pytest.raises(
        ValueError,
        match=r"^Invalid uv_mode 'astral'\. Expected one of \['velocity', 'polarization'\]$",
    ):
        rlic.convolve(img, u, v, kernel=kernel, uv_mode="astral")


@pytest.mark.parametrize("shape", [(16, 16, 16)])
def test_invalid_texture_ndim(shape):
    with RaisesGroup(
        RaisesExc(ValueError,
                  match=r"^Expected a texture with exactly two dimensions\. Got texture\.ndim=3$"),
        RaisesExc(ValueError,
                  match=r"^Shape mismatch: expected texture, u and v with identical shapes\."),
        match=r"^Invalid inputs were received\.",
    ):
        rlic.convolve(np.ones(shape), u, v, kernel=kernel)


def test_invalid_texture_values():
    with pytest.raises(
        ValueError,
        match=r"^Found invalid texture element\(s\)\. Expected only positive values\.$",
    ):
        rlic.convolve(-np.ones((64, 64)), v, v, kernel=kernel)


def test_negative_zero_texture_is_valid_input():
    # -0.0 < 0 is False: accepted (the check is `texture < 0`)
    tex = np.full((4, 4), -0.0)
    out = rlic.convolve(tex, np.ones((4, 4)), np.ones((4, 4)), kernel=np.ones(3), iterations=0)
    assert np.array_equal(out, tex)


@pytest.mark.parametrize(
    "texture_shape, u_shape, v_shape",
    [
        ((64, 64), (65, 64), (64, 64)),
        ((64, 64), (64, 64), (63, 64)),
        ((64, 66), (64, 64), (64, 64)),
    ],
)
def test_mismatched_shapes(texture_shape, u_shape, v_shape):
    prng = np.random.default_rng(0)
    texture = prng.random(texture_shape)
    uu = prng.random(u_shape)
    vv = prng.random(v_shape)
    with pytest.raises(
        ValueError,
        match=(
            r"^Shape mismatch: expected texture, u and v with identical shapes\. "
            rf"Got texture.shape=\({texture.shape[0]}, {texture.shape[1]}\), "
            rf"u.shape=\({uu.shape[0]}, {uu.shape[1]}\), "
            rf"v.shape=\({vv.shape[0]}, {vv.shape[1]}\)$"
        ),
    ):
        rlic.convolve(texture, uu, vv, kernel=kernel)


def test_invalid_kernel_ndim():
    with pytest.raises(
        ValueError, match=r"^Expected a kernel with exactly one dimension\. Got kernel\.ndim=2$"
    ):
        rlic.convolve(img, u, v, kernel=np.ones((5, 5)))


@pytest.mark.parametrize("polluting_value", [-np.inf, np.inf, np.nan])
def test_non_finite_kernel(polluting_value):
    k = np.ones(11)
    k[5] = polluting_value
    with pytest.raises(ValueError, match=r"^Found non-finite value\(s\) in kernel\.$"):
        rlic.convolve(img, u, v, kernel=k)


def test_invalid_texture_dtype():
    with RaisesGroup(
        RaisesExc(
            TypeError,
            match=(
                r"^Found unsupported data type\(s\): \[dtype\('complex128'\)\]\. "
                r"Expected texture, u, v and kernel with identical dtype, from "
                r"\[dtype\('float32'\), dtype\('float64'\)\]\. "
                r"Got texture\.dtype=dtype\('complex128'\), u\.dtype=dtype\('float64'\), "
                r"v\.dtype=dtype\('float64'\), kernel\.dtype=dtype\('float64'\)$"
            ),
        ),
        RaisesExc(TypeError, match=r"^Data types mismatch"),
        match=r"^Invalid inputs were received\.",
    ):
        rlic.convolve(np.ones((64, 64), dtype="complex128"), u, v, kernel=kernel)


def test_invalid_kernel_dtype():
    with RaisesGroup(
        RaisesExc(
            TypeError,
            match=(
                r"^Found unsupported data type\(s\): \[dtype\('complex128'\)\]\. "
                r"Expected texture, u, v and kernel with identical dtype, from "
                r"\[dtype\('float32'\), dtype\('float64'\)\]\. "
                r"Got texture\.dtype=dtype\('float64'\), u\.dtype=dtype\('float64'\), "
                r"v\.dtype=dtype\('float64'\), kernel\.dtype=dtype\('complex128'\)$"
            ),
        ),
        RaisesExc(TypeError, match=r"^Data types mismatch"),
        match=r"^Invalid inputs were received\.",
    ):
        rlic.convolve(img, u, v, kernel=-np.ones(5, dtype="complex128"))


def test_mismatched_dtypes():
    with pytest.raises(
        TypeError,
        match=(
            r"^Data types mismatch. "
            r"Expected texture, u, v and kernel with identical dtype, from "
            r"\[dtype\('float32'\), dtype\('float64'\)\]\. "
            r"Got texture\.dtype=dtype\('float32'\), u\.dtype=dtype\('float64'\), "
            r"v\.dtype=dtype\('float64'\), kernel\.dtype=dtype\('float64'\)$"
        ),
    ):
        rlic.convolve(np.ones((64, 64), dtype="float32"), u, v, kernel=kernel)


def test_all_validators_before_returns():
    with pytest.raises(ValueError, match=r"^Found non-finite value\(s\) in kernel\.$"):
        rlic.convolve(img, u, v, kernel=np.full(11, np.nan), iterations=0)


def test_invalid_boundary_type():
    with pytest.raises(TypeError, match=r"^Invalid boundary specification "):
        rlic.convolve(img, u, v, kernel=kernel, boundaries=None)


def test_many_problems_are_grouped():
    with RaisesGroup(
        RaisesExc(ValueError, match=r"^Invalid number of iterations: -2"),
        RaisesExc(ValueError, match=r"^Invalid uv_mode 'x'"),
        RaisesExc(ValueError, match=r"^Unknown left x boundary 'open'$"),
        match=r"^Invalid inputs were received\.",
    ):
        rlic.convolve(img, u, v, kernel=kernel, iterations=-2, uv_mode="x",
                      boundaries={"x": ("open", "closed"), "y": "closed"})


def test_iterations_zero_returns_a_copy():
    # tests/test_convolution.py:42-44; src/rlic/_lib.py:208-209
    out = rlic.convolve(img, u, v, kernel=kernel, iterations=0)
    assert np.array_equal(out, img)
    assert out is not img and not np.shares_memory(out, img)


def test_drop_in_module_surface():
    import rlic_amd
    from rlic_amd import _core, _typing

    assert rlic_amd.__all__ == ["convolve"]
    assert callable(_core.convolve_f32) and callable(_core.convolve_f64)
    for name in ("F", "FArray1D", "FArray2D", "Pair", "PairSpec", "UVMode", "D1", "D2"):
        assert hasattr(_typing, name)


# ---------------------------------------------------------------------------
# tests/test_boundaries.py
# ---------------------------------------------------------------------------
@pytest.mark.parametrize(
    "spec, expected",
    [
        pytest.param("a", BoundarySet(x=("a", "a"), y=("a", "a")), id="expand-all"),
        pytest.param({"x": "a", "y": "b"}, BoundarySet(x=("a", "a"), y=("b", "b")),
                     id="expand-keys"),
        pytest.param({"x": ("a", "b"), "y": ("c", "w")}, BoundarySet(x=("a", "b"), y=("c", "w")),
                     id="already-expanded"),
        pytest.param({"x": ["a", "b"], "y": ["c", "w"]}, BoundarySet(x=("a", "b"), y=("c", "w")),
                     id="lists-to-tuples"),
    ],
)
def test_from_spec(spec, expected):
    assert BoundarySet.from_spec(spec) == expected


@pytest.mark.parametrize(
    "spec",
    [123, None, ["a", "b"], {"x": "a"}, {"y": "b"}, {"x": "a", "y": "b", "z": "c"},
     {"x": ("a", "b", "c"), "y": "a"}, {"x": 1, "y": "a"}],
)
def test_from_spec_invalid_type(spec):
    assert BoundarySet.from_spec(spec) is None


@pytest.mark.parametrize(
    "bound_input",
    [BoundarySet(x=(b1, b1), y=(b2, b2)) for (b1, b2) in product(SUPPORTED_BOUNDS, SUPPORTED_BOUNDS)],
)
def test_validate_simple_bounds(bound_input):
    assert not bound_input.collect_exceptions()
    assert bound_input.validate() is None


@pytest.mark.parametrize(
    "bound_input",
    [BoundarySet(x=(b1, b2), y=(b1, b1))
     for (b1, b2) in product(COMBO_DISALLOWED_BOUNDS, SUPPORTED_BOUNDS) if b1 != b2],
)
def test_invalid_combos(bound_input):
    with pytest.raises(ValueError, match=r"^left x boundary '\w+' cannot be combined"):
        bound_input.validate()
    reversed_x = replace(bound_input, x=reversed(bound_input.x))
    with pytest.raises(ValueError, match=r"^right x boundary '\w+' cannot be combined"):
        reversed_x.validate()
    swapped = BoundarySet(x=bound_input.y, y=bound_input.x)
    with pytest.raises(ValueError, match=r"^left y boundary '\w+' cannot be combined"):
        swapped.validate()
    swapped_reversed_y = replace(swapped, y=reversed(swapped.y))
    with pytest.raises(ValueError, match=r"^right y boundary '\w+' cannot be combined"):
        swapped_reversed_y.validate()


def test_combo_message_names_the_other_side():
    with pytest.raises(
        ValueError,
        match=r"^left x boundary 'periodic' cannot be combined with a different boundary "
              r"\('closed'\)$",
    ):
        BoundarySet(x=("periodic", "closed"), y=("closed", "closed")).validate()


@pytest.mark.parametrize(
    "bound_input",
    [BoundarySet(x=(b1, b2), y=(b3, b4))
     for (b1, b2, b3, b4) in set(permutations(("periodic", "periodic", "periodic", "unknown")))],
)
def test_unknown_bounds(bound_input):
    with pytest.raises(ValueError, match=r"^Unknown (left|right) (x|y) boundary 'unknown'$"):
        bound_input.validate()


def test_multiple_errors():
    with RaisesGroup(
        RaisesExc(ValueError, match=r"^Unknown left x boundary 'unknown'$"),
        RaisesExc(ValueError, match=r"^left y boundary 'periodic' cannot be combined"),
        match="Found multiple issues with boundary specifications",
    ):
        BoundarySet(x=("unknown", "periodic"), y=("periodic", "closed")).validate()


def test_allowed_combo_closed_pairs():
    assert COMBO_ALLOWED_BOUNDS == frozenset({"closed"})
    assert not BoundarySet(x=("closed", "closed"), y=("closed", "closed")).collect_exceptions()


class _DLPackOnly:
    """An array of a foreign library: it exports DLPack and nothing else."""

    def __init__(self, arr):
        self._arr = arr

    def __dlpack__(self, *args, **kwargs):
        return self._arr.__dlpack__(*args, **kwargs)

    def __dlpack_device__(self):
        return self._arr.__dlpack_device__()


def test_dlpack_producers_are_viewed_in_place():
    from rlic_amd._lib import _from_dlpack

    a = np.arange(12.0).reshape(3, 4)[:, ::2]  # strided view
    t = _from_dlpack(_DLPackOnly(a))
    assert type(t).__module__.startswith("torch")
    assert t.data_ptr() == a.ctypes.data  # no copy
    assert tuple(t.stride()) == tuple(s // a.itemsize for s in a.strides)
    assert _from_dlpack(a) is a
    plain = [1.0, 2.0]
    assert _from_dlpack(plain) is plain


def test_dlpack_inputs_are_validated():
    img = _DLPackOnly(np.zeros((4, 4), dtype="int64"))
    k = _DLPackOnly(np.ones(3, dtype="int64"))
    with pytest.raises(TypeError, match=r"Found unsupported data type\(s\)"):
        rlic.convolve(img, img, img, kernel=k)


def test_reference_import_name():
    """``import rlic`` (the reference's name) gives this build's convolve and
    the submodules the reference's tests import."""
    import rlic
    import rlic_amd
    from rlic._boundaries import BoundarySet, SUPPORTED_BOUNDS
    from rlic._core import convolve_f32, convolve_f64
    from rlic._typing import D2, F, FArray1D, FArray2D, f64  # noqa: F401

    assert rlic.__all__ == ["convolve"] and rlic.convolve is rlic_amd.convolve
    assert BoundarySet is rlic_amd._boundaries.BoundarySet
    assert SUPPORTED_BOUNDS == {"closed", "periodic"}
    assert convolve_f32 is rlic_amd._core.convolve_f32 and callable(convolve_f64)
    with pytest.raises(ValueError, match="^Invalid number of iterations"):
        rlic.convolve(np.ones((4, 4)), np.ones((4, 4)), np.ones((4, 4)), kernel=np.ones(3),
                      iterations=-1)
