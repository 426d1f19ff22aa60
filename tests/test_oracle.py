"""Pins the CPU oracle to the reference's semantics (CPU only).

Checks, in order of strength:
1. the reference's own known-answer tests (``src/lib.rs:139-297``), for all
   four arithmetic profiles (CI runs them on the whole feature matrix,
   ``.github/workflows/ci.yml:103-123``);
2. hand-derived closed forms: step times at signed zeros / infinities,
   single steps at every edge, uniform-row / uniform-column / zero / NaN
   fields (sums of exact small integers, so order and fusion cannot matter);
3. an exact geometric walker for unit fields (u, v in {-1, 0, +1}) written
   from the geometry of the step rule with rational arithmetic — this pins
   the tie-break and the regression inputs of the reference's
   ``tests/test_regressions.py`` (whose ``vectorplot`` oracle is absent);
4. the independent exact-arithmetic restatement :mod:`oracle.pyref` on
   random inputs for every dtype x profile x mode x boundary combination;
5. invariants from the reference's ``tests/test_convolution.py``.

A plausible mistake in the oracle (a swapped axis, a dropped ``abs``, a
wrong tie-break, a fused product the reference does not fuse, a wrong
boundary side) fails at least one of these; ``test_planted_mistakes`` shows
that for several such mistakes planted into copies of the exact restatement.
"""

from __future__ import annotations

import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import cpu, pyref

ALL_PROFILES = [0, 1, 2, 3]
DTYPES = [np.float32, np.float64]


# --------------------------------------------------------------------------
# 1. the reference's known-answer tests (src/lib.rs:139-297)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("profile", ALL_PROFILES)
def test_time_to_next_pixel_known_answers(profile):
    # src/lib.rs:187-190, 192-195
    assert cpu.time_to_next_pixel(1.0, 0.0, profile) == 1.0
    assert cpu.time_to_next_pixel(-1.0, 1.0, profile) == 1.0
    # src/lib.rs:197-205 (f32 and f64)
    assert cpu.time_to_next_pixel(0.0, 0.5, profile, np.float32) == np.float32(np.inf)
    assert cpu.time_to_next_pixel(0.0, 0.5, profile, np.float64) == np.inf


@pytest.mark.parametrize("profile", ALL_PROFILES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_advance_zero_velocity(profile, dtype):
    # src/lib.rs:279-296
    i, j, fx, fy = cpu.advance(0.0, 0.0, 5, 5, 0.5, 0.5, shape=(10, 10), profile=profile,
                               dtype=dtype)
    assert (i, j, fx, fy) == (5, 5, 0.5, 0.5)


def test_select_pixel_row_major():
    # src/lib.rs:144-150: arr[[1, 1]] of [[1, 2], [3, 4]] is 4 -- (i, j) = (row, col).
    # A 2x2 image with a u-only field and kernel [0, 0, 1] (kmid = 1, only the
    # forward tap weighs): out[i, j] = X at the cell one step downstream,
    # X[i, j + 1] with the closed edge clamping j.
    tex = np.array([[1.0, 2.0], [3.0, 4.0]])
    out = cpu.convolve(tex, np.ones((2, 2)), np.zeros((2, 2)), kernel=np.array([0.0, 0.0, 1.0]))
    assert out.tolist() == [[2.0, 2.0], [4.0, 4.0]]


# --------------------------------------------------------------------------
# 2. closed forms
# --------------------------------------------------------------------------
def _ttnp_geometric(vel, f):
    """Mathematical time to the next edge (exact, for |vel| > 0, f in [0, 1])."""
    if vel > 0:
        return (1 - Fraction(f)) / Fraction(vel)
    return Fraction(f) / -Fraction(vel)


@pytest.mark.parametrize("profile", ALL_PROFILES)
@pytest.mark.parametrize("dtype", DTYPES)
def test_time_to_next_pixel_dyadic_exact(profile, dtype):
    # for dyadic velocity magnitudes and fractions every profile is exact
    for vel, f in itertools.product([0.25, 0.5, 1.0, 2.0, 4.0, -0.25, -1.0, -2.0],
                                    [0.0, 0.25, 0.5, 0.75, 1.0]):
        got = cpu.time_to_next_pixel(vel, f, profile, dtype)
        exact = _ttnp_geometric(vel, f)
        if profile & 2:  # branchless: absolute value
            exact = abs(exact)
        assert Fraction(float(got)) == exact, (vel, f)


def test_time_to_next_pixel_signed_zero_cases():
    inf = np.inf
    for dtype in DTYPES:
        for profile in (0, 1):  # branching profiles: +-0 velocity -> +inf always
            for vel in (0.0, -0.0):
                for f in (0.0, 0.5, 1.0):
                    assert cpu.time_to_next_pixel(vel, f, profile, dtype) == inf
        for profile in (2, 3):  # branchless: |r / 0|, NaN when r == 0
            # +0: factor 2, r = 2*(0.5 - f) + f = 1 - f
            assert cpu.time_to_next_pixel(0.0, 0.5, profile, dtype) == inf
            assert cpu.time_to_next_pixel(0.0, 0.0, profile, dtype) == inf
            assert np.isnan(cpu.time_to_next_pixel(0.0, 1.0, profile, dtype))
            # -0: factor 0, r = f
            assert cpu.time_to_next_pixel(-0.0, 0.5, profile, dtype) == inf
            assert cpu.time_to_next_pixel(-0.0, 1.0, profile, dtype) == inf
            assert np.isnan(cpu.time_to_next_pixel(-0.0, 0.0, profile, dtype))
        # infinite velocity: zero time, sign per profile
        for profile in ALL_PROFILES:
            assert cpu.time_to_next_pixel(np.inf, 0.5, profile, dtype) == 0.0
            assert cpu.time_to_next_pixel(-np.inf, 0.5, profile, dtype) == 0.0


def test_time_to_next_pixel_out_of_range_fraction():
    # a fraction that drifted above 1 gives a negative time in the branching
    # profiles and its absolute value in the branchless ones (src/lib.rs:161, 178)
    f = 1.0 + 2.0**-20
    for profile in (0, 1):
        assert cpu.time_to_next_pixel(1.0, f, profile) == -(2.0**-20)
    for profile in (2, 3):
        assert cpu.time_to_next_pixel(1.0, f, profile) == 2.0**-20


def test_profiles_differ_only_at_ulp_level():
    """Branching `(1 - f) / v` and branchless `fma(2, 0.5 - f, f) / v` differ by
    one ulp for some binary32 fractions below 0.25 (0.5 - f is inexact there)."""
    r32 = pyref.BINARY32.round
    witnesses = [np.float32(0.06839993), np.float32(0.06360278)]
    for f in witnesses:
        exact_f = Fraction(float(f))
        branching = r32(1 - exact_f)  # one rounding
        branchless = r32(2 * Fraction(r32(Fraction(1, 2) - exact_f)) + exact_f)
        assert branching != branchless
        assert cpu.time_to_next_pixel(1.0, f, 1, np.float32) == np.float32(branching)
        assert cpu.time_to_next_pixel(1.0, f, 3, np.float32) == np.float32(branchless)
        assert abs(branching - branchless) == np.spacing(np.float32(branching))
    # for fractions in [0.25, 1] 0.5 - f is exact and the profiles agree
    for f in np.linspace(0.25, 1.0, 257):
        assert cpu.time_to_next_pixel(0.7, f, 1) == cpu.time_to_next_pixel(0.7, f, 3)


@pytest.mark.parametrize("profile", ALL_PROFILES)
def test_advance_edges(profile):
    """Hand-derived single steps, including both boundary kinds on each side."""
    shape = (4, 5)
    # +x from the centre of a cell: tx = 0.5 < ty = inf -> j+1, fx=0, fy unchanged
    assert cpu.advance(1.0, 0.0, 1, 1, 0.5, 0.5, shape=shape, profile=profile) == (1, 2, 0.0, 0.5)
    # -y: i-1, fy=1
    assert cpu.advance(0.0, -1.0, 1, 1, 0.5, 0.5, shape=shape, profile=profile) == (0, 1, 0.5, 1.0)
    # tie (u = v = 1 from the centre) goes to y; fx advanced by t*u = 0.5
    assert cpu.advance(1.0, 1.0, 1, 1, 0.5, 0.5, shape=shape, profile=profile) == (2, 1, 1.0, 0.0)
    # x edges
    for bx, j_lo, j_hi in (("closed", 0, 4), ("periodic", 4, 0)):
        b = {"x": bx, "y": "closed"}
        assert cpu.advance(-1.0, 0.0, 2, 0, 0.5, 0.5, shape=shape, boundaries=b,
                           profile=profile)[:2] == (2, j_lo)
        assert cpu.advance(1.0, 0.0, 2, 4, 0.5, 0.5, shape=shape, boundaries=b,
                           profile=profile)[:2] == (2, j_hi)
    # y edges
    for by, i_lo, i_hi in (("closed", 0, 3), ("periodic", 3, 0)):
        b = {"x": "closed", "y": by}
        assert cpu.advance(0.0, -1.0, 0, 2, 0.5, 0.5, shape=shape, boundaries=b,
                           profile=profile)[:2] == (i_lo, 2)
        assert cpu.advance(0.0, 1.0, 3, 2, 0.5, 0.5, shape=shape, boundaries=b,
                           profile=profile)[:2] == (i_hi, 2)
    # sides of one axis are independent: left periodic, right closed
    b = {"x": ("periodic", "closed"), "y": "closed"}
    assert cpu.advance(-1.0, 0.0, 2, 0, 0.5, 0.5, shape=shape, boundaries=b,
                       profile=profile)[:2] == (2, 4)
    assert cpu.advance(1.0, 0.0, 2, 4, 0.5, 0.5, shape=shape, boundaries=b,
                       profile=profile)[:2] == (2, 4)


def _int_data(rng, shape, klen):
    tex = rng.integers(0, 1000, size=shape).astype(np.float64)
    k = rng.integers(1, 50, size=klen).astype(np.float64)
    return tex, k


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("profile", ALL_PROFILES)
@pytest.mark.parametrize("bound", ["closed", "periodic"])
def test_uniform_rows_closed_form(dtype, profile, bound):
    """u = 1, v = 0: forward walks +x one cell per tap, backward -x."""
    rng = np.random.default_rng(1)
    ny, nx, klen = 6, 9, 13
    tex, k = _int_data(rng, (ny, nx), klen)
    kmid = klen // 2
    cols = np.arange(nx)
    exp = np.zeros((ny, nx))
    for t in range(klen):
        jj = cols + (t - kmid)
        jj = np.mod(jj, nx) if bound == "periodic" else np.clip(jj, 0, nx - 1)
        exp += k[t] * tex[:, jj]
    got = cpu.convolve(tex.astype(dtype), np.ones((ny, nx), dtype), np.zeros((ny, nx), dtype),
                       kernel=k.astype(dtype), boundaries=bound, profile=profile)
    assert np.array_equal(got, exp.astype(dtype))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("profile", ALL_PROFILES)
@pytest.mark.parametrize("bound", ["closed", "periodic"])
def test_uniform_columns_closed_form(dtype, profile, bound):
    """u = 0, v = -2: forward walks -y (rows), backward +y."""
    rng = np.random.default_rng(2)
    ny, nx, klen = 8, 5, 11
    tex, k = _int_data(rng, (ny, nx), klen)
    kmid = klen // 2
    rows = np.arange(ny)
    exp = np.zeros((ny, nx))
    for t in range(klen):
        ii = rows - (t - kmid)
        ii = np.mod(ii, ny) if bound == "periodic" else np.clip(ii, 0, ny - 1)
        exp += k[t] * tex[ii, :]
    got = cpu.convolve(tex.astype(dtype), np.zeros((ny, nx), dtype), np.full((ny, nx), -2.0, dtype),
                       kernel=k.astype(dtype), boundaries=bound, profile=profile)
    assert np.array_equal(got, exp.astype(dtype))


@pytest.mark.parametrize("profile", ALL_PROFILES)
def test_zero_field_closed_form(profile):
    """u = v = 0 (either sign): every tap re-samples the start cell."""
    rng = np.random.default_rng(3)
    tex, k = _int_data(rng, (5, 7), 9)
    u = np.where(rng.random((5, 7)) < 0.5, 0.0, -0.0)
    v = np.where(rng.random((5, 7)) < 0.5, 0.0, -0.0)
    got = cpu.convolve(tex, u, v, kernel=k, profile=profile, iterations=2)
    assert np.array_equal(got, k.sum() ** 2 * tex)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [0, 1, 5])
def test_nan_field_closed_form(dtype, n):
    """All-NaN vectors stop every line at once: out = K[kmid]**n * X."""
    rng = np.random.default_rng(4)
    tex = rng.random((9, 9)).astype(dtype)
    nan = np.full_like(tex, np.nan)
    k = np.linspace(0, 1, 11, dtype=dtype)
    got = cpu.convolve(tex, nan, nan, kernel=k, iterations=n)
    scale = got / tex
    assert np.ptp(scale) == 0.0
    assert scale[0, 0] == k[5] ** n


def test_kernel_length_one_and_two():
    rng = np.random.default_rng(5)
    tex = rng.random((6, 6))
    u, v = rng.random((6, 6)) - 0.5, rng.random((6, 6)) - 0.5
    # L = 1: both directions empty
    assert np.array_equal(cpu.convolve(tex, u, v, kernel=np.array([3.0])), 3.0 * tex)
    # L = 2: kmid = 1, no forward tap, one backward tap (k = 0)
    k = np.array([2.0, 1.0])
    tex_i, _ = _int_data(rng, (6, 6), 2)
    got = cpu.convolve(tex_i, np.ones((6, 6)), np.zeros((6, 6)), kernel=k)
    exp = 1.0 * tex_i + 2.0 * tex_i[:, np.clip(np.arange(6) - 1, 0, 5)]
    assert np.array_equal(got, exp)


def test_empty_kernel_is_rejected():
    with pytest.raises(ValueError):
        cpu.convolve(np.ones((3, 3)), np.ones((3, 3)), np.ones((3, 3)), kernel=np.ones(0))
    # nothing to compute: the kernel is never read (src/lib.rs:373-374)
    e = np.ones((0, 3))
    assert cpu.convolve(e, e, e, kernel=np.ones(0)).shape == (0, 3)


# --------------------------------------------------------------------------
# 3. exact geometric walker for unit fields (tie-break, regression inputs)
# --------------------------------------------------------------------------
def _walk_unit_field(tex, u, v, k, uv_mode, bounds):
    """LIC for u, v in {-1, 0, 1} with exact rational fractions.

    Written from the geometry: a line starts at the cell centre, moves at
    constant speed, and enters the neighbour across whichever edge it
    reaches first; reaching both at once counts as crossing y.  In
    polarization mode a vector pointing against the previous one is
    flipped.  Values are summed exactly and rounded once.
    """
    ny, nx = tex.shape
    kmid = len(k) // 2
    (xl, xr), (yl, yr) = bounds
    out = np.zeros_like(tex)
    for i0, j0 in itertools.product(range(ny), range(nx)):
        total = Fraction(float(k[kmid])) * Fraction(float(tex[i0, j0]))
        for sense, taps in ((1, range(kmid + 1, len(k))), (-1, range(kmid - 1, -1, -1))):
            i, j = i0, j0
            fx = fy = Fraction(1, 2)
            prev = (0, 0)
            for t in taps:
                pu, pv = int(u[i, j]), int(v[i, j])
                if uv_mode == "polarization":
                    if pu * prev[0] + pv * prev[1] < 0:
                        pu, pv = -pu, -pv
                    prev = (pu, pv)
                mu, mv = sense * pu, sense * pv
                if mu or mv:
                    tx = ((1 - fx) / mu if mu > 0 else fx / -mu) if mu else None
                    ty = ((1 - fy) / mv if mv > 0 else fy / -mv) if mv else None
                    cross_x = ty is None or (tx is not None and tx < ty)
                    if cross_x:
                        j += 1 if mu > 0 else -1
                        fx = Fraction(0) if mu > 0 else Fraction(1)
                        fy += tx * mv
                    else:
                        i += 1 if mv >= 0 else -1
                        fy = Fraction(0) if mv > 0 else Fraction(1)
                        fx += ty * mu
                    if j < 0:
                        j = 0 if xl == "closed" else nx - 1
                    elif j >= nx:
                        j = nx - 1 if xr == "closed" else 0
                    if i < 0:
                        i = 0 if yl == "closed" else ny - 1
                    elif i >= ny:
                        i = ny - 1 if yr == "closed" else 0
                total += Fraction(float(k[t])) * Fraction(float(tex[i, j]))
        out[i0, j0] = float(total)
    return out


def _regression_inputs(dtype=np.float32):
    """The reference's tests/test_regressions.py:19-40 inputs, restated."""
    shape = (31, 33)
    ny_cols = shape[1]
    nx_rows = shape[0]
    ones = np.ones(shape, dtype)
    col = np.broadcast_to(np.arange(ny_cols), shape)
    u1 = np.where(col < ny_cols / 2, -ones, ones)
    row = np.broadcast_to(np.arange(nx_rows)[:, None], shape)
    v1 = np.where(row < nx_rows / 2, -ones, ones)
    zero = np.zeros(shape, dtype)
    return u1, v1, zero


@pytest.mark.parametrize("uv_mode", ["velocity", "polarization"])
@pytest.mark.parametrize("which", ["00", "u0", "0v", "uv"])
def test_unit_fields_against_geometric_walker(uv_mode, which):
    u1, v1, zero = _regression_inputs(np.float64)
    u = u1 if which[0] == "u" else zero
    v = v1 if which[1] == "v" else zero
    # integer texture and kernel keep every sum exact, so the comparison is
    # about the visited cells only (the tie-break, the flips, the edges)
    rng = np.random.default_rng(6)
    tex = rng.integers(0, 100, size=u.shape).astype(np.float64)
    k = rng.integers(1, 20, size=17).astype(np.float64)
    for bound in ("closed", "periodic", {"x": "periodic", "y": "closed"}):
        bounds = pyref._norm_bounds(bound)
        exp = _walk_unit_field(tex, u, v, k, uv_mode, bounds)
        for profile in ALL_PROFILES:
            got = cpu.convolve(tex, u, v, kernel=k, uv_mode=uv_mode, boundaries=bound,
                               profile=profile)
            assert np.array_equal(got, exp), (which, bound, profile)


# --------------------------------------------------------------------------
# 4. the exact-arithmetic restatement, every combination
# --------------------------------------------------------------------------
COMBOS = list(itertools.product(
    DTYPES, ALL_PROFILES, ["velocity", "polarization"],
    ["closed", "periodic", {"x": ("periodic", "periodic"), "y": ("closed", "closed")}],
))


def _random_case(rng, dtype, shape=(5, 6), klen=9):
    tex = rng.random(shape).astype(dtype)
    u = (rng.random(shape) - 0.5).astype(dtype)
    v = (rng.random(shape) - 0.5).astype(dtype)
    u[1, 2] = 0.0
    v[1, 2] = 0.0
    v[3, :] = 0.0
    u[0, 0] = -0.0
    v[2, 4] = np.nan
    k = (rng.random(klen) - 0.2).astype(dtype)
    return tex, u, v, k


@pytest.mark.parametrize("dtype,profile,uv_mode,bound", COMBOS)
def test_oracle_matches_exact_restatement(dtype, profile, uv_mode, bound):
    rng = np.random.default_rng(hash((np.dtype(dtype).name, profile, uv_mode, str(bound))) % 2**32)
    tex, u, v, k = _random_case(rng, dtype)
    a = cpu.convolve(tex, u, v, kernel=k, uv_mode=uv_mode, boundaries=bound, iterations=2,
                     profile=profile)
    r = pyref.convolve(tex, u, v, kernel=k, uv_mode=uv_mode, boundaries=bound, iterations=2,
                       profile=profile)
    assert np.array_equal(a.view(np.uint8), r.view(np.uint8))


def test_exact_arithmetic_primitives():
    """The emulator's own arithmetic against known IEEE results."""
    f64, f32 = pyref.BINARY64, pyref.BINARY32
    assert f64.add(0.1, 0.2) == 0.1 + 0.2
    assert f64.div(1.0, 3.0) == 1.0 / 3.0
    assert f32.div(1.0, 3.0) == float(np.float32(1.0) / np.float32(3.0))
    assert f32.mul(0.1, 3.0) == float(np.float32(0.1) * np.float32(3.0))
    # fma rounds once: (1 + 2^-30)(1 - 2^-30) - 1 = -2^-60 exactly
    a = 1.0 + 2.0**-30
    b = 1.0 - 2.0**-30
    assert f64.fma(a, b, -1.0) == -(2.0**-60)
    assert a * b - 1.0 == 0.0  # two roundings lose it
    assert math.copysign(1.0, f64.add(-0.0, -0.0)) == -1.0
    assert math.copysign(1.0, f64.add(-0.0, 0.0)) == 1.0
    assert math.isnan(f64.div(0.0, 0.0)) and f64.div(1.0, -0.0) == -math.inf
    # binary32 subnormal rounding
    tiny = float(np.float32(1e-45))
    assert f32.mul(tiny, 0.5) == 0.0 and f32.mul(tiny, 1.5) == float(np.float32(2 * tiny))


# --------------------------------------------------------------------------
# 5. invariants of the reference's tests/test_convolution.py
# --------------------------------------------------------------------------
def _args(rng, n=32, klen=11, dtype=np.float64):
    return (rng.random((n, n), dtype=dtype), rng.random((n, n), dtype=dtype),
            rng.random((n, n), dtype=dtype), np.linspace(0, 1, klen, dtype=dtype))


def test_uv_transpose_symmetry():
    # tests/test_convolution.py:69-82
    img, u, v, k = _args(np.random.default_rng(0))
    a = cpu.convolve(img, u, v, kernel=k)
    b = cpu.convolve(img.T, v.T, u.T, kernel=k).T
    assert np.array_equal(a, b)


def test_iterations_differ_and_chain():
    # tests/test_convolution.py:60-66 and chaining of passes (src/lib.rs:432-440)
    img, u, v, k = _args(np.random.default_rng(1))
    outs = [cpu.convolve(img, u, v, kernel=k, iterations=n) for n in range(4)]
    for o1, o2 in itertools.combinations(outs[:3], 2):
        assert np.all(o1 != o2)
    assert np.array_equal(cpu.convolve(outs[1], u, v, kernel=k, iterations=2), outs[3])


def test_polarization_sign_flip_symmetry():
    # tests/test_convolution.py:93-110 and :125-162
    n = 32
    img = np.random.default_rng(2).random((n, n))
    ii = np.broadcast_to(np.arange(n), (n, n))
    u1 = np.where(ii < n / 2, 1.0, -1.0)
    z = np.zeros((n, n))
    k = np.ones(5)
    vel1 = cpu.convolve(img, u1, z, kernel=k)
    vel2 = cpu.convolve(img, -u1, z, kernel=k)
    pol1 = cpu.convolve(img, u1, z, kernel=k, uv_mode="polarization")
    pol2 = cpu.convolve(img, -u1, z, kernel=k, uv_mode="polarization")
    np.testing.assert_allclose(vel2, vel1, atol=1e-14)
    np.testing.assert_allclose(pol2, pol1, atol=1e-14)
    assert np.ptp(vel2 - pol2) > 1
    eye = np.eye(5)
    one, zero = np.ones((5, 5)), np.zeros((5, 5))
    for uu, vv in ((one, zero), (zero, one)):
        a = cpu.convolve(eye, uu, vv, kernel=k, uv_mode="polarization")
        b = cpu.convolve(eye, -uu, -vv, kernel=k, uv_mode="polarization")
        np.testing.assert_allclose(a, b)


def test_polarization_equals_velocity_for_short_kernels_on_positive_fields():
    # tests/test_convolution.py:113-122 (holds for non-negative fields)
    img, u, v, _ = _args(np.random.default_rng(3))
    for klen in (3, 4):
        k = np.ones(klen)
        assert np.array_equal(cpu.convolve(img, u, v, kernel=k),
                              cpu.convolve(img, u, v, kernel=k, uv_mode="polarization"))


def test_boundaries_change_every_pixel():
    # tests/test_convolution.py:180-207
    rng = np.random.default_rng(4)
    n = 64
    img = rng.random((n, n))
    k = np.linspace(0, 1, 128)
    ii = np.broadcast_to(np.arange(n), (n, n))
    U = np.where(ii < n / 2, -1.0, 1.0)
    V = np.broadcast_to(np.sin(np.linspace(0, np.pi, n)), (n, n))
    closed = cpu.convolve(img, U, V, kernel=k, boundaries="closed")
    period = cpu.convolve(img, U, V, kernel=k, boundaries="periodic")
    c_p = cpu.convolve(img, U, V, kernel=k, boundaries={"x": "closed", "y": "periodic"})
    p_c = cpu.convolve(img, U, V, kernel=k, boundaries={"x": "periodic", "y": "closed"})
    assert np.all(np.abs(closed - period) > 0)
    assert np.all(np.abs(c_p - period) > 0)
    assert np.all(np.abs(p_c - closed) > 0)
    assert np.all(np.abs(c_p - p_c) > 0)


# --------------------------------------------------------------------------
# planted mistakes: each must be caught by one of the checks above
# --------------------------------------------------------------------------
def _tie_to_x_step(state, mu, mv, fma, branchless, fmt, shape, bounds):
    """pyref.step with one mistake: a tie crosses x instead of y."""
    if mu == 0 and mv == 0:
        return
    tx = pyref.time_to_next_pixel(mu, state["fx"], fma, branchless, fmt)
    ty = pyref.time_to_next_pixel(mv, state["fy"], fma, branchless, fmt)
    if tx <= ty:
        state["j"] += 1 if mu >= 0 else -1
        state["fx"] = 0.0 if mu >= 0 else 1.0
        state["fy"] = fmt.fma(tx, mv, state["fy"])
    else:
        state["i"] += 1 if mv >= 0 else -1
        state["fy"] = 0.0 if mv >= 0 else 1.0
        state["fx"] = fmt.fma(ty, mu, state["fx"])
    (xl, xr), (yl, yr) = bounds
    state["j"] = pyref._wrap(state["j"], shape[1], xl, xr)
    state["i"] = pyref._wrap(state["i"], shape[0], yl, yr)


def test_planted_tie_mistake_is_caught(monkeypatch):
    u1, v1, _ = _regression_inputs(np.float64)
    rng = np.random.default_rng(8)
    tex = rng.integers(0, 100, size=(7, 9)).astype(np.float64)
    u, v = u1[:7, :9].copy(), v1[:7, :9].copy()
    k = rng.integers(1, 20, size=7).astype(np.float64)
    bounds = pyref._norm_bounds("closed")
    walker = _walk_unit_field(tex, u, v, k, "velocity", bounds)
    assert np.array_equal(pyref.convolve(tex, u, v, kernel=k), walker)
    monkeypatch.setattr(pyref, "step", _tie_to_x_step)
    assert not np.array_equal(pyref.convolve(tex, u, v, kernel=k), walker)


def test_planted_boundary_side_mistake_is_caught(monkeypatch):
    orig_wrap = pyref._wrap
    # mistake: the low-side rule is applied at the high side and vice versa
    monkeypatch.setattr(pyref, "_wrap", lambda c, size, low, high: orig_wrap(c, size, high, low))
    state = {"i": 2, "j": 0, "fx": 0.5, "fy": 0.5}
    pyref.step(state, -1.0, 0.0, True, True, pyref.BINARY64, (4, 5),
               (("periodic", "closed"), ("closed", "closed")))
    # hand-derived (test_advance_edges): left periodic sends j = -1 to 4
    assert state["j"] != 4


def test_planted_missing_abs_is_caught(monkeypatch):
    orig = pyref.time_to_next_pixel

    def no_abs(velocity, frac, fma, branchless, fmt):
        if not branchless:
            return orig(velocity, frac, fma, branchless, fmt)
        factor = fmt.add(1.0, -1.0 if fmt.signbit(velocity) else 1.0)
        return fmt.div(fmt.fma(factor, fmt.sub(0.5, frac), frac), velocity)

    monkeypatch.setattr(pyref, "time_to_next_pixel", no_abs)
    # the reference's own known answer (src/lib.rs:192-195): T(-1, 1) == 1
    assert pyref.time_to_next_pixel(-1.0, 1.0, True, True, pyref.BINARY64) != 1.0
