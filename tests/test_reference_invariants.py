"""The scenarios of the reference's tests/test_convolution.py (:42-207), run
through the reference's import name (``import rlic``, the shim over
rlic_amd) on the GPU path: the same input recipes (128^2 random fields and
linspace(0, 1, 11) kernels, half-plane fields, NaN fields, the boundary
field), the same invariants, and every computed output also equal, bit for
bit, to the oracle's."""

from __future__ import annotations

import itertools

import numpy as np
import pytest

import rlic
from oracle_check import oracle_convolve
from rlic._typing import FArray1D, FArray2D  # noqa: F401  (the reference's test imports)

pytestmark = pytest.mark.gpu

NX = 128


def _args(seed, nx=NX, klen=11, dtype=np.float64):
    rng = np.random.default_rng(seed)
    return (rng.random((nx, nx), dtype=dtype), rng.random((nx, nx), dtype=dtype),
            rng.random((nx, nx), dtype=dtype), np.linspace(0, 1, klen, dtype=dtype))


def _run(img, u, v, **kw):
    """rlic.convolve on the GPU, checked against the oracle."""
    out = rlic.convolve(img, u, v, **kw)
    exp = oracle_convolve(img, u, v, kernel=kw["kernel"], uv_mode=kw.get("uv_mode", "velocity"),
                          boundaries=kw.get("boundaries", "closed"),
                          iterations=kw.get("iterations", 1))
    assert np.array_equal(out, exp, equal_nan=True), "differs from the oracle"
    return out


def test_no_iterations(gpu_device):
    img, u, v, k = _args(0)
    out = _run(img, u, v, kernel=k, iterations=0)
    assert np.array_equal(out, img) and out is not img


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_single_iteration(gpu_device, dtype):
    img, u, v, k = _args(1, dtype=dtype)
    assert np.array_equal(_run(img, u, v, kernel=k), _run(img, u, v, kernel=k, iterations=1))


def test_multiple_iterations(gpu_device):
    img, u, v, k = _args(2)
    outs = [_run(img, u, v, kernel=k, iterations=n) for n in range(3)]
    for o1, o2 in itertools.combinations(outs, 2):
        assert np.all(o2 != o1)


def test_uv_symmetry(gpu_device):
    img, u, v, k = _args(3)
    out1 = _run(img, u, v, kernel=k)
    out2 = _run(img.T, v.T, u.T, kernel=k).T
    assert np.array_equal(out2, out1)


def test_uv_mode_default(gpu_device):
    img, u, v, k = _args(4)
    assert np.array_equal(_run(img, u, v, kernel=k), _run(img, u, v, kernel=k, uv_mode="velocity"))


def test_uv_modes_diff(gpu_device):
    img = _args(5)[0]
    kernel = np.ones(5)
    u0 = np.ones((NX, NX))
    ii = np.broadcast_to(np.arange(NX), (NX, NX))
    u1 = np.where(ii < NX / 2, u0, -u0)
    u2 = np.where(ii < NX / 2, -u0, u0)
    v = np.zeros((NX, NX))
    vel1 = _run(img, u1, v, kernel=kernel, uv_mode="velocity")
    vel2 = _run(img, u2, v, kernel=kernel, uv_mode="velocity")
    np.testing.assert_allclose(vel2, vel1, atol=1e-14)
    pol1 = _run(img, u1, v, kernel=kernel, uv_mode="polarization")
    pol2 = _run(img, u2, v, kernel=kernel, uv_mode="polarization")
    np.testing.assert_allclose(pol2, pol1, atol=1e-14)
    assert np.ptp(vel2 - pol2) > 1


@pytest.mark.parametrize("kernel_size", [3, 4])
def test_uv_modes_equiv(gpu_device, kernel_size):
    # the fields are non-negative, so polarization never flips a vector
    img, u, v, _ = _args(6)
    kernel = np.ones(kernel_size)
    assert np.array_equal(_run(img, u, v, kernel=kernel, uv_mode="polarization"),
                          _run(img, u, v, kernel=kernel, uv_mode="velocity"))


def test_uv_mode_polarization_sym(gpu_device):
    n = 5
    kernel = np.ones(5)
    img = np.eye(n)
    zero, one = np.zeros((n, n)), np.ones((n, n))
    for uu, vv in ((one, zero), (zero, one)):
        fwd = _run(img, uu, vv, kernel=kernel, uv_mode="polarization")
        bwd = _run(img, -uu, -vv, kernel=kernel, uv_mode="polarization")
        np.testing.assert_allclose(bwd, fwd)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("niterations", [0, 1, 5])
def test_nan_vectors(gpu_device, dtype, niterations):
    img, _, _, k = _args(7, dtype=dtype)
    nan = np.full_like(img, np.nan)
    out = _run(img, nan, nan, kernel=k, iterations=niterations)
    scaling = out / img
    assert np.ptp(scaling) == 0.0
    assert scaling[0, 0] == k[len(k) // 2] ** niterations


def test_boundaries(gpu_device):
    img, _, _, k = _args(8, nx=64, klen=128)
    nx, ny = img.shape
    ii = np.broadcast_to(np.arange(nx), img.shape)
    U = np.where(ii < nx / 2, -1.0, 1.0)
    V = np.broadcast_to(np.sin(np.linspace(0, np.pi, ny)).T, img.shape)
    closed = _run(img, U, V, kernel=k, boundaries="closed")
    period = _run(img, U, V, kernel=k, boundaries="periodic")
    assert np.all(np.abs(closed - period) > 0)
    c_p = _run(img, U, V, kernel=k, boundaries={"x": "closed", "y": "periodic"})
    p_c = _run(img, U, V, kernel=k, boundaries={"y": "closed", "x": "periodic"})
    assert np.all(np.abs(c_p - period) > 0)
    assert np.all(np.abs(p_c - closed) > 0)
    assert np.all(np.abs(c_p - p_c) > 0)
