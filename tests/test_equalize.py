"""Histogram equalization (the reference's declared, unimplemented
``equalize_histogram_f32/_f64(image, nbins)``, src/rlic/_core.pyi:30-37).

CPU: the oracle (oracle/equalize.py) pinned by what the mathematics fixes:
the closed form on ramps and constant images, two-level images,
monotonicity, permutation invariance, a brute-force count, non-finite
pass-through, the overflow path, and a planted mistake that the closed
form catches.  Parity against the reference is unpinned: it has no
implementation.  GPU: the HIP kernels against the oracle, bit for bit,
including the output of a device convolve call equalized on the device."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import rlic_amd
from oracle import equalize as eq
from rlic_amd import _native


@pytest.mark.parametrize("n", [2, 3, 16, 255, 1000])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ramp_closed_form(n, dtype):
    # n distinct values with nbins = n: every value is alone in its bin, so
    # out = (rank + 1) / n exactly
    rng = np.random.default_rng(n)
    ranks = rng.permutation(n)
    img = (ranks.astype(dtype) * dtype(0.37) - dtype(5)).reshape(1, n)
    out = eq.equalize_histogram(img, n)
    assert np.array_equal(out, ((ranks + 1) / n).astype(dtype).reshape(1, n))


def test_constant_and_two_level_images():
    assert np.array_equal(eq.equalize_histogram(np.full((3, 4), 7.0), 10), np.ones((3, 4)))
    img = np.array([[1.0, 1.0, 3.0], [3.0, 3.0, 1.0]], dtype=np.float32)
    for nbins in (1, 2, 5, 100):
        out = eq.equalize_histogram(img, nbins)
        lower = np.float32(0.5) if nbins > 1 else np.float32(1.0)
        assert np.array_equal(out, np.where(img == 1.0, lower, np.float32(1.0)))


@pytest.mark.parametrize("nbins", [1, 2, 7, 64, 1000])
def test_monotone_and_permutation_invariant(nbins):
    rng = np.random.default_rng(nbins)
    img = rng.normal(size=(37, 23)) ** 3
    out = eq.equalize_histogram(img, nbins)
    order = np.argsort(img, axis=None, kind="stable")
    assert np.all(np.diff(out.ravel()[order]) >= 0)
    assert out.max() == 1.0 and out.min() > 0
    perm = rng.permutation(img.size)
    shuffled = img.ravel()[perm].reshape(img.shape)
    assert np.array_equal(eq.equalize_histogram(shuffled, nbins).ravel(), out.ravel()[perm])


@pytest.mark.parametrize("nbins", [1, 3, 10, 50])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bruteforce_count(nbins, dtype):
    rng = np.random.default_rng(nbins + 7)
    img = rng.random((9, 11)).astype(dtype)
    img[2, 3] = img[5, 5]  # equal values share a bin
    assert np.array_equal(eq.equalize_histogram(img, nbins),
                          eq.equalize_histogram_bruteforce(img, nbins))


def test_non_finite_pixels_pass_through():
    img = np.array([[np.nan, 1.0, np.inf], [2.0, -np.inf, 3.0]])
    out = eq.equalize_histogram(img, 3)
    assert np.isnan(out[0, 0]) and out[0, 2] == np.inf and out[1, 1] == -np.inf
    assert np.array_equal(out[[0, 1, 1], [1, 0, 2]], [1 / 3, 2 / 3, 1.0])
    allnan = np.full((2, 2), np.nan)
    assert np.all(np.isnan(eq.equalize_histogram(allnan, 4)))


def test_range_overflow_uses_halves():
    # vmax - vmin overflows float64: the halved form bins the extremes apart
    img = np.array([[-1.5e308, 0.0, 1.5e308, 1.0e308]])
    out = eq.equalize_histogram(img, 4)
    # halves: t = (x/2 + 0.75e308) * (4 / 1.5e308): bins 0, 2, 3, 3
    assert np.array_equal(out, [[0.25, 0.5, 1.0, 1.0]])


def test_planted_mistake_is_caught(monkeypatch):
    # rounding to the nearest bin instead of flooring: the ramp closed form fails
    real_floor = np.floor
    monkeypatch.setattr(eq.np, "floor", np.round)
    try:
        img = np.arange(16, dtype=np.float64).reshape(4, 4)
        out = eq.equalize_histogram(img, 5)
        bins = [min(int(real_floor(k * 5 / 15)), 4) for k in range(16)]
        counts = np.bincount(bins, minlength=5).cumsum()
        expected = np.array([counts[b] / 16 for b in bins]).reshape(4, 4)
        assert not np.array_equal(out, expected)
    finally:
        monkeypatch.undo()
    assert np.array_equal(eq.equalize_histogram(img, 5), expected)


def test_argument_errors():
    with pytest.raises(ValueError, match="nbins"):
        rlic_amd.equalize_histogram(np.ones((3, 3)), 0)
    with pytest.raises(TypeError):
        rlic_amd.equalize_histogram(np.ones((3, 3), dtype=np.int32), 4)
    with pytest.raises(TypeError):
        rlic_amd.equalize_histogram(np.ones(3), 4)
    with pytest.raises(TypeError):
        rlic_amd.equalize_histogram(np.ones((3, 3)), 2.5)
    lib = _native.load()
    a = np.ones((2, 2))
    assert lib.rlic_equalize_histogram_f64(a.ctypes.data, 2, 1, 2, 2, 1 << 25, a.ctypes.data,
                                           0, None) == _native.RLIC_ERR_INVALID_ARGUMENT
    assert "nbins must be in" in _native.last_error()
    # an empty image needs no device
    assert lib.rlic_equalize_histogram_f64(a.ctypes.data, 2, 1, 0, 2, 4, a.ctypes.data,
                                           0, None) == 0


def _images(rng, dtype, device):
    """(name, host image, the same image as a device tensor or view)."""
    import torch

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(device)

    img = rng.random((97, 131)).astype(dtype)
    yield "uniform", img, dev(img)
    img = (rng.normal(size=(64, 200)) ** 3).astype(dtype)
    yield "normal cubed", img, dev(img)
    img = rng.random((50, 70)).astype(dtype)
    img[rng.random(img.shape) < 0.05] = np.nan
    img[3, 4], img[7, 8] = np.inf, -np.inf
    yield "non-finite pixels", img, dev(img)
    img = np.full((33, 17), 2.5, dtype=dtype)
    yield "constant", img, dev(img)
    base = rng.random((120, 45)).astype(dtype)
    yield "transposed view", base.T, dev(base).T
    base = rng.random((90, 180)).astype(dtype)
    yield "strided view", base[::3, ::2], dev(base)[::3, ::2]
    img = np.round(rng.random((128, 128)) * 6).astype(dtype)
    yield "few levels", img, dev(img)
    img = np.array([[4.0]], dtype=dtype)
    yield "one pixel", img, dev(img)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nbins", [1, 2, 7, 256, 4099, 16384, 20000])
def test_gpu_matches_oracle(gpu_device, dtype, nbins):
    """Host arrays and device tensors (strided views read in place); nbins
    up to and past the LDS histogram's 16384 bins (global atomics)."""
    rng = np.random.default_rng(nbins)
    for name, img, t in _images(rng, dtype, gpu_device):
        exp = eq.equalize_histogram(img, nbins)
        got = rlic_amd.equalize_histogram(img, nbins)
        assert np.array_equal(got, exp, equal_nan=True), (name, nbins)
        assert tuple(t.shape) == img.shape
        dev = rlic_amd.equalize_histogram(t, nbins)
        assert dev.device == t.device
        assert np.array_equal(dev.cpu().numpy(), exp, equal_nan=True), (name, nbins, "device")


@pytest.mark.gpu
def test_gpu_equalizes_convolve_output_on_the_device(gpu_device):
    """The LIC output of a device convolve call, equalized on the device (it
    never goes to the host in between), equals the oracle's equalization of
    the oracle's convolution."""
    import torch

    from oracle import cpu

    rng = np.random.default_rng(3)
    shape = (300, 260)
    tex = rng.random(shape).astype(np.float32)
    u = (rng.random(shape) - 0.5).astype(np.float32)
    v = (rng.random(shape) - 0.5).astype(np.float32)
    k = (1 - np.abs(np.linspace(-1, 1, 41))).astype(np.float32)
    t = [torch.from_numpy(a).to(gpu_device) for a in (tex, u, v, k)]
    lic = rlic_amd.convolve(t[0], t[1], t[2], kernel=t[3], boundaries="periodic", iterations=3)
    out = rlic_amd.equalize_histogram(lic, 256)
    exp_lic = cpu.convolve(tex, u, v, kernel=k, boundaries="periodic", iterations=3)
    exp = eq.equalize_histogram(exp_lic, 256)
    assert np.array_equal(out.cpu().numpy(), exp)


@pytest.mark.gpu
def test_gpu_in_place_through_the_c_abi(gpu_device):
    import torch

    rng = np.random.default_rng(4)
    img = rng.random((128, 96))
    t = torch.from_numpy(img).to(gpu_device)
    lib = _native.load()
    stream = torch.cuda.current_stream(gpu_device).cuda_stream
    status = lib.rlic_equalize_histogram_f64(t.data_ptr(), 96, 1, 128, 96, 100, t.data_ptr(),
                                             _native.RLIC_DEVICE, ctypes.c_void_p(stream))
    assert status == 0, _native.last_error()
    assert np.array_equal(t.cpu().numpy(), eq.equalize_histogram(img, 100))
