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


# ---------------------------------------------------------------------------
# Adaptive equalization: sliding tile and tile interpolation (the reference's
# declared, unimplemented stubs, src/rlic/_core.pyi:38-57).
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sliding_tile_reduces_to_global_and_to_one(dtype):
    rng = np.random.default_rng(11)
    img = rng.random((13, 21)).astype(dtype)
    img[4, 5] = np.nan
    g = eq.equalize_histogram(img, 32)
    # a window of 2n - 1 (or more) covers the whole image from every pixel
    for shape in [(25, 41), (26, 60), (100, 100)]:
        assert np.array_equal(eq.equalize_histogram_sliding_tile(img, 32, shape), g,
                              equal_nan=True), shape
    # a 1 x 1 window holds only the pixel itself
    one = eq.equalize_histogram_sliding_tile(img, 32, (1, 1))
    assert np.array_equal(one, np.where(np.isnan(img), img, dtype(1)), equal_nan=True)


@pytest.mark.parametrize("shape", [(3, 3), (4, 5), (2, 7), (1, 6), (9, 2)])
def test_sliding_tile_bruteforce(shape):
    rng = np.random.default_rng(sum(shape))
    img = rng.random((11, 9))
    img[rng.random(img.shape) < 0.1] = np.nan
    img[2, 2] = img[6, 7]
    for nbins in (1, 5, 40):
        brute = eq.equalize_histogram_sliding_tile_bruteforce(img, nbins, shape)
        for method in ("shifted", "integral"):
            assert np.array_equal(eq.equalize_histogram_sliding_tile(img, nbins, shape, method),
                                  brute, equal_nan=True), (shape, nbins, method)


def test_tile_interpolation_closed_forms():
    rng = np.random.default_rng(12)
    img = rng.random((10, 14))
    g = eq.equalize_histogram(img, 16)
    for shape in [(10, 14), (64, 64)]:  # one tile: the global mapping
        assert np.array_equal(eq.equalize_histogram_tile_interpolation(img, 16, shape), g)
    # two constant tiles side by side, 0 on the left and 1 on the right,
    # nbins = 2: the left tile maps both bins to 1, the right one bin 0 to 0
    # and bin 1 to 1.  Tile centres at columns 1.5 and 5.5: a left pixel in
    # column j gets lerp(1, 0, w) = 1 - w, w = clamp((j - 1.5) / 4); right
    # pixels get 1 everywhere.
    two = np.zeros((3, 8))
    two[:, 4:] = 1.0
    out = eq.equalize_histogram_tile_interpolation(two, 2, (3, 4))
    w = np.clip((np.arange(4) - 1.5) / 4.0, 0.0, None)
    assert np.array_equal(out[:, :4], np.broadcast_to(1.0 - w, (3, 4)))
    assert np.array_equal(out[:, 4:], np.ones((3, 4)))
    # constant image: every mapping is 1
    assert np.array_equal(eq.equalize_histogram_tile_interpolation(np.full((7, 9), 3.0), 8,
                                                                   (2, 4)), np.ones((7, 9)))
    # at the centre of a full odd tile the weights vanish: the tile's own
    # equalization (with the global bins) of that pixel
    img = rng.random((15, 15))
    out = eq.equalize_histogram_tile_interpolation(img, 8, (5, 5))
    b, _, _ = eq.bins(img, 8)
    for ty in range(3):
        for tx in range(3):
            i, j = 5 * ty + 2, 5 * tx + 2
            tile = b[5 * ty:5 * ty + 5, 5 * tx:5 * tx + 5]
            assert out[i, j] == np.count_nonzero(tile <= b[i, j]) / 25.0


def test_tile_interpolation_non_finite_and_empty_tiles():
    img = np.random.default_rng(13).random((8, 8))
    img[:4, :4] = np.nan  # a tile without finite pixels takes the global mapping
    img[6, 6] = np.inf
    out = eq.equalize_histogram_tile_interpolation(img, 4, (4, 4))
    assert np.all(np.isnan(out[:4, :4])) and out[6, 6] == np.inf
    assert np.all((out[np.isfinite(img)] > 0) & (out[np.isfinite(img)] <= 1))


def test_tile_interpolation_planted_mistake_is_caught(monkeypatch):
    # tile centres without the half-pixel shift: the two-tile closed form fails
    two = np.zeros((3, 8))
    two[:, 4:] = 1.0
    real = eq._centres

    def shifted(n, t):
        return real(n, t) + 0.5

    monkeypatch.setattr(eq, "_centres", shifted)
    out = eq.equalize_histogram_tile_interpolation(two, 2, (3, 4))
    monkeypatch.undo()
    w = np.clip((np.arange(4) - 1.5) / 4.0, 0.0, None)
    assert not np.array_equal(out[:, :4], np.broadcast_to(1.0 - w, (3, 4)))


def test_adaptive_argument_errors():
    for fn in (rlic_amd.equalize_histogram_sliding_tile,
               rlic_amd.equalize_histogram_tile_interpolation):
        with pytest.raises(ValueError, match="tile_shape"):
            fn(np.ones((3, 3)), 4, (0, 2))
        with pytest.raises(TypeError, match="tile_shape"):
            fn(np.ones((3, 3)), 4, 3)
        with pytest.raises(ValueError, match="nbins"):
            fn(np.ones((3, 3)), 0, (2, 2))
    lib = _native.load()
    a = np.ones((2, 2))
    for name in ("rlic_equalize_histogram_sliding_tile_f64",
                 "rlic_equalize_histogram_tile_interpolation_f64"):
        fn = getattr(lib, name)
        assert fn(a.ctypes.data, 2, 1, 2, 2, 4, 0, 3, a.ctypes.data, 0, None) == \
            _native.RLIC_ERR_INVALID_ARGUMENT
        assert "tile_shape" in _native.last_error()
        assert fn(a.ctypes.data, 2, 1, 0, 2, 4, 2, 2, a.ctypes.data, 0, None) == 0  # empty
    big = np.ones((4096, 4096))
    st = lib.rlic_equalize_histogram_tile_interpolation_f64(big.ctypes.data, 4096, 1, 4096, 4096,
                                                            1 << 24, 8, 8, big.ctypes.data, 0,
                                                            None)
    assert st == _native.RLIC_ERR_INVALID_ARGUMENT and "tile counters" in _native.last_error()


_ADAPTIVE_SHAPES = [(1, 1), (5, 5), (8, 3), (31, 17), (64, 64), (400, 7), (301, 301)]


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nbins", [1, 16, 256, 20000])
def test_gpu_sliding_tile_matches_oracle(gpu_device, dtype, nbins):
    """Windows from 1 x 1 to larger than the images (301 x 301: its staged
    bins exceed the LDS, so the window is counted from global memory)."""
    rng = np.random.default_rng(nbins + 100)
    for name, img, t in _images(rng, dtype, gpu_device):
        for shape in _ADAPTIVE_SHAPES:
            if nbins == 20000 and shape[0] * shape[1] > 64:
                continue  # many distinct bins under a big window: a slow oracle, same kernels
            exp = eq.equalize_histogram_sliding_tile(img, nbins, shape)
            got = rlic_amd.equalize_histogram_sliding_tile(img, nbins, shape)
            assert np.array_equal(got, exp, equal_nan=True), (name, shape, nbins)
            dev = rlic_amd.equalize_histogram_sliding_tile(t, nbins, shape)
            assert np.array_equal(dev.cpu().numpy(), exp, equal_nan=True), (name, shape, "device")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("nbins", [1, 16, 256, 20000])
def test_gpu_tile_interpolation_matches_oracle(gpu_device, dtype, nbins):
    """Tiles from 1 x 1 (every pixel its own tile) to larger than the
    images; 20000 bins take the global-atomic tile histograms."""
    rng = np.random.default_rng(nbins + 200)
    for name, img, t in _images(rng, dtype, gpu_device):
        for shape in _ADAPTIVE_SHAPES:
            if nbins == 20000 and shape in ((1, 1), (5, 5), (8, 3)):
                continue  # tiles x bins beyond what the oracle builds quickly
            exp = eq.equalize_histogram_tile_interpolation(img, nbins, shape)
            got = rlic_amd.equalize_histogram_tile_interpolation(img, nbins, shape)
            assert np.array_equal(got, exp, equal_nan=True), (name, shape, nbins)
            dev = rlic_amd.equalize_histogram_tile_interpolation(t, nbins, shape)
            assert np.array_equal(dev.cpu().numpy(), exp, equal_nan=True), (name, shape, "device")


@pytest.mark.gpu
def test_gpu_adaptive_on_convolve_output_in_place(gpu_device):
    """A device LIC image equalized in place by both adaptive kernels
    through the C ABI, against the oracle on the same image."""
    import torch

    rng = np.random.default_rng(14)
    shape = (257, 383)
    t = [torch.from_numpy(a).to(gpu_device) for a in (
        rng.random(shape).astype(np.float32), (rng.random(shape) - 0.5).astype(np.float32),
        (rng.random(shape) - 0.5).astype(np.float32), np.ones(31, dtype=np.float32))]
    lic = rlic_amd.convolve(t[0], t[1], t[2], kernel=t[3], boundaries="closed", iterations=2)
    host = lic.cpu().numpy()
    lib = _native.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream(gpu_device).cuda_stream)
    for name, oracle, tile in (("sliding_tile", eq.equalize_histogram_sliding_tile, (33, 65)),
                               ("tile_interpolation", eq.equalize_histogram_tile_interpolation,
                                (64, 48))):
        work = lic.clone()
        fn = getattr(lib, f"rlic_equalize_histogram_{name}_f32")
        status = fn(work.data_ptr(), shape[1], 1, shape[0], shape[1], 256, tile[0], tile[1],
                    work.data_ptr(), _native.RLIC_DEVICE, stream)
        assert status == 0, _native.last_error()
        assert np.array_equal(work.cpu().numpy(), oracle(host, 256, tile)), name
