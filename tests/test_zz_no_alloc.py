"""The allocation-free entry points rlic_convolve_noalloc_f32/_f64: same
bits as rlic_convolve_f32/_f64 on device arrays (GPU), two calls in a row
sharing the fixed scratch array, the second smaller; a move record that
does not fit the array (every pass traced); sizes beyond the array
rejected before any work (CPU)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle_check import oracle_convolve
from rlic_amd import _native

LIMIT = 64 << 20  # RLIC_NOALLOC_SCRATCH_BYTES


def _args(tex, u, v, kernel, out, *, uv_mode=0, bounds=(1, 1, 0, 0), iterations=1):
    """Arguments of rlic_convolve_* / rlic_convolve_noalloc_* (without
    location) for C-contiguous torch tensors or NumPy arrays."""
    def ptr(a):
        return a.data_ptr() if hasattr(a, "data_ptr") else a.ctypes.data

    ny, nx = tex.shape
    return [ptr(tex), nx, 1, ptr(u), nx, 1, ptr(v), nx, 1, ny, nx, ptr(kernel),
            kernel.shape[0], uv_mode, *bounds, iterations, ptr(out)]


def _fields(rng, shape, klen, dtype):
    tex = rng.random(shape).astype(dtype)
    u = (rng.random(shape) - 0.5).astype(dtype)
    v = (rng.random(shape) - 0.5).astype(dtype)
    u[rng.random(shape) < 0.01] = np.nan
    return tex, u, v, rng.random(klen).astype(dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_gives_the_bits_of_convolve(gpu_device, dtype):
    import torch

    lib = _native.load()
    f32 = dtype == np.float32
    ref = lib.rlic_convolve_f32 if f32 else lib.rlic_convolve_f64
    noalloc = lib.rlic_convolve_noalloc_f32 if f32 else lib.rlic_convolve_noalloc_f64
    stream = torch.cuda.current_stream(gpu_device)
    rng = np.random.default_rng(70 if f32 else 71)
    # two calls in a row on the shared array, the second smaller; then
    # one-pass and zero-pass calls
    cases = [((300, 260), 65, 4, 1, (1, 1, 0, 0)), ((70, 90), 33, 3, 0, (0, 0, 1, 1)),
             ((129, 77), 17, 1, 1, (1, 0, 0, 1)), ((40, 50), 9, 0, 0, (0, 0, 0, 0))]
    for (shape, klen, iters, mode, bounds) in cases:
        arrays = _fields(rng, shape, klen, dtype)
        tex, u, v, k = (torch.from_numpy(a).to(gpu_device) for a in arrays)
        want = torch.empty_like(tex)
        got = torch.full_like(tex, np.nan)
        a_ref = _args(tex, u, v, k, want, uv_mode=mode, bounds=bounds, iterations=iters)
        a_na = _args(tex, u, v, k, got, uv_mode=mode, bounds=bounds, iterations=iters)
        assert ref(*a_ref, _native.RLIC_DEVICE, 3, ctypes.c_void_p(stream.cuda_stream)) == 0
        assert noalloc(*a_na, 3, ctypes.c_void_p(stream.cuda_stream)) == 0, \
            _native.last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, want), (shape, klen, iters)
        if shape[0] * shape[1] <= 70 * 90:
            exp = oracle_convolve(*arrays[:3], kernel=arrays[3],
                                  uv_mode=("velocity", "polarization")[mode],
                                  boundaries={"x": tuple(("closed", "periodic")[b] for b in bounds[:2]),
                                              "y": tuple(("closed", "periodic")[b] for b in bounds[2:])},
                                  iterations=iters)
            assert np.array_equal(got.cpu().numpy(), exp, equal_nan=True), (shape, klen, iters)


@pytest.mark.gpu
def test_noalloc_traces_every_pass_when_the_record_does_not_fit(gpu_device):
    """1024 x 1536 float32, L = 257, 3 passes: weights, cells and ping-pong
    buffer fit the fixed array (31.5 MB), the 201 MB move record does not,
    so every pass is traced; the bits are those of the replaying call."""
    import torch

    lib = _native.load()
    stream = torch.cuda.current_stream(gpu_device)
    rng = np.random.default_rng(72)
    arrays = _fields(rng, (1024, 1536), 257, np.float32)
    tex, u, v, k = (torch.from_numpy(a).to(gpu_device) for a in arrays)
    want = torch.empty_like(tex)
    got = torch.empty_like(tex)
    s = ctypes.c_void_p(stream.cuda_stream)
    assert lib.rlic_convolve_f32(*_args(tex, u, v, k, want, iterations=3),
                                 _native.RLIC_DEVICE, 3, s) == 0
    assert lib.rlic_convolve_noalloc_f32(*_args(tex, u, v, k, got, iterations=3), 3, s) == 0, \
        _native.last_error()
    torch.cuda.synchronize()
    assert torch.equal(got, want)


def test_noalloc_rejects_sizes_beyond_its_array():
    """Checked before any HIP call: no GPU needed."""
    lib = _native.load()
    a = np.ones((4, 4), dtype=np.float32)
    args = _args(a, a, a, np.ones(3, dtype=np.float32), a, iterations=2)
    # 2048 x 2048 float32: 67 MB of guarded cells alone
    args[9], args[10] = 2048, 2048
    assert lib.rlic_convolve_noalloc_f32(*args, 3, None) == _native.RLIC_ERR_INVALID_ARGUMENT
    msg = _native.last_error()
    assert "bytes of scratch" in msg and str(LIMIT) in msg
    # argument errors come first, as in rlic_convolve_*
    assert lib.rlic_convolve_noalloc_f32(*args, 9, None) == _native.RLIC_ERR_INVALID_ARGUMENT
    assert "unknown arithmetic profile" in _native.last_error()
    args[9] = 0  # an empty image: nothing to do, no device touched
    assert lib.rlic_convolve_noalloc_f32(*args, 3, None) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_call_captured_in_one_graph(gpu_device, dtype):
    """A whole allocation-free call — kernel copy, pack, trace with its move
    record and the LDS-staged replays of passes 2..3 — captured in one HIP
    graph (torch.cuda.graph) and replayed: each replay, with the texture
    changed in place and eager work and synchronizations in between, gives
    the bits of an eager call on the same inputs (and of the oracle)."""
    import torch

    lib = _native.load()
    noalloc = lib.rlic_convolve_noalloc_f32 if dtype == np.float32 else lib.rlic_convolve_noalloc_f64
    rng = np.random.default_rng(72)
    shape, klen, iters = (512, 512), 65, 3  # 2^18 pixels: the replays stage LDS boxes
    arrays = _fields(rng, shape, klen, dtype)
    tex, u, v, k = (torch.from_numpy(a).to(gpu_device) for a in arrays)
    out = torch.empty_like(tex)
    eager_out = torch.empty_like(tex)
    bounds = (1, 1, 0, 0)

    def call(dst, stream):
        a = _args(tex, u, v, k, dst, uv_mode=1, bounds=bounds, iterations=iters)
        return noalloc(*a, 3, ctypes.c_void_p(stream.cuda_stream))

    side = torch.cuda.Stream(gpu_device)
    side.wait_stream(torch.cuda.current_stream(gpu_device))
    with torch.cuda.stream(side):
        assert call(out, side) == 0, _native.last_error()  # first call: scratch address
    side.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        status = call(out, torch.cuda.current_stream(gpu_device))
    assert status == 0, _native.last_error()
    stream = torch.cuda.current_stream(gpu_device)
    for replay in range(3):
        new_tex = rng.random(shape).astype(dtype)
        tex.copy_(torch.from_numpy(new_tex))
        out.fill_(float("nan"))
        graph.replay()
        torch.cuda.synchronize()
        assert call(eager_out, stream) == 0, _native.last_error()
        torch.cuda.synchronize()
        assert torch.equal(out, eager_out), replay
        if replay == 0:
            exp = oracle_convolve(new_tex, arrays[1], arrays[2], kernel=arrays[3],
                                  uv_mode="polarization",
                                  boundaries={"x": "periodic", "y": "closed"}, iterations=iters)
            assert np.array_equal(out.cpu().numpy(), exp, equal_nan=True)
