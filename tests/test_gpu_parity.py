"""GPU parity: the HIP path against the oracle, bit for bit.

Every case runs the product path (``librlic_hip.so`` through the C ABI) on
the GPU and compares with the CPU oracle on identical inputs.  Both fp64
and fp32 are required to match exactly: the kernels use the reference's
operations in the reference's order with IEEE rounding, so there is no
tolerance to grant.
"""

from __future__ import annotations

import itertools
import os

import numpy as np
import pytest

import rlic_amd
from oracle import cpu
from rlic_amd import _core

pytestmark = pytest.mark.gpu

PROFILES = [0, 1, 2, 3]
MODES = ["velocity", "polarization"]
BOUNDS = [
    "closed",
    "periodic",
    {"x": "periodic", "y": "closed"},
    {"x": ("closed", "closed"), "y": "periodic"},
]


def _fields(rng, shape, dtype, kind="random"):
    ny, nx = shape
    tex = rng.random(shape).astype(dtype)
    if kind == "random":
        u = (rng.random(shape) - 0.5).astype(dtype)
        v = (rng.random(shape) - 0.5).astype(dtype)
    elif kind == "holes":  # NaN holes, exact zeros, signed zeros
        u = (rng.random(shape) - 0.5).astype(dtype)
        v = (rng.random(shape) - 0.5).astype(dtype)
        u[rng.random(shape) < 0.05] = np.nan
        v[rng.random(shape) < 0.05] = np.nan
        zero = rng.random(shape) < 0.1
        u[zero] = 0.0
        v[zero] = -0.0
        v[:, nx // 3] = 0.0
        u[ny // 2, :] = -0.0
    elif kind == "ties":  # unit half-plane fields: exact tx == ty ties
        jj = np.broadcast_to(np.arange(nx), shape)
        ii = np.broadcast_to(np.arange(ny)[:, None], shape)
        u = np.where(jj < nx / 2, -1.0, 1.0).astype(dtype)
        v = np.where(ii < ny / 2, -1.0, 1.0).astype(dtype)
    else:
        raise ValueError(kind)
    return tex, u, v


def _native(dtype):
    return _core.convolve_f32 if np.dtype(dtype) == np.float32 else _core.convolve_f64


def _bpair(b):
    if isinstance(b, str):
        return ((b, b), (b, b))
    bx = (b["x"], b["x"]) if isinstance(b["x"], str) else tuple(b["x"])
    by = (b["y"], b["y"]) if isinstance(b["y"], str) else tuple(b["y"])
    return (bx, by)


def _assert_same(got, expected, label=""):
    assert got.dtype == expected.dtype and got.shape == expected.shape
    if not np.array_equal(got, expected, equal_nan=True):
        diff = ~((got == expected) | (np.isnan(got) & np.isnan(expected)))
        n = int(np.count_nonzero(diff))
        i, j = np.argwhere(diff)[0]
        raise AssertionError(
            f"{label}: {n}/{got.size} pixels differ; first at ({i},{j}): "
            f"{got[i, j]!r} vs oracle {expected[i, j]!r}"
        )


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("profile", PROFILES)
@pytest.mark.parametrize("uv_mode", MODES)
def test_random_fields_all_profiles(dtype, profile, uv_mode):
    rng = np.random.default_rng(100 + profile)
    for b, (klen, iters) in zip(BOUNDS, [(11, 1), (12, 2), (33, 3), (5, 1)]):
        tex, u, v = _fields(rng, (37, 53), dtype)
        k = rng.random(klen).astype(dtype)
        exp = cpu.convolve(tex, u, v, kernel=k, uv_mode=uv_mode, boundaries=b,
                           iterations=iters, profile=profile)
        got = _native(dtype)(tex, (u, v, uv_mode), k, _bpair(b), iters, profile=profile)
        _assert_same(got, exp, f"{dtype.__name__} p{profile} {uv_mode} {b} L={klen} n={iters}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["holes", "ties"])
@pytest.mark.parametrize("profile", [1, 3])
def test_special_fields(dtype, kind, profile):
    rng = np.random.default_rng(7)
    for uv_mode, b in itertools.product(MODES, BOUNDS[:3]):
        tex, u, v = _fields(rng, (31, 33), dtype, kind)
        k = np.sin(np.arange(31) * np.pi / 31).astype(dtype)
        exp = cpu.convolve(tex, u, v, kernel=k, uv_mode=uv_mode, boundaries=b,
                           iterations=2, profile=profile)
        got = _native(dtype)(tex, (u, v, uv_mode), k, _bpair(b), 2, profile=profile)
        _assert_same(got, exp, f"{kind} {uv_mode} {b}")


@pytest.mark.parametrize("klen", [1, 2, 3, 4, 5, 64, 65, 129])
def test_kernel_lengths(klen):
    rng = np.random.default_rng(klen)
    tex, u, v = _fields(rng, (40, 24), np.float64)
    k = rng.random(klen)
    exp = cpu.convolve(tex, u, v, kernel=k, uv_mode="polarization", boundaries="periodic",
                       iterations=2)
    got = rlic_amd.convolve(tex, u, v, kernel=k, uv_mode="polarization",
                            boundaries="periodic", iterations=2)
    _assert_same(got, exp, f"L={klen}")


@pytest.mark.parametrize("shape", [(1, 1), (1, 17), (17, 1), (2, 3), (16, 16), (17, 15), (5, 300)])
def test_ragged_shapes(shape):
    rng = np.random.default_rng(sum(shape))
    for dtype in (np.float32, np.float64):
        tex, u, v = _fields(rng, shape, dtype)
        k = rng.random(9).astype(dtype)
        for b in ("closed", "periodic"):
            exp = cpu.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
            got = rlic_amd.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
            _assert_same(got, exp, f"{shape} {b}")


def test_empty_images():
    for shape in [(0, 5), (5, 0), (0, 0)]:
        tex = np.zeros(shape)
        out = rlic_amd.convolve(tex, tex, tex, kernel=np.ones(3))
        assert out.shape == shape
        # with nothing to compute the reference never reads the kernel
        out = _core.convolve_f64(tex, (tex, tex, "velocity"), np.ones(0),
                                 (("closed",) * 2,) * 2, 2)
        assert out.shape == shape


def test_empty_kernel_raises():
    tex = np.ones((4, 4))
    with pytest.raises(ValueError, match="kernel is empty"):
        rlic_amd.convolve(tex, tex, tex, kernel=np.ones(0))


def test_strided_inputs():
    """Transposed, sliced, negative-stride and broadcast views (any strides)."""
    rng = np.random.default_rng(3)
    big = rng.random((80, 90))
    tex = big[5:45, 10:70:2]  # sliced and strided
    u = (rng.random((30, 40)) - 0.5).T  # F-order
    v = (rng.random((40, 30))[::-1] - 0.5)  # negative row stride
    k = rng.random(13)
    exp = cpu.convolve(np.ascontiguousarray(tex), np.ascontiguousarray(u),
                       np.ascontiguousarray(v), kernel=k, iterations=2)
    got = rlic_amd.convolve(tex, u, v, kernel=k, iterations=2)
    _assert_same(got, exp, "strided")
    ub = np.broadcast_to(rng.random(30) - 0.5, (40, 30))
    vb = np.broadcast_to((rng.random(40) - 0.5)[:, None], (40, 30))
    exp = cpu.convolve(np.ascontiguousarray(tex), np.ascontiguousarray(ub),
                       np.ascontiguousarray(vb), kernel=k, boundaries="periodic")
    got = rlic_amd.convolve(tex, ub, vb, kernel=k, boundaries="periodic")
    _assert_same(got, exp, "broadcast")


@pytest.mark.parametrize("boundaries", [
    "closed",
    "periodic",
    {"x": "periodic", "y": "closed"},
])
def test_retrace_path_matches_replay(boundaries):
    """Iterations re-traced every pass (no move record) give the same bits;
    between passes the texture of every cell is refreshed, including the
    guard cells that stand for the boundary-rule targets."""
    rng = np.random.default_rng(11)
    tex, u, v = _fields(rng, (50, 61), np.float32, "holes")
    k = rng.random(21).astype(np.float32)
    kw = dict(kernel=k, uv_mode="polarization", iterations=4, boundaries=boundaries)
    exp = cpu.convolve(tex, u, v, **kw)
    got_replay = rlic_amd.convolve(tex, u, v, **kw)
    os.environ["RLIC_REPLAY"] = "0"
    try:
        got_retrace = rlic_amd.convolve(tex, u, v, **kw)
    finally:
        del os.environ["RLIC_REPLAY"]
    _assert_same(got_replay, exp, "replay")
    _assert_same(got_retrace, exp, "retrace")


def test_readme_example_config():
    """The reference README example: 256x256 float64, L=65, periodic, 1 iteration."""
    shape = (256, 256)
    texture = np.random.default_rng(0).random(shape)
    x = np.linspace(0, np.pi, 256)
    U = np.broadcast_to(np.cos(2 * x), shape)
    V = np.broadcast_to(np.sin(x).T, shape)
    kernel = 1 - np.abs(np.linspace(-1, 1, 65))
    exp = cpu.convolve(texture, U, V, kernel=kernel, boundaries="periodic")
    for n in (1, 5):
        got = rlic_amd.convolve(texture, U, V, kernel=kernel, boundaries="periodic",
                                iterations=n)
        if n == 1:
            _assert_same(got, exp, "README example")
        else:
            _assert_same(got, cpu.convolve(texture, U, V, kernel=kernel,
                                           boundaries="periodic", iterations=n), "n=5")


def test_torch_device_tensors(gpu_device):
    import torch

    rng = np.random.default_rng(5)
    for dtype in (np.float32, np.float64):
        tex, u, v = _fields(rng, (45, 70), dtype, "holes")
        k = rng.random(17).astype(dtype)
        exp = cpu.convolve(tex, u, v, kernel=k, uv_mode="polarization",
                           boundaries={"x": "periodic", "y": "closed"}, iterations=3)
        t = [torch.from_numpy(np.ascontiguousarray(a)).to(gpu_device) for a in (tex, u, v, k)]
        got = rlic_amd.convolve(t[0], t[1], t[2], kernel=t[3], uv_mode="polarization",
                                boundaries={"x": "periodic", "y": "closed"}, iterations=3)
        assert got.device == gpu_device
        _assert_same(got.cpu().numpy(), exp, f"torch {dtype.__name__}")
        # transposed device views are read in place
        got_t = rlic_amd.convolve(t[0].T, t[2].T, t[1].T, kernel=t[3],
                                  boundaries={"x": "closed", "y": "periodic"}, iterations=1)
        exp_t = cpu.convolve(np.ascontiguousarray(tex.T), np.ascontiguousarray(v.T),
                             np.ascontiguousarray(u.T), kernel=k,
                             boundaries={"x": "closed", "y": "periodic"})
        _assert_same(got_t.cpu().numpy(), exp_t, "torch transposed")


@pytest.mark.parametrize("wave_w_log2", ["3", "4", "5"])
@pytest.mark.parametrize("band", ["1", "16"])
def test_launch_shapes_do_not_change_results(wave_w_log2, band, monkeypatch):
    """Wave tile shape and tile ordering only move work between lanes."""
    monkeypatch.setenv("RLIC_WAVE_W_LOG2", wave_w_log2)
    monkeypatch.setenv("RLIC_TILE_BAND", band)
    rng = np.random.default_rng(int(wave_w_log2) * 10 + int(band))
    tex, u, v = _fields(rng, (75, 131), np.float32, "holes")
    k = rng.random(25).astype(np.float32)
    for b in ("closed", {"x": "periodic", "y": ("closed", "closed")}):
        exp = cpu.convolve(tex, u, v, kernel=k, uv_mode="polarization", boundaries=b,
                           iterations=3)
        got = rlic_amd.convolve(tex, u, v, kernel=k, uv_mode="polarization", boundaries=b,
                                iterations=3)
        _assert_same(got, exp, f"wave 2^{wave_w_log2} band {band} {b}")


def test_library_and_torch_share_one_hip_runtime():
    """Host-array calls before `import torch` must leave torch's device usable."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import numpy as np, rlic_amd\n"
        "a = np.random.default_rng(0).random((16, 16))\n"
        "rlic_amd.convolve(a, a, a, kernel=np.ones(5))\n"
        "import torch\n"
        "assert torch.cuda.is_available(), 'torch lost the device'\n"
        "t = torch.from_numpy(a).cuda()\n"
        "out = rlic_amd.convolve(t, t, t, kernel=torch.ones(5, dtype=t.dtype, device=t.device))\n"
        "assert np.array_equal(out.cpu().numpy(), rlic_amd.convolve(a, a, a, kernel=np.ones(5)))\n"
        "print('ok')\n" % str(root)
    )
    proc = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          timeout=240)
    assert proc.returncode == 0 and proc.stdout.strip().endswith("ok"), proc.stdout + proc.stderr


@pytest.mark.parametrize("lds", ["0", "2"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_replay_paths_match(lds, dtype, monkeypatch):
    """Replay from LDS-staged tiles and from global gathers give the same bits,
    including lines that wrap across periodic edges of images smaller than
    the staged box."""
    monkeypatch.setenv("RLIC_REPLAY_LDS", lds)
    rng = np.random.default_rng(21 + int(lds))
    for shape, klen in (((70, 90), 33), ((20, 24), 65), ((33, 47), 64)):
        tex, u, v = _fields(rng, shape, dtype, "holes")
        k = rng.random(klen).astype(dtype)
        for b in ("periodic", {"x": "periodic", "y": "closed"}, "closed"):
            exp = cpu.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
            got = rlic_amd.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
            _assert_same(got, exp, f"lds={lds} {shape} L={klen} {b}")


@pytest.mark.parametrize("hybrid", ["2", "0"])
@pytest.mark.parametrize("dtype,klen", [
    (np.float32, 65),   # 64x64 LDS tiles
    (np.float32, 137),  # hybrid; without it 32x64 LDS tiles
    (np.float32, 161),  # hybrid; without it 32x32 LDS tiles
    (np.float32, 181),  # hybrid; without it global gathers
    (np.float64, 73),   # hybrid; without it 64x64
    (np.float64, 81),   # hybrid; without it 32x64
])
def test_replay_tile_shapes(dtype, klen, hybrid, monkeypatch):
    monkeypatch.setenv("RLIC_REPLAY_HYBRID", hybrid)
    rng = np.random.default_rng(klen)
    tex, u, v = _fields(rng, (70, 130), dtype, "holes")
    k = rng.random(klen).astype(dtype)
    for b in ("periodic", {"x": ("closed", "closed"), "y": "periodic"}):
        exp = cpu.convolve(tex, u, v, kernel=k, uv_mode="polarization", boundaries=b,
                           iterations=2)
        got = rlic_amd.convolve(tex, u, v, kernel=k, uv_mode="polarization", boundaries=b,
                                iterations=2)
        _assert_same(got, exp, f"{np.dtype(dtype).name} L={klen} {b} hybrid={hybrid}")


@pytest.mark.parametrize("shape", [(33, 1), (14, 107), (65, 40), (63, 129), (130, 3), (128, 64)])
def test_lds_box_wraps_around_small_images(shape):
    """The LDS replay stages a box of tile + kmid halo wrapped modulo the
    image: with one add or subtract when the box lies within one image size
    of the image, else with a modulo.  Images around and below the box size,
    periodic on both axes so that lines cross the edges, both dtypes."""
    rng = np.random.default_rng(shape[0] * 1000 + shape[1])
    for dtype in (np.float32, np.float64):
        tex, u, v = _fields(rng, shape, dtype, "holes")
        for klen in (3, 17, 33, 65):
            k = rng.random(klen).astype(dtype)
            for b in ("periodic", {"x": "periodic", "y": "closed"}):
                exp = cpu.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
                got = rlic_amd.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
                _assert_same(got, exp, f"{shape} {np.dtype(dtype).name} L={klen} {b}")


@pytest.mark.parametrize("profile", [2, 3])
def test_extreme_velocity_magnitudes(profile):
    """Float, branchless profiles: velocities of extreme magnitude (2^-75 to
    2^75), exact zeros, denormals, values near the float maximum and NaN
    holes scattered over a field, so that step times over- and underflow
    and fractions leave [0, 1]; every mode and boundary set must give the
    oracle's bits."""
    rng = np.random.default_rng(60 + profile)
    shape = (96, 160)
    tex, u, v = _fields(rng, shape, np.float32, "holes")
    scale = np.exp2(rng.uniform(-75, 75, shape))
    pick = rng.random(shape)
    u = np.where(pick < 0.08, u * scale, u).astype(np.float32)
    v = np.where((pick > 0.92), v * scale, v).astype(np.float32)
    u[rng.random(shape) < 0.01] = np.float32(1e-42)   # denormal
    v[rng.random(shape) < 0.01] = np.float32(3e38)
    for klen in (9, 33, 65):
        k = rng.random(klen).astype(np.float32)
        for uv_mode in MODES:
            for b in BOUNDS:
                exp = cpu.convolve(tex, u, v, kernel=k, uv_mode=uv_mode, boundaries=b,
                                   iterations=2, profile=profile)
                got = _core.convolve_f32(tex, (u, v, uv_mode), k, _bpair(b), 2, profile=profile)
                _assert_same(got, exp, f"profile={profile} L={klen} {uv_mode} {b}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_replay_from_unaligned_output_buffer(dtype, gpu_device):
    """With an even number of passes, pass 3 reads the caller's output
    buffer.  The LDS replay stages its box with 16-byte copies only when that
    buffer is 16-byte aligned (and the row length and kernel half-length are
    multiples of the copy); an output one element past an aligned address
    takes the 4-byte copies.  Both must give the oracle's bits."""
    import ctypes

    import torch

    from rlic_amd import _native

    rng = np.random.default_rng(77)
    shape, klen = (150, 256), 65           # nx and kmid multiples of 4
    tex, u, v = _fields(rng, shape, dtype, "holes")
    k = rng.random(klen).astype(dtype)
    lib = _native.load()
    fn = lib.rlic_convolve_f32 if np.dtype(dtype) == np.float32 else lib.rlic_convolve_f64
    t = [torch.from_numpy(a).to(gpu_device) for a in (tex, u, v, k)]
    n = tex.size
    for iterations in (4, 6):
        exp = cpu.convolve(tex, u, v, kernel=k, boundaries="periodic", iterations=iterations)
        for shift in (0, 1):
            store = torch.full((n + 4,), float("nan"), dtype=t[0].dtype, device=gpu_device)
            out = store[shift:shift + n].view(shape)
            assert (out.data_ptr() % 16 == 0) == (shift == 0)
            stream = torch.cuda.current_stream(gpu_device).cuda_stream
            status = fn(t[0].data_ptr(), shape[1], 1, t[1].data_ptr(), shape[1], 1,
                        t[2].data_ptr(), shape[1], 1, shape[0], shape[1],
                        t[3].data_ptr(), klen, 0, 1, 1, 1, 1, iterations, out.data_ptr(),
                        _native.RLIC_DEVICE, 3, ctypes.c_void_p(stream))
            _native.raise_for_status(status)
            torch.cuda.synchronize(gpu_device)
            _assert_same(out.cpu().numpy(), exp, f"{np.dtype(dtype).name} n={iterations} shift={shift}")
            # the elements around the output are untouched
            rest = torch.cat([store[:shift], store[shift + n:]]).cpu().numpy()
            assert np.isnan(rest).all()


def test_dlpack_device_arrays(gpu_device):
    """Device arrays of another library that exports DLPack are read in
    place; the result is a device tensor with the oracle's bits."""
    import torch

    class DLPackOnly:
        def __init__(self, t):
            self._t = t

        def __dlpack__(self, *args, **kwargs):
            return self._t.__dlpack__(*args, **kwargs)

        def __dlpack_device__(self):
            return self._t.__dlpack_device__()

    rng = np.random.default_rng(91)
    tex, u, v = _fields(rng, (70, 90), np.float32, "holes")
    k = rng.random(33).astype(np.float32)
    exp = cpu.convolve(tex, u, v, kernel=k, boundaries="periodic", iterations=3)
    t = [torch.from_numpy(a).to(gpu_device) for a in (tex, u, v, k)]
    got = rlic_amd.convolve(*(DLPackOnly(a) for a in t[:3]), kernel=DLPackOnly(t[3]),
                            boundaries="periodic", iterations=3)
    assert isinstance(got, torch.Tensor) and got.device == gpu_device
    _assert_same(got.cpu().numpy(), exp, "dlpack")


@pytest.mark.parametrize("klen", [33, 34, 35, 36, 40, 66])
def test_lds_box_kernel_lengths(klen):
    """LDS replay with kernel half-lengths 16 to 33, odd and even lengths
    (an even length has one forward step fewer than backward steps):
    straight lines (unit fields) reach the box's outer rows and columns;
    both dtypes, with 4-byte and 16-byte box copies (odd and even kmid)."""
    rng = np.random.default_rng(klen)
    for dtype in (np.float32, np.float64):
        for kind in ("ties", "random"):
            tex, u, v = _fields(rng, (150, 200), dtype, kind)
            k = rng.random(klen).astype(dtype)
            exp = cpu.convolve(tex, u, v, kernel=k, boundaries="periodic", iterations=3)
            got = rlic_amd.convolve(tex, u, v, kernel=k, boundaries="periodic", iterations=3)
            _assert_same(got, exp, f"{np.dtype(dtype).name} {kind} L={klen}")


@pytest.mark.parametrize("dtype, klen", [
    (np.float32, 131), (np.float32, 167), (np.float32, 200), (np.float32, 257),  # halo 64
    (np.float64, 111), (np.float64, 150), (np.float64, 257),                     # halo 32
])
def test_hybrid_replay(dtype, klen, monkeypatch):
    """Kernel lengths whose whole LDS box does not fit replay with the
    hybrid plan: the first 64 (float) or 32 (double) steps of each direction
    from the LDS box, the rest gathered from the image.  It must give the
    oracle's bits and the L2-gather replay's, on fields larger and smaller
    than the box, with NaN holes, periodic and closed sides."""
    rng = np.random.default_rng(klen)
    k = rng.random(klen).astype(dtype)
    for shape in ((150, 200), (40, 33), (130, 3)):
        for kind in ("holes", "ties"):
            tex, u, v = _fields(rng, shape, dtype, kind)
            for b in ("periodic", {"x": "periodic", "y": "closed"}):
                exp = cpu.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
                monkeypatch.setenv("RLIC_REPLAY_HYBRID", "2")  # also below 2^18 pixels
                got = rlic_amd.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
                _assert_same(got, exp, f"hybrid {shape} {kind} {b}")
                monkeypatch.setenv("RLIC_REPLAY_HYBRID", "0")
                l2 = rlic_amd.convolve(tex, u, v, kernel=k, boundaries=b, iterations=3)
                _assert_same(l2, exp, f"l2 {shape} {kind} {b}")


def test_hybrid_is_the_default_from_2_18_pixels(gpu_device, monkeypatch):
    """Untuned, a float field of at least 2^18 pixels whose whole box does
    not fit replays with the hybrid plan (a lic_replay_lds launch per pass);
    it gives the oracle's bits and the L2-gather replay's."""
    import torch

    from rlic_amd import _native

    rng = np.random.default_rng(201)
    shape, klen = (520, 530), 201
    tex, u, v = _fields(rng, shape, np.float32, "holes")
    k = rng.random(klen).astype(np.float32)
    t = [torch.from_numpy(a).to(gpu_device) for a in (tex, u, v, k)]
    kw = dict(kernel=t[3], boundaries={"x": "periodic", "y": "closed"}, iterations=3)
    monkeypatch.setenv("RLIC_AUTOTUNE", "0")
    monkeypatch.setenv("RLIC_REPLAY_LDS", "1")
    monkeypatch.setenv("RLIC_REPLAY_HYBRID", "1")
    _native.profiler_enable(True)
    _native.profiler_read()
    try:
        got = rlic_amd.convolve(t[0], t[1], t[2], **kw).cpu().numpy()
        prof = _native.profiler_read()
    finally:
        _native.profiler_enable(False)
    assert prof.get("lic_replay_lds", {}).get("count") == 2 and "lic_replay" not in prof
    monkeypatch.setenv("RLIC_REPLAY_HYBRID", "0")
    l2 = rlic_amd.convolve(t[0], t[1], t[2], **kw).cpu().numpy()
    _assert_same(got, l2, "hybrid vs L2")
    exp = cpu.convolve(tex, u, v, kernel=k, boundaries={"x": "periodic", "y": "closed"},
                       iterations=3, threads=8)
    _assert_same(got, exp, "hybrid vs oracle")


@pytest.mark.parametrize("shape, kernel_name", [
    ((300, 400), "lic_replay"),        # below 2^18 pixels: L2 gathers
    ((520, 530), "lic_replay_lds"),    # from 2^18 pixels: the 64x64 LDS box
])
def test_default_replay_plan_by_image_size(shape, kernel_name, gpu_device, monkeypatch):
    """With the default switches, images below 2^18 pixels replay with L2
    gathers and larger ones from LDS boxes; both give the oracle's bits."""
    import torch

    from rlic_amd import _native

    rng = np.random.default_rng(shape[0])
    tex, u, v = _fields(rng, shape, np.float32, "random")
    k = rng.random(65).astype(np.float32)
    t = [torch.from_numpy(a).to(gpu_device) for a in (tex, u, v, k)]
    monkeypatch.setenv("RLIC_AUTOTUNE", "0")
    monkeypatch.setenv("RLIC_REPLAY_LDS", "1")
    monkeypatch.setenv("RLIC_REPLAY_HYBRID", "1")
    _native.profiler_enable(True)
    _native.profiler_read()
    try:
        got = rlic_amd.convolve(t[0], t[1], t[2], kernel=t[3], boundaries="periodic",
                                iterations=2).cpu().numpy()
        prof = _native.profiler_read()
    finally:
        _native.profiler_enable(False)
    replays = {name for name in prof if name.startswith("lic_replay")}
    assert replays == {kernel_name}
    exp = cpu.convolve(tex, u, v, kernel=k, boundaries="periodic", iterations=2, threads=8)
    _assert_same(got, exp, f"{shape}")
