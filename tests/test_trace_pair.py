"""The paired trace (``lic_trace_pair``: the two directions of a pixel's
line on two lanes of one wave, for fields too small to fill the GPU) gives
the bits of the one-lane trace (``lic_trace``) and of the oracle: with and
without the move record the later passes replay, over every arithmetic
profile, both uv modes, the boundary kinds and the special fields (NaN
holes, zero vectors, exact crossing ties).  Where it does not apply (even
kernel lengths, whose two directions take different step counts, and
kernels whose backward values overflow the LDS) the call falls back to the
one-lane trace.  ``RLIC_TRACE_PAIR_MAX`` (pixels) switches it on or off.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu
from rlic_amd import _core, _native

from test_gpu_parity import BOUNDS, MODES, PROFILES, _assert_same, _bpair, _fields

pytestmark = pytest.mark.gpu

ALWAYS = str(1 << 30)


def _run(monkeypatch, pair_max, call):
    """call() with the paired trace allowed up to pair_max pixels; the
    result and the names of the kernels the call launched."""
    monkeypatch.setenv("RLIC_TRACE_PAIR_MAX", pair_max)
    _native.profiler_enable(True)
    _native.profiler_read()
    try:
        out = call()
        names = set(_native.profiler_read())
    finally:
        _native.profiler_enable(False)
    return out, names


def _check(monkeypatch, call, expected, label, paired=True):
    got, names = _run(monkeypatch, ALWAYS, call)
    assert ("lic_trace_pair" in names) == paired, f"{label}: kernels {sorted(names)}"
    _assert_same(got, expected, f"{label} (paired trace)")
    one, names = _run(monkeypatch, "0", call)
    assert "lic_trace_pair" not in names
    _assert_same(one, expected, f"{label} (one-lane trace)")


def _fn(dtype):
    return _core.convolve_f32 if np.dtype(dtype) == np.float32 else _core.convolve_f64


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("profile", PROFILES)
@pytest.mark.parametrize("uv_mode", MODES)
def test_paired_trace_matches(monkeypatch, dtype, profile, uv_mode):
    rng = np.random.default_rng(300 + profile)
    for b, (klen, iters) in zip(BOUNDS, [(33, 1), (17, 3), (65, 2), (9, 1)]):
        tex, u, v = _fields(rng, (45, 70), dtype)
        k = rng.random(klen).astype(dtype)
        exp = cpu.convolve(tex, u, v, kernel=k, uv_mode=uv_mode, boundaries=b,
                           iterations=iters, profile=profile)
        _check(monkeypatch,
               lambda: _fn(dtype)(tex, (u, v, uv_mode), k, _bpair(b), iters, profile=profile),
               exp, f"{np.dtype(dtype).name} p{profile} {uv_mode} {b} L={klen} n={iters}")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["holes", "ties"])
def test_paired_trace_special_fields(monkeypatch, dtype, kind):
    rng = np.random.default_rng(11)
    for uv_mode in MODES:
        for b in BOUNDS[:3]:
            for iters in (1, 2):
                tex, u, v = _fields(rng, (31, 33), dtype, kind)
                k = np.sin(np.arange(31) * np.pi / 31).astype(dtype)
                exp = cpu.convolve(tex, u, v, kernel=k, uv_mode=uv_mode, boundaries=b,
                                   iterations=iters, profile=3)
                _check(monkeypatch,
                       lambda: _fn(dtype)(tex, (u, v, uv_mode), k, _bpair(b), iters, profile=3),
                       exp, f"{kind} {uv_mode} {b} n={iters}")


# a one-wave block holds 32 x (kmid + 1) backward values, 8 bytes each in
# float64: kmid 639 fills 160 KiB of LDS, kmid 641 does not fit (the
# one-lane trace takes it)
@pytest.mark.parametrize("klen,paired", [(1, True), (3, True), (4, False), (64, False),
                                         (129, True), (1279, True), (1283, False)])
def test_paired_trace_kernel_lengths(monkeypatch, klen, paired):
    rng = np.random.default_rng(klen)
    tex, u, v = _fields(rng, (40, 24), np.float64)
    k = rng.random(klen)
    exp = cpu.convolve(tex, u, v, kernel=k, uv_mode="polarization", boundaries="periodic",
                       iterations=2)
    _check(monkeypatch,
           lambda: _core.convolve_f64(tex, (u, v, "polarization"), k,
                                      _bpair("periodic"), 2),
           exp, f"L={klen}", paired=paired)


@pytest.mark.parametrize("shape", [(1, 1), (1, 17), (17, 1), (3, 700), (250, 190)])
def test_paired_trace_shapes(monkeypatch, shape):
    rng = np.random.default_rng(sum(shape))
    for dtype in (np.float32, np.float64):
        tex, u, v = _fields(rng, shape, dtype)
        k = rng.random(41).astype(dtype)
        for b in ("closed", "periodic"):
            exp = cpu.convolve(tex, u, v, kernel=k, boundaries=b, iterations=1)
            _check(monkeypatch, lambda: _fn(dtype)(tex, (u, v, "velocity"), k, _bpair(b), 1),
                   exp, f"{shape} {np.dtype(dtype).name} {b}")


def test_paired_trace_readme_example(monkeypatch):
    """The README example (256 x 256 float64, L = 65, one pass), which the
    default threshold sends through the paired trace."""
    rng = np.random.default_rng(0)
    n = 256
    tex = rng.random((n, n))
    x = np.linspace(0, np.pi, n)
    U = np.broadcast_to(np.cos(2 * x), (n, n)).copy()
    V = np.broadcast_to(np.sin(x), (n, n)).T.copy()
    k = 1 - np.abs(np.linspace(-1, 1, 65))
    exp = cpu.convolve(tex, U, V, kernel=k, boundaries="periodic", iterations=1)
    _check(monkeypatch, lambda: _core.convolve_f64(tex, (U, V, "velocity"), k,
                                                   _bpair("periodic"), 1), exp, "README")
    monkeypatch.delenv("RLIC_TRACE_PAIR_MAX")
    _native.profiler_enable(True)
    _native.profiler_read()
    try:
        got = _core.convolve_f64(tex, (U, V, "velocity"), k, _bpair("periodic"), 1)
        names = set(_native.profiler_read())
    finally:
        _native.profiler_enable(False)
    assert "lic_trace_pair" in names
    _assert_same(got, exp, "README (default threshold)")
