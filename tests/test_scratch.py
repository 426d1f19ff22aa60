"""Call scratch reuse: calls on one stream share a cached block (grown as
needed), host threads on one stream take turns, separate streams get
separate blocks; every result matches the oracle (GPU)."""

from __future__ import annotations

import threading

import numpy as np
import pytest

import rlic_amd
from oracle import cpu


def _case(seed, ny, nx, klen, iterations):
    rng = np.random.default_rng(seed)
    tex = rng.random((ny, nx))
    u = rng.random((ny, nx)) - 0.5
    v = rng.random((ny, nx)) - 0.5
    k = rng.random(klen)
    exp = cpu.convolve(tex, u, v, kernel=k, uv_mode="velocity", boundaries="periodic",
                       iterations=iterations)
    return (tex, u, v, k, iterations), exp


def _run(dev, case, stream=None):
    import torch

    (tex, u, v, k, it) = case
    t = [torch.from_numpy(a).to(dev) for a in (tex, u, v, k)]
    if stream is None:
        out = rlic_amd.convolve(t[0], t[1], t[2], kernel=t[3], boundaries="periodic", iterations=it)
    else:
        with torch.cuda.stream(stream):
            out = rlic_amd.convolve(t[0], t[1], t[2], kernel=t[3], boundaries="periodic",
                                    iterations=it)
        stream.synchronize()
    return out.cpu().numpy()


@pytest.mark.gpu
def test_scratch_reuse_sizes_threads_streams(gpu_device):
    import torch

    cases = [_case(s, ny, nx, klen, it) for s, (ny, nx, klen, it) in enumerate(
        [(40, 50, 17, 3), (130, 90, 41, 2), (30, 30, 9, 4), (64, 64, 33, 1)])]
    # growing and shrinking on the current stream, twice
    for _ in range(2):
        for case, exp in cases:
            assert np.array_equal(_run(gpu_device, case), exp)
    # host threads on one stream
    errors = []

    def worker(idx):
        try:
            for _ in range(3):
                case, exp = cases[idx % len(cases)]
                if not np.array_equal(_run(gpu_device, case), exp):
                    errors.append(idx)
        except Exception as exc:  # pragma: no cover - reported below
            errors.append(repr(exc))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors
    # separate streams
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for case, exp in cases:
        assert np.array_equal(_run(gpu_device, case, s1), exp)
        assert np.array_equal(_run(gpu_device, case, s2), exp)


@pytest.mark.gpu
def test_scratch_release_then_reuse(gpu_device):
    from rlic_amd import _native

    case, exp = _case(9, 50, 60, 21, 3)
    assert np.array_equal(_run(gpu_device, case), exp)
    _native.raise_for_status(_native.load().rlic_scratch_release())
    assert np.array_equal(_run(gpu_device, case), exp)  # allocates again
