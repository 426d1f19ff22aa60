"""On-device input generation (``rlic_amd.generate``).

The noise kernel must reproduce, bit for bit, a NumPy restatement of its
counter-based hash (splitmix64, Steele, Lea and Flood, OOPSLA 2014), which
is itself pinned to the published output sequence of splitmix64 for seed 0.
The Fourier-mode kernel is checked against the host generator
``rlic_amd.workloads.fourier_modes`` (same draws, NumPy's cos and sin): the
two differ only by the last-place rounding of cos and sin, so the check is
to 1e-12 in double precision (values are sums of 8 modes of amplitude
N(0, 1)) and to one float32 spacing after rounding to float32.
"""

from __future__ import annotations

import numpy as np
import pytest

from rlic_amd.workloads import fourier_modes as host_fourier_modes

_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def _splitmix64(z: np.ndarray) -> np.ndarray:
    """splitmix64's output function on uint64 states (wrapping arithmetic)."""
    z = z.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _host_noise(shape, seed: int, dtype) -> np.ndarray:
    """The noise kernel's values: hash of seed + (p + 1) * gamma, top bits
    as a binary fraction in [0, 1)."""
    n = int(np.prod(shape))
    with np.errstate(over="ignore"):
        state = np.uint64(seed) + (np.arange(n, dtype=np.uint64) + np.uint64(1)) * _GAMMA
    z = _splitmix64(state)
    if np.dtype(dtype) == np.float64:
        vals = (z >> np.uint64(11)).astype(np.float64) * 2.0**-53
    else:
        vals = (z >> np.uint64(40)).astype(np.float32) * np.float32(2.0**-24)
    return vals.reshape(shape).astype(dtype)


def test_splitmix64_restatement_known_answers():
    # the published sequence of splitmix64 seeded with 0: states k * gamma
    states = np.arange(1, 5, dtype=np.uint64) * _GAMMA
    got = [int(v) for v in _splitmix64(states)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F,
                   0xF88BB8A8724C81EC]


def test_host_noise_is_uniform_in_unit_interval():
    for dtype in (np.float32, np.float64):
        a = _host_noise((300, 400), 7, dtype)
        assert a.min() >= 0 and a.max() < 1
        assert abs(a.mean() - 0.5) < 0.005
        assert len(np.unique(a)) > 0.99 * a.size


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(0, 5), (1, 1), (37, 129), (256, 256), (3, 100_003)])
@pytest.mark.parametrize("seed", [0, 12345, 2**64 - 1])
def test_noise_matches_restatement(dtype, shape, seed, gpu_device):
    from rlic_amd import generate

    got = generate.noise(shape, seed=seed, dtype=dtype, device=gpu_device)
    assert tuple(got.shape) == shape and got.device == gpu_device
    np.testing.assert_array_equal(got.cpu().numpy(), _host_noise(shape, seed, dtype))


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 1), (64, 48), (130, 257)])
def test_fourier_modes_match_host_generator(dtype, shape, gpu_device):
    from rlic_amd import generate

    exp = host_fourier_modes(np.random.default_rng(11), shape)
    got = generate.random_fourier_modes(np.random.default_rng(11), shape, dtype=dtype,
                                        device=gpu_device).cpu().numpy()
    assert got.dtype == dtype
    if np.dtype(dtype) == np.float64:
        np.testing.assert_allclose(got, exp, rtol=0, atol=1e-12)
    else:
        exp32 = exp.astype(np.float32)
        assert np.all(np.abs(got - exp32) <= np.spacing(np.abs(exp32)))


@pytest.mark.gpu
def test_generated_inputs_convolve_like_host_ones(gpu_device):
    """Device-generated inputs run through convolve like any device tensors:
    the result equals the oracle on the same values copied to the host."""
    import torch

    import rlic_amd
    from oracle import cpu
    from rlic_amd import generate

    shape = (96, 160)
    tex = generate.noise(shape, seed=3, dtype=np.float64, device=gpu_device)
    u = generate.random_fourier_modes(np.random.default_rng(1), shape, dtype=np.float64,
                                      device=gpu_device)
    v = generate.random_fourier_modes(np.random.default_rng(2), shape, dtype=np.float64,
                                      device=gpu_device)
    k = np.hanning(33)
    got = rlic_amd.convolve(tex, u, v, kernel=torch.from_numpy(k).to(gpu_device),
                            boundaries="periodic", iterations=2)
    exp = cpu.convolve(tex.cpu().numpy(), u.cpu().numpy(), v.cpu().numpy(), kernel=k,
                       boundaries="periodic", iterations=2)
    np.testing.assert_array_equal(got.cpu().numpy(), exp)


def test_mode_count_is_bounded():
    from rlic_amd import generate

    with pytest.raises(ValueError, match="at most 32 modes"):
        generate.fourier_modes((4, 4), amplitude=np.ones(33), kx=np.ones(33), ky=np.ones(33),
                               phase=np.zeros(33), device="cpu")
