"""Seeded random configurations against the oracle, bit for bit, through
every path a call can take: the one-lane or paired trace, the L2-gather,
LDS-box and hybrid replays (``RLIC_REPLAY_LDS=2`` / ``RLIC_REPLAY_HYBRID=2``
put the LDS plans on small fields too), re-tracing every pass
(``RLIC_REPLAY=0``), and the batched call (one persistent launch, or
single-field launches on side streams).  Each configuration draws the
dtype, shape, kernel length (odd and even), arithmetic profile, uv mode,
each side's boundary, the number of passes and special values (NaN holes,
zero and signed-zero vectors)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu
from rlic_amd import _core
from rlic_amd.batch import Field, convolve_batch

from test_gpu_parity import _assert_same

pytestmark = pytest.mark.gpu

N_CONFIGS = 48
BOUNDS = ("closed", "periodic")


def _config(seed):
    rng = np.random.default_rng(1000 + seed)
    dtype = np.float32 if rng.random() < 0.5 else np.float64
    ny, nx = (int(v) for v in rng.integers(1, 161, size=2))
    klen = int(rng.integers(1, 202))
    tex = rng.random((ny, nx)).astype(dtype)
    u = (rng.random((ny, nx)) - 0.5).astype(dtype)
    v = (rng.random((ny, nx)) - 0.5).astype(dtype)
    if rng.random() < 0.4:
        u[rng.random((ny, nx)) < 0.03] = np.nan
        zero = rng.random((ny, nx)) < 0.05
        u[zero] = 0.0
        v[zero] = -0.0
    return dict(
        tex=tex, u=u, v=v, kernel=rng.random(klen).astype(dtype),
        profile=int(rng.integers(0, 4)),
        uv_mode="polarization" if rng.random() < 0.5 else "velocity",
        bounds=tuple(tuple(BOUNDS[int(b)] for b in rng.integers(0, 2, size=2)) for _ in range(2)),
        iterations=int(rng.integers(1, 5)),
    )


def _expected(c):
    (xl, xr), (yl, yr) = c["bounds"]
    return cpu.convolve(c["tex"], c["u"], c["v"], kernel=c["kernel"], uv_mode=c["uv_mode"],
                        boundaries={"x": (xl, xr), "y": (yl, yr)},
                        iterations=c["iterations"], profile=c["profile"])


def _label(seed, c):
    return (f"seed {seed}: {c['tex'].dtype.name} {c['tex'].shape} L={len(c['kernel'])} "
            f"p{c['profile']} {c['uv_mode']} {c['bounds']} n={c['iterations']}")


# (RLIC_REPLAY_LDS, RLIC_REPLAY_HYBRID, RLIC_REPLAY, RLIC_TRACE_PAIR_MAX)
PATHS = {
    "default": ("1", "1", "1", "131072"),
    "one-lane trace, L2 replay": ("0", "0", "1", "0"),
    "LDS and hybrid replays": ("2", "2", "1", "131072"),
    "whole LDS boxes only": ("2", "0", "1", "0"),
    "re-trace every pass": ("1", "1", "0", "131072"),
}


@pytest.mark.parametrize("path", list(PATHS))
def test_random_configs_single_calls(monkeypatch, path):
    lds, hybrid, replay, pair = PATHS[path]
    monkeypatch.setenv("RLIC_REPLAY_LDS", lds)
    monkeypatch.setenv("RLIC_REPLAY_HYBRID", hybrid)
    monkeypatch.setenv("RLIC_REPLAY", replay)
    monkeypatch.setenv("RLIC_TRACE_PAIR_MAX", pair)
    monkeypatch.setenv("RLIC_AUTOTUNE", "0")
    for seed in range(N_CONFIGS):
        c = _config(seed)
        fn = _core.convolve_f32 if c["tex"].dtype == np.float32 else _core.convolve_f64
        got = fn(c["tex"], (c["u"], c["v"], c["uv_mode"]), c["kernel"], c["bounds"],
                 c["iterations"], profile=c["profile"])
        _assert_same(got, _expected(c), f"{path}, {_label(seed, c)}")


@pytest.mark.parametrize("single_min", ["262144", "1"])
def test_random_configs_batched(monkeypatch, single_min):
    """All configurations of one dtype in one batched call (the profile is
    the batch's, so each field is checked against the oracle under it; an
    axis takes its left side's boundary on both sides, since the batched
    call validates as ``convolve`` does, which rejects a periodic side
    beside a closed one)."""
    monkeypatch.setenv("RLIC_BATCH_SINGLE_MIN", single_min)
    for dtype in (np.float32, np.float64):
        cfgs = []
        for s in range(N_CONFIGS):
            c = _config(s)
            if c["tex"].dtype == dtype:
                (xl, _), (yl, _) = c["bounds"]
                cfgs.append((s, dict(c, bounds=((xl, xl), (yl, yl)))))
        profile = 3
        fields = []
        for _, c in cfgs:
            (xl, xr), (yl, yr) = c["bounds"]
            fields.append(Field(c["tex"], c["u"], c["v"], c["kernel"], uv_mode=c["uv_mode"],
                                boundaries={"x": (xl, xr), "y": (yl, yr)},
                                iterations=c["iterations"]))
        outs = convolve_batch(fields, profile=profile)
        for (seed, c), out in zip(cfgs, outs):
            c = dict(c, profile=profile)
            _assert_same(out, _expected(c), f"batched, {_label(seed, c)}")
