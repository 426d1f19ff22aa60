"""Divergence measurement (RLIC_TRACE_STATS=1): live and executed
lane-steps of a tracing pass against values derived by hand (GPU)."""

from __future__ import annotations

import numpy as np
import pytest

import rlic_amd
from rlic_amd import _native


def _run(monkeypatch, *args, **kw):
    monkeypatch.setenv("RLIC_TRACE_STATS", "1")
    _native.trace_stats_read()
    out = rlic_amd.convolve(*args, **kw)
    stats = _native.trace_stats_read()
    monkeypatch.delenv("RLIC_TRACE_STATS")
    return out, stats


@pytest.mark.gpu
@pytest.mark.parametrize("iterations", [1, 3])
def test_uniform_flow_with_a_nan_column(monkeypatch, iterations):
    ny, nx, c, klen = 40, 64, 30, 21
    n_fwd = n_bwd = klen // 2
    rng = np.random.default_rng(1)
    tex = rng.random((ny, nx))
    u = np.ones((ny, nx))
    v = np.zeros((ny, nx))
    u[:, c] = np.nan
    kernel = rng.random(klen)
    out, stats = _run(monkeypatch, tex, u, v, kernel=kernel, boundaries="closed",
                      iterations=iterations)
    assert np.array_equal(out, rlic_amd.convolve(tex, u, v, kernel=kernel, boundaries="closed",
                                                 iterations=iterations))
    # u = 1, v = 0: every step moves one column right (forward) or left
    # (backward); a line ends at the step that reads the NaN column
    # (src/lib.rs:336-338): d steps were accumulated before it
    j = np.arange(nx)
    d_fwd = np.where(j == c, 0, np.where(j < c, np.minimum(n_fwd, c - j), n_fwd))
    d_bwd = np.where(j == c, 0, np.where(j > c, np.minimum(n_bwd, j - c), n_bwd))
    live = ny * int(d_fwd.sum() + d_bwd.sum())
    # a wave is an 8 x 8 pixel square; it runs groups of 8 steps until the
    # last of its lines has ended
    executed = 0
    for d, n in ((d_fwd, n_fwd), (d_bwd, n_bwd)):
        for j0 in range(0, nx, 8):
            dd = d[j0:j0 + 8]
            groups = [x // 8 if x < n else 10**9 for x in dd]
            steps = min(n, 8 * (max(groups) + 1))
            executed += steps * 8 * ny  # every square column holds 8 rows of pixels
    assert stats["live_lane_steps"] == live  # one tracing pass per call
    assert stats["executed_lane_steps"] == executed
    assert stats["lane_efficiency"] < 1.0


@pytest.mark.gpu
def test_no_nan_field_loses_nothing(monkeypatch):
    rng = np.random.default_rng(2)
    ny, nx, klen = 70, 90, 33
    tex, u, v = rng.random((ny, nx)), rng.random((ny, nx)) - 0.5, rng.random((ny, nx)) - 0.5
    _, stats = _run(monkeypatch, tex, u, v, kernel=rng.random(klen), boundaries="periodic",
                    iterations=2)
    assert stats["live_lane_steps"] == stats["executed_lane_steps"] == ny * nx * (klen - 1)
