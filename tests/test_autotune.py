"""Replay autotuner: the first multi-pass call for a new key times every
candidate plan on its second pass and keeps the fastest; every plan gives
the same bits (GPU), and the table round-trips through its text cache."""

from __future__ import annotations

import numpy as np
import pytest

import rlic_amd
from rlic_amd import _native


@pytest.mark.gpu
@pytest.mark.parametrize("klen, lds_fits", [(65, True), (301, False)])
def test_tuned_calls_give_the_same_bits(monkeypatch, tmp_path, klen, lds_fits):
    import torch

    rng = np.random.default_rng(klen)
    ny, nx = 520, 530  # >= 2^18 pixels: tuned
    t = lambda a: torch.from_numpy(a.astype(np.float32)).to("cuda:0")  # noqa: E731
    tex, u, v = t(rng.random((ny, nx))), t(rng.random((ny, nx)) - 0.5), t(rng.random((ny, nx)) - 0.5)
    kernel = t(rng.random(klen))
    kw = dict(kernel=kernel, boundaries={"x": "periodic", "y": "closed"}, iterations=4)
    monkeypatch.setenv("RLIC_AUTOTUNE", "0")
    reference = rlic_amd.convolve(tex, u, v, **kw).cpu().numpy()
    cache = tmp_path / "plans.txt"
    monkeypatch.setenv("RLIC_AUTOTUNE", "1")
    monkeypatch.setenv("RLIC_AUTOTUNE_CACHE", str(cache))
    _native.autotune_clear()
    first = rlic_amd.convolve(tex, u, v, **kw).cpu().numpy()    # tunes (pass 2 x candidates)
    second = rlic_amd.convolve(tex, u, v, **kw).cpu().numpy()   # uses the tuned plan
    assert np.array_equal(first, reference)
    assert np.array_equal(second, reference)
    table = _native.autotune_table()
    entry = [e for e in table if (e["ny"], e["nx"], e["kernel_len"]) == (ny, nx, klen)]
    assert len(entry) == 1 and entry[0]["pass_ms"] > 0
    if not lds_fits:
        assert entry[0]["replay"] in ("l2", "lds_hybrid")
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("rlic-replay-plan 4 520 530")
    # every candidate plan, loaded one at a time, gives the same bits
    monkeypatch.delenv("RLIC_AUTOTUNE_CACHE")
    # hybrid plans carry the box halo in the band field: 64 (the default
    # for float) and 32
    plans = (["1 64 64 0 0", "1 32 64 0 0", "1 32 32 0 0"] if lds_fits
             else ["2 64 64 64 0", "2 64 64 32 0"]) + \
        ["0 0 0 16 3", "0 0 0 4 3", "0 0 0 8 4"]
    for plan in plans:
        one = tmp_path / "one.txt"
        one.write_text(f"rlic-replay-plan 4 {ny} {nx} {klen} 3 {plan} 1.0\n")
        _native.autotune_clear()
        assert _native.autotune_load(one) == 1
        out = rlic_amd.convolve(tex, u, v, **kw).cpu().numpy()
        assert np.array_equal(out, reference), plan
    _native.autotune_clear()


def test_table_load_and_read_without_a_gpu(tmp_path):
    path = tmp_path / "plans.txt"
    path.write_text("rlic-replay-plan 8 100 200 33 3 1 32 64 0 0 0.5\n"
                    "rlic-replay-plan 4 64 64 9 1 0 0 0 16 3 0.25\n"
                    "rlic-replay-plan 4 80 90 257 3 2 64 64 64 0 0.75\n")
    _native.autotune_clear()
    assert _native.autotune_load(path) == 3
    table = _native.autotune_table()
    assert {(e["dtype_size"], e["ny"], e["replay"]) for e in table} == \
        {(8, 100, "lds"), (4, 64, "l2"), (4, 80, "lds_hybrid")}
    with pytest.raises(ValueError, match="cannot open"):
        _native.autotune_load(tmp_path / "missing.txt")
    _native.autotune_clear()
    assert _native.autotune_table() == []
