"""Host threads, each on a stream of its own, calling the library at the
same time: single calls of assorted shapes, kernel lengths, profiles and
plans (paired and one-lane traces, L2, LDS-box and hybrid replays, the
autotuner on a field large enough to tune) and batched calls (persistent
launch and side streams) overlap on the GPU, and every result still
equals the oracle's."""

from __future__ import annotations

import threading

import numpy as np
import pytest

from rlic_amd import _core
from rlic_amd.batch import Field, convolve_batch

from test_gpu_parity import _assert_same
from test_random_configs import _config, _expected

pytestmark = pytest.mark.gpu


def _single(dev, stream, c):
    import torch

    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (c["tex"], c["u"], c["v"],
                                                                        c["kernel"])]
    fn = _core.convolve_f32 if c["tex"].dtype == np.float32 else _core.convolve_f64
    with torch.cuda.stream(stream):
        out = fn(t[0], (t[1], t[2], c["uv_mode"]), t[3], c["bounds"], c["iterations"],
                 profile=c["profile"])
    stream.synchronize()
    return out.cpu().numpy()


def _batch(dev, stream, cfgs):
    import torch

    fields = []
    for c in cfgs:
        (xl, _), (yl, _) = c["bounds"]
        t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev)
             for a in (c["tex"], c["u"], c["v"], c["kernel"])]
        fields.append(Field(t[0], t[1], t[2], t[3], uv_mode=c["uv_mode"],
                            boundaries={"x": (xl, xl), "y": (yl, yl)},
                            iterations=c["iterations"]))
    with torch.cuda.stream(stream):
        outs = convolve_batch(fields, profile=3)
    stream.synchronize()
    return [o.cpu().numpy() for o in outs]


def test_threads_on_separate_streams(gpu_device, monkeypatch):
    import torch

    monkeypatch.setenv("RLIC_BATCH_SINGLE_MIN", "4096")  # some batched fields as single launches
    rng = np.random.default_rng(77)
    # a field large enough for the autotuner (>= 2^18 pixels, >= 3 passes)
    big = dict(_config(0), profile=3, uv_mode="velocity", iterations=3,
               bounds=(("periodic", "periodic"), ("closed", "closed")))
    big["tex"] = rng.random((520, 520)).astype(np.float32)
    big["u"] = (rng.random((520, 520)) - 0.5).astype(np.float32)
    big["v"] = (rng.random((520, 520)) - 0.5).astype(np.float32)
    big["kernel"] = rng.random(65).astype(np.float32)
    singles = [_config(s) for s in range(24)] + [big]
    expected = [_expected(c) for c in singles]
    batch_sets = [[_config(s) for s in range(24, 36)], [_config(s) for s in range(36, 48)]]
    batch_sets = [[c for c in bs if c["tex"].dtype == np.float32] for bs in batch_sets]
    batch_expected = []
    for bs in batch_sets:
        exp = []
        for c in bs:
            (xl, _), (yl, _) = c["bounds"]
            exp.append(_expected(dict(c, profile=3, bounds=((xl, xl), (yl, yl)))))
        batch_expected.append(exp)

    errors: list[str] = []

    def single_worker(k):
        stream = torch.cuda.Stream(device=gpu_device)
        try:
            for rep in range(2):
                for idx in range(k, len(singles), 6):
                    got = _single(gpu_device, stream, singles[idx])
                    try:
                        _assert_same(got, expected[idx], f"thread {k} config {idx} rep {rep}")
                    except AssertionError as exc:
                        errors.append(str(exc))
        except Exception as exc:  # reported below
            errors.append(repr(exc))

    def batch_worker(k):
        stream = torch.cuda.Stream(device=gpu_device)
        try:
            for rep in range(2):
                outs = _batch(gpu_device, stream, batch_sets[k])
                for j, (got, exp) in enumerate(zip(outs, batch_expected[k])):
                    try:
                        _assert_same(got, exp, f"batch thread {k} field {j} rep {rep}")
                    except AssertionError as exc:
                        errors.append(str(exc))
        except Exception as exc:
            errors.append(repr(exc))

    threads = [threading.Thread(target=single_worker, args=(k,)) for k in range(6)]
    threads += [threading.Thread(target=batch_worker, args=(k,)) for k in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[:5]
