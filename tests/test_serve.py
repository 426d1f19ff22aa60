"""Serving: requests batched at pass granularity give, each, the oracle's
bits (and those of a single convolve call) — alone, in a mix, arriving
between steps, under a tight memory budget, as follow-ups of kept fields,
after eviction and re-admission, and on the serving workload's distribution
(GPU); the budget arithmetic and the C ABI's argument checks (CPU)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import rlic_amd
from oracle_check import assert_matches_oracle, oracle_convolve, same_bits, to_numpy
from rlic_amd import _native


def _requests(rng, n, max_side=100):
    reqs = []
    for k in range(n):
        dtype = np.float32 if k % 2 else np.float64
        ny, nx = int(rng.integers(1, max_side)), int(rng.integers(1, max_side))
        klen = int(rng.integers(1, 60))
        tex = rng.random((ny, nx)).astype(dtype)
        u = (rng.random((ny, nx)) - 0.5).astype(dtype)
        v = (rng.random((ny, nx)) - 0.5).astype(dtype)
        if k % 5 == 3:
            u[rng.random((ny, nx)) < 0.05] = np.nan
        reqs.append(dict(
            texture=tex, u=u, v=v, kernel=rng.random(klen).astype(dtype),
            uv_mode=("velocity", "polarization")[k % 3 == 0],
            boundaries={"x": ("periodic", "closed")[k % 2], "y": ("closed", "periodic")[k % 4 < 2]},
            iterations=int(rng.integers(0, 6)),
        ))
    return reqs


def _expected(r, iterations=None):
    """The oracle's result for request `r` with `iterations` passes in all
    (src/lib.rs:432-440 chains the passes: a follow-up of k passes after n
    equals one call with n + k); the GPU single call must give the same
    bits, and served results are compared with both."""
    n = r["iterations"] if iterations is None else iterations
    exp = oracle_convolve(r["texture"], r["u"], r["v"], kernel=r["kernel"],
                          uv_mode=r["uv_mode"], boundaries=r["boundaries"], iterations=n)
    single = rlic_amd.convolve(r["texture"], r["u"], r["v"], kernel=r["kernel"],
                               uv_mode=r["uv_mode"], boundaries=r["boundaries"], iterations=n)
    assert np.array_equal(to_numpy(single), exp, equal_nan=True), "single call vs oracle"
    return exp


def _submit(server, r, **kw):
    return server.submit(r["texture"], r["u"], r["v"], kernel=r["kernel"], uv_mode=r["uv_mode"],
                         boundaries=r["boundaries"], iterations=r["iterations"], **kw)


def test_field_bytes_arithmetic():
    lib = _native.load()
    # 64 x 64 f32, L = 33: weights 256 + guarded cells (66 x 66 x 16 B, 256-aligned)
    # + two buffers 16 KiB each
    assert lib.rlic_server_field_bytes_f32(64, 64, 33, 0, 1) == 256 + 69888 + 2 * 16384
    # keeping the move record adds 2 chunks (16 + 16 steps) of 16 B per pixel
    assert lib.rlic_server_field_bytes_f32(64, 64, 33, 1, 1) == 256 + 69888 + 2 * 16384 + 131072
    assert lib.rlic_server_field_bytes_f32(64, 64, 33, 0, 2) == 256 + 69888 + 2 * 16384 + 131072
    assert lib.rlic_server_field_bytes_f64(10, 10, 1, 0, 1) == 256 + 4608 + 2 * 1024
    assert lib.rlic_server_field_bytes_f32(-1, 4, 3, 0, 1) == -1


def test_server_abi_argument_errors():
    lib = _native.load()
    assert not lib.rlic_server_create(-1, 3, None)
    assert "negative budget" in _native.last_error()
    assert not lib.rlic_server_create(1 << 20, 9, None)
    assert "unknown arithmetic profile" in _native.last_error()
    n = ctypes.c_int64(0)
    assert lib.rlic_server_step(None, None, 0, ctypes.byref(n)) < 0
    assert lib.rlic_server_release(None, 1) < 0
    assert lib.rlic_server_extend(None, 1, 1) < 0


@pytest.mark.gpu
def test_requests_alone_and_in_a_mix_give_the_same_bits(gpu_device):
    from rlic_amd.serve import Server

    rng = np.random.default_rng(11)
    reqs = _requests(rng, 14)
    with Server(1 << 30, device=gpu_device) as mixed:
        handles = [_submit(mixed, r) for r in reqs[:8]]
        mixed.step()  # the rest arrive while the first ones run
        handles += [_submit(mixed, r) for r in reqs[8:]]
        mixed.run()
    for k, (r, h) in enumerate(zip(reqs, handles)):
        assert h.state == "done", k
        with Server(1 << 30, device=gpu_device) as alone:
            a = _submit(alone, r)
            alone.run()
        exp = _expected(r)
        assert np.array_equal(a.result, exp, equal_nan=True), f"request {k} alone"
        assert np.array_equal(h.result, exp, equal_nan=True), f"request {k} in the mix"
        assert h.admitted is not None and h.finished >= h.admitted >= h.arrival


@pytest.mark.gpu
def test_follow_ups_continue_kept_fields(gpu_device):
    from rlic_amd.serve import Server

    rng = np.random.default_rng(12)
    reqs = [r for r in _requests(rng, 6) if r["iterations"] > 0]
    with Server(1 << 30, device=gpu_device) as srv:
        first = [_submit(srv, r, keep=True) for r in reqs]
        srv.run()
        more = [srv.follow_up(h, 3) for h in first]
        srv.run()
        again = [srv.follow_up(h, 1) for h in more]
        srv.run()
        assert srv.readmissions == 0
    for k, r in enumerate(reqs):
        n = r["iterations"]
        assert np.array_equal(first[k].result, _expected(r, n), equal_nan=True), k
        assert np.array_equal(more[k].result, _expected(r, n + 3), equal_nan=True), k
        assert np.array_equal(again[k].result, _expected(r, n + 4), equal_nan=True), k


@pytest.mark.gpu
def test_tight_budget_admits_in_turn_and_evicts(gpu_device):
    from rlic_amd.serve import Server

    rng = np.random.default_rng(13)
    reqs = [r for r in _requests(rng, 10, max_side=64) if r["iterations"] > 0]
    lib = _native.load()
    biggest = max(lib.rlic_server_field_bytes_f64(64, 64, 60, 1, 2) for _ in reqs)
    budget = 2 * biggest  # room for about two fields at a time
    with Server(budget, device=gpu_device) as srv:
        hs = [_submit(srv, r, keep=True) for r in reqs]
        peak = 0
        while srv.queued or srv.active:
            srv.step()
            peak = max(peak, srv.usage()["used_bytes"])
        assert peak <= budget
        assert srv.evictions > 0
        # following up a request whose field was evicted: it is admitted
        # again from its inputs
        k_ev = next(k for k, h in enumerate(hs) if not srv.is_resident(h))
        f = srv.follow_up(hs[k_ev], 2)
        srv.run()
        assert srv.readmissions == 1
    for k, (r, h) in enumerate(zip(reqs, hs)):
        assert np.array_equal(h.result, _expected(r), equal_nan=True), k
    exp = _expected(reqs[k_ev], reqs[k_ev]["iterations"] + 2)
    assert np.array_equal(f.result, exp, equal_nan=True)


@pytest.mark.gpu
def test_released_blocks_are_reused_with_fresh_contents(gpu_device):
    """A server hands a released field's device block to the next field of
    the same size class (and trims its cache to stay within the budget):
    three rounds of requests over the same few shapes, each round with new
    textures, kernels and pass counts, every result equal to the oracle's
    (nothing of an earlier field's cells, weights or move record shows)."""
    from rlic_amd.serve import Server

    rng = np.random.default_rng(99)
    shapes = [(40, 56), (56, 40), (33, 33)]
    lib = _native.load()
    budget = 3 * int(lib.rlic_server_field_bytes_f64(56, 56, 41, 1, 3))
    with Server(budget, device=gpu_device) as srv:
        for rnd in range(3):
            reqs = []
            for k in range(6):
                dtype = np.float64 if k % 2 else np.float32
                ny, nx = shapes[(k + rnd) % 3]
                reqs.append(dict(
                    texture=rng.random((ny, nx)).astype(dtype),
                    u=(rng.random((ny, nx)) - 0.5).astype(dtype),
                    v=(rng.random((ny, nx)) - 0.5).astype(dtype),
                    kernel=rng.random(int(rng.choice([9, 25, 41]))).astype(dtype),
                    uv_mode=("velocity", "polarization")[(k + rnd) % 2],
                    boundaries={"x": "periodic", "y": ("closed", "periodic")[k % 2]},
                    iterations=int(rng.integers(1, 4))))
            hs = [_submit(srv, r) for r in reqs]
            srv.run()
            assert srv.usage()["used_bytes"] == 0 and srv.usage()["resident"] == 0
            for k, (r, h) in enumerate(zip(reqs, hs)):
                assert np.array_equal(h.result, _expected(r), equal_nan=True), (rnd, k)


@pytest.mark.gpu
def test_device_tensors_and_edge_cases(gpu_device):
    import torch

    from rlic_amd.serve import Server

    rng = np.random.default_rng(14)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu_device)  # noqa: E731
    tex = rng.random((40, 30))
    u = rng.random((40, 30)) - 0.5
    v = rng.random((40, 30)) - 0.5
    k = rng.random(9)
    with Server(1 << 28, device=gpu_device) as srv:
        a = srv.submit(t(tex), t(u), t(v), kernel=t(k), iterations=3)
        b = srv.submit(t(tex), t(u), t(v), kernel=t(k), iterations=0)  # a copy, no work
        c = srv.submit(t(np.ones((0, 5))), t(np.ones((0, 5))), t(np.ones((0, 5))),
                       kernel=t(k), iterations=2)  # empty field
        with pytest.raises(ValueError, match="^Found invalid texture element"):
            srv.submit(t(-tex), t(u), t(v), kernel=t(k))
        srv.run()
    assert isinstance(a.result, torch.Tensor) and a.result.device == torch.device(gpu_device)
    exp = _expected(dict(texture=tex, u=u, v=v, kernel=k, uv_mode="velocity",
                         boundaries="closed", iterations=3))
    assert np.array_equal(a.result.cpu().numpy(), exp)
    assert np.array_equal(b.result.cpu().numpy(), tex)
    assert c.result.shape == (0, 5)


@pytest.mark.gpu
def test_background_loop_with_concurrent_submitters(gpu_device):
    import threading
    import time

    from rlic_amd.serve import Server

    rng = np.random.default_rng(15)
    reqs = [r for r in _requests(rng, 18) if r["iterations"] > 0]
    results = {}
    with Server(1 << 30, device=gpu_device) as srv:
        srv.start()

        def client(idx):
            for k in range(idx, len(reqs), 3):
                h = _submit(srv, reqs[k], keep=(k % 2 == 0))
                time.sleep(0.002)  # arrivals while earlier requests run
                assert h.wait(60), k
                results[k] = h
                if h.keep:
                    f = srv.follow_up(h, 2)
                    assert f.wait(60), k
                    results[(k, "f")] = f

        threads = [threading.Thread(target=client, args=(i,)) for i in range(3)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        srv.run()
        assert srv.loop_error is None
        srv.stop()
    for k, r in enumerate(reqs):
        assert results[k].state == "done"
        assert np.array_equal(results[k].result, _expected(r), equal_nan=True), k
        if (k, "f") in results:
            exp = _expected(r, r["iterations"] + 2)
            assert np.array_equal(results[(k, "f")].result, exp, equal_nan=True), k


@pytest.mark.gpu
@pytest.mark.parametrize("streams", ["1", "4"])
def test_large_fields_replay_as_single_launches(gpu_device, monkeypatch, streams):
    """Replay passes of fields of at least 2^18 pixels run as single-field
    launches (LDS box or hybrid replay) inside the serving step, on the
    server's stream or spread over side streams, next to the persistent
    launch of the small fields; every request, follow-ups included, gives
    the bits of one convolve call."""
    from rlic_amd import _native
    from rlic_amd.serve import Server

    monkeypatch.setenv("RLIC_BATCH_STREAMS", streams)
    rng = np.random.default_rng(13)
    reqs = _requests(rng, 4)
    for k, (dtype, klen) in enumerate([(np.float32, 65), (np.float32, 201), (np.float64, 99)]):
        shape = (520, 530)
        u = (rng.random(shape) - 0.5).astype(dtype)
        u[rng.random(shape) < 0.02] = np.nan
        reqs.append(dict(
            texture=rng.random(shape).astype(dtype), u=u,
            v=(rng.random(shape) - 0.5).astype(dtype), kernel=rng.random(klen).astype(dtype),
            uv_mode=("velocity", "polarization")[k % 2],
            boundaries={"x": "periodic", "y": ("closed", "periodic")[k % 2]}, iterations=3))
    _native.profiler_enable(True)
    _native.profiler_read()
    try:
        with Server(1 << 32, device=gpu_device) as srv:
            handles = [_submit(srv, r, keep=True) for r in reqs]
            srv.run()
            more = [srv.follow_up(h, 2) for h in handles[-3:]]
            srv.run()
        prof = _native.profiler_read()
    finally:
        _native.profiler_enable(False)
    assert prof.get("lic_replay_lds", {}).get("count", 0) >= 3 * 4  # 3 fields x (2 + 2) passes
    for k, (r, h) in enumerate(zip(reqs, handles)):
        assert np.array_equal(h.result, _expected(r), equal_nan=True), k
    for r, h in zip(reqs[-3:], more):
        assert np.array_equal(h.result, _expected(r, 5), equal_nan=True)


@pytest.mark.gpu
def test_serving_workload_distribution(gpu_device):
    """The serving workload's distribution under its 64 GiB budget: fields
    of sides 128..2048 (aspect 1:4..4:1), float32 and float64 alternating,
    odd kernel lengths 17..257, random axes and uv_mode, 1..20 passes;
    requests arrive between steps (Poisson arrivals, counted in steps), half
    of them keep their field and a quarter of the arrivals follow up a
    finished kept request.  Every result equals the single call with the
    request's total passes, bit for bit; small fields equal the oracle
    whole, one-pass large fields on a band of rows."""
    import torch

    from rlic_amd.serve import Server
    from rlic_amd.workloads import serving_pool, serving_request

    pool = serving_pool(24, seed=5, device=gpu_device)
    rng = np.random.default_rng(55)
    n_req = 40
    # arrival step of each request: Poisson process of 1.5 arrivals per step
    arrive_at = np.floor(np.cumsum(rng.exponential(1 / 1.5, size=n_req))).astype(int)
    plans = [(serving_request(rng, pool), rng.random() < 0.25, rng.random() < 0.5,
              int(rng.integers(1, 21))) for _ in range(n_req)]
    done_reqs = []  # (request, spec dict, total passes)
    kept = []
    with Server(64 << 30, device=gpu_device) as srv:
        k = 0
        step = 0
        while k < n_req or srv.queued or srv.active:
            while k < n_req and arrive_at[k] <= step:
                r, follow, keep, more = plans[k]
                ready = [(q, spec) for q, spec in kept if q.state == "done"]
                if follow and ready:
                    base, spec = ready[int(rng.integers(len(ready)))]
                    kept.remove((base, spec))
                    q = srv.follow_up(base, more)
                    done_reqs.append((q, spec))
                    kept.append((q, spec))
                else:
                    q = _submit(srv, r, keep=keep)
                    done_reqs.append((q, r))
                    if keep:
                        kept.append((q, r))
                k += 1
            srv.step()
            step += 1
        assert srv.queued == 0 and srv.active == 0
    assert sum(q.base is not None for q, _ in done_reqs) >= 3
    assert {q.texture.dtype for q, _ in done_reqs} == {torch.float32, torch.float64}
    checked = {"whole": 0, "band": 0, None: 0}
    for idx, (q, spec) in enumerate(done_reqs):
        assert q.state == "done", idx
        n = q.total_iterations
        single = rlic_amd.convolve(spec["texture"], spec["u"], spec["v"], kernel=spec["kernel"],
                                   uv_mode=spec["uv_mode"], boundaries=spec["boundaries"],
                                   iterations=n)
        assert same_bits(q.result, single), f"request {idx} differs from its single call"
        if spec["texture"].numel() <= 256 * 256 or n == 1:
            how = assert_matches_oracle(q.result, spec["texture"], spec["u"], spec["v"],
                                        kernel=spec["kernel"], uv_mode=spec["uv_mode"],
                                        boundaries=spec["boundaries"], iterations=n,
                                        what=f"request {idx}")
            checked[how] += 1
    assert checked["whole"] + checked["band"] >= 4, checked
