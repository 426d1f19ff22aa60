#!/usr/bin/env python
"""Throughput benchmark of the LIC hot path on one MI355X.

Workload (default): plain LIC at scale — 4096x4096 float32, texture U[0,1),
u and v sums of 8 random Fourier modes, Gaussian kernel of 129 taps
(sigma 32), boundaries x periodic / y closed, velocity mode, 5 iterations.
One step = one full ``convolve`` call (pack, trace, replays) on inputs that
are already resident in HBM as device tensors.

Prints ONE JSON line:
  value     Gpixel-steps/s = H*W*(L-1)*iterations per step / step time
  roofline  for the kernel with the most total time in the timed region
            (lic_trace): it is bound by VALU instruction issue, so `achieved`
            is its VALU lane-instructions per launch (wave64 instructions from
            the committed rocprofv3 counter summary, x 64) / its average
            launch time measured live with HIP events on the launch stream,
            against MI355X's FP32 vector peak expressed in lane-instructions
            (157.3 TFLOP/s / 2 flops per FMA = 78.6 T/s: 128 lanes per CU per
            clock x 256 CUs x 2.4 GHz); `hbm` = counter-measured HBM bytes per
            launch / launch time against 8 TB/s; `logical_gather_gbs` = the
            reference algorithm's gathers ((3*(L-1)+2)*sizeof(T) per pixel)
            per launch / launch time, served mostly by caches and LDS.
  extras    the other configurations, each timed here: the README example
            (256^2 float64), the polarization fields (16384^2 float32 and
            8192^2 float64, generated on the device), the 512-field uneven
            batch (batched vs per-call speedup) and serving (Poisson arrivals:
            throughput, p50/p99 latency).  --no-extras skips them.
  cpu_baseline  the CPU oracle (a restatement of the reference's
            single-threaded algorithm) on a bounded band of rows of pass 1.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                       [--no-extras]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# MI355X FP32 vector peak, 157.3 TFLOP/s, in VALU lane-instructions per second:
# one FMA lane-instruction is 2 flops, i.e. 128 lanes x 256 CUs x 2.4 GHz
VALU_PEAK_TLANE = 157.3 / 2


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str, workload: str) -> float | None:
    """HBM bytes per launch from the committed counter summary, if any."""
    path = ROOT / "profiles" / f"traffic_{kernel}_{workload}.json"
    if not path.is_file():
        return None
    try:
        data = json.loads(path.read_text())
        return float(data["hbm_bytes_per_launch"])
    except (ValueError, KeyError):
        return None


def valu_from_profiles(kernel: str, workload: str) -> dict | None:
    """Wave64 VALU instructions and waves per launch from the committed
    counter summary (SQ_INSTS_VALU, SQ_WAVES), if any."""
    path = ROOT / "profiles" / f"valu_{kernel}_{workload}.json"
    if not path.is_file():
        return None
    try:
        data = json.loads(path.read_text())
        return {"insts": float(data["valu_insts_per_launch"]),
                "waves": float(data["waves_per_launch"])}
    except (ValueError, KeyError):
        return None


def measure_copy_gbs(dev, repeats: int = 10) -> float:
    """Measured HBM bandwidth of a plain device copy (STREAM copy): PyTorch's
    copy of 2^28 float32 (1 GiB read, 1 GiB written), timed with events on
    the stream it runs on."""
    import torch

    src = torch.rand((1 << 28,), device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    stream = torch.cuda.current_stream(dev)
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(repeats):
        dst.copy_(src)
    end.record(stream)
    end.synchronize()
    avg_s = start.elapsed_time(end) / repeats / 1e3
    del src, dst
    torch.cuda.empty_cache()
    return 2 * (1 << 30) / avg_s / 1e9


def cpu_baseline(wl, rows: int) -> dict:
    """Time the CPU oracle on pass 1 of a band of `rows` rows (one thread)."""
    from oracle import cpu

    ny, nx = wl.shape
    r0 = max(0, ny // 2 - rows // 2)
    r1 = min(ny, r0 + rows)
    t0 = time.perf_counter()
    cpu.convolve(wl.texture, wl.u, wl.v, kernel=wl.kernel, uv_mode=wl.uv_mode,
                 boundaries=wl.boundaries, iterations=1, rows=(r0, r1), threads=1)
    dt = time.perf_counter() - t0
    steps = (r1 - r0) * nx * (wl.kernel_len - 1)
    # the same band with OpenMP over rows on the host's cores (at most 16:
    # one GPU's share of the host), a parallel extension of the reference's
    # single-threaded loop
    threads = max(1, min(16, cpu.max_threads()))
    t0 = time.perf_counter()
    cpu.convolve(wl.texture, wl.u, wl.v, kernel=wl.kernel, uv_mode=wl.uv_mode,
                 boundaries=wl.boundaries, iterations=1, rows=(r0, r1), threads=threads)
    dt_par = time.perf_counter() - t0
    return {
        "value": steps / dt / 1e9,
        "unit": "Gpixel-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"pass 1 of rows [{r0},{r1}) x {nx} cols "
                  f"({steps / 1e6:.1f} M pixel-steps, {dt:.2f} s), single thread, "
                  f"host CPU: {os.cpu_count()} logical cores visible",
        "parallel": {"value": steps / dt_par / 1e9, "cores": threads,
                     "sample": f"the same rows, OpenMP over rows, {dt_par:.2f} s"},
    }


def batch_measure(dev, steps: int, warmup: int, small: bool = False) -> dict:
    """512 uneven fields (sides 128..2048, or 32..256 with `small`): one
    batched call vs a per-call loop of plain GPU calls, both on
    device-resident inputs, timed from a synchronized start to a
    synchronized end over `steps` repetitions after `warmup`."""
    import torch

    from rlic_amd import _core, _native
    from rlic_amd._boundaries import BoundarySet
    from rlic_amd.batch import convolve_batch
    from rlic_amd.workloads import batch_uneven, pixel_steps

    log(f"bench: generating 512 uneven fields on the device (sides "
        f"{'32..256' if small else '128..2048'})")
    fields = (batch_uneven(512, seed=4, device=dev, min_side=32, max_side=256) if small
              else batch_uneven(512, seed=4, device=dev))
    steps_total = pixel_steps(fields)
    bpairs = []
    for f in fields:
        bs = BoundarySet.from_spec(f["boundaries"])
        bpairs.append((bs.x, bs.y))

    def per_call():
        for f, bp in zip(fields, bpairs):
            _core.convolve_f32(f["texture"], (f["u"], f["v"], f["uv_mode"]), f["kernel"], bp,
                               f["iterations"])

    def batched():
        convolve_batch(fields)

    timings = {}
    for name, fn in (("per_call", per_call), ("batched", batched)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        timings[name] = (time.perf_counter() - t0) / steps
        log(f"bench: {name} {timings[name] * 1e3:.2f} ms per batch")
    # kernel times from one more, untimed batched call: the launch profiler's
    # events add host work, so it stays off while timing
    _native.profiler_enable(True)
    _native.profiler_read()
    batched()
    torch.cuda.synchronize()
    prof = _native.profiler_read()
    _native.profiler_enable(False)
    del fields
    torch.cuda.empty_cache()
    return {
        "workload": "batch_uneven_512_f32" + ("_small" if small else ""),
        "fields": 512, "sides": "32..256" if small else "128..2048",
        "pixel_steps_per_batch": steps_total,
        "batched": {"value": round(steps_total / timings["batched"] / 1e9, 3),
                    "ms_per_batch": round(timings["batched"] * 1e3, 3)},
        "per_call": {"value": round(steps_total / timings["per_call"] / 1e9, 3),
                     "ms_per_batch": round(timings["per_call"] * 1e3, 3)},
        "batched_vs_per_call_speedup": round(timings["per_call"] / timings["batched"], 3),
        "kernel_ms_per_batch": {k: round(v["total_ms"], 3) for k, v in prof.items()},
        "steps": steps, "warmup": warmup,
    }


def bench_batch(args, dev) -> int:
    """The uneven batch as the benchmark's one line."""
    m = batch_measure(dev, args.steps, args.warmup, small=args.workload == "batch512small")
    line = {
        "metric": "Gpixel-steps/s (batched uneven fields)",
        "value": m["batched"]["value"],
        "unit": "Gpixel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["batched"]["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": m["workload"], "fields": m["fields"], "sides": m["sides"],
                   "pixel_steps_per_step": m["pixel_steps_per_batch"]},
        "per_call": m["per_call"],
        "batched_vs_per_call_speedup": m["batched_vs_per_call_speedup"],
        "kernel_ms_per_step": m["kernel_ms_per_batch"],
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def serve_measure(dev, burst: int, n_requests: int, pool_fields: int = 96) -> dict:
    """Serving: Poisson arrivals of convolve requests on a continuous-batching
    server (64 GiB budget, fields of the uneven batch's distribution, float32
    and float64, 1..20 passes, a quarter of the arrivals following up a
    finished kept request).  The arrival rate is 70% of the capacity
    measured on a burst of `burst` requests arriving at once."""
    import torch

    from rlic_amd.serve import Server
    from rlic_amd.workloads import serving_pool, serving_request

    budget = 64 << 30
    log("bench: generating the serving field pool on the device")
    pool = serving_pool(pool_fields, seed=5, device=dev)
    rng = np.random.default_rng(5)

    def submit(srv, r, arrival=None, keep=False):
        return srv.submit(r["texture"], r["u"], r["v"], kernel=r["kernel"],
                          uv_mode=r["uv_mode"], boundaries=r["boundaries"],
                          iterations=r["iterations"], keep=keep, arrival=arrival)

    # capacity: a burst of requests all arriving at once, timed on a fresh
    # server after the same requests ran once untimed on another (a
    # process's first run of a burst is slower: 510-730 against 755-796
    # Gpixel-steps/s measured for the same burst run again)
    warm_in = [serving_request(rng, pool) for _ in range(8)]
    reqs_in = [serving_request(rng, pool) for _ in range(burst)]
    with Server(budget, device=dev) as srv:
        for r in warm_in + reqs_in:
            submit(srv, r)
        srv.run()
    with Server(budget, device=dev) as srv:
        for r in warm_in:
            submit(srv, r)
        srv.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reqs = [submit(srv, r) for r in reqs_in]
        srv.run()
        torch.cuda.synchronize()
        cap_s = time.perf_counter() - t0
    cap_work = sum(q.pixel_steps for q in reqs)
    capacity = cap_work / cap_s
    mean_work = cap_work / len(reqs)
    lam = 0.7 * capacity / mean_work
    log(f"bench: capacity {capacity / 1e9:.1f} Gpixel-steps/s over {len(reqs)} requests; "
        f"Poisson rate {lam:.1f} requests/s (70% load)")

    # Poisson arrivals injected in real time by this thread while a
    # background thread steps the server; a quarter of the arrivals follow
    # up a finished kept request
    n = n_requests
    gaps = rng.exponential(1.0 / lam, size=n)
    offsets = np.cumsum(gaps)
    plans = [(serving_request(rng, pool), rng.random() < 0.25, rng.random() < 0.5,
              int(rng.integers(1, 21))) for _ in range(n)]
    submitted: list = []
    kept: list = []
    with Server(budget, device=dev) as srv:
        torch.cuda.synchronize()
        srv.start()
        t_start = time.perf_counter()
        for k in range(n):
            delay = offsets[k] - (time.perf_counter() - t_start)
            if delay > 0:
                time.sleep(delay)
            r, follow, keep, more = plans[k]
            arrival = t_start + offsets[k]
            ready = [q for q in kept if q.state == "done"]
            if follow and ready:
                base = ready[int(rng.integers(len(ready)))]
                kept.remove(base)
                q = srv.follow_up(base, more, arrival=arrival)
            else:
                q = submit(srv, r, arrival=arrival, keep=keep)
            submitted.append(q)
            if q.keep:
                kept.append(q)
        srv.run()
        t_end = time.perf_counter()
        srv.stop()
        steps_run, evictions, readmissions = srv.steps, srv.evictions, srv.readmissions
    done = [q for q in submitted if q.state == "done"]
    lat = np.array([q.latency for q in done]) * 1e3
    work = sum(q.pixel_steps for q in done)
    span = t_end - t_start
    del pool
    torch.cuda.empty_cache()
    return {
        "workload": "serving_poisson", "requests": n, "completed": len(done),
        "value": round(work / span / 1e9, 3), "unit": "Gpixel-steps/s",
        "span_s": round(span, 3), "server_steps": steps_run,
        "rate_per_s": round(lam, 2), "budget_bytes": budget,
        "capacity_gpixel_steps_per_s": round(capacity / 1e9, 3),
        "followups": sum(q.base is not None for q in done),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                       "p99": round(float(np.percentile(lat, 99)), 2),
                       "max": round(float(lat.max()), 2)},
        "evictions": evictions, "readmissions": readmissions,
    }


def bench_serve(args, dev) -> int:
    """Serving as the benchmark's one line."""
    m = serve_measure(dev, args.serve_burst, args.serve_requests)
    line = {
        "metric": "Gpixel-steps/s (serving, Poisson arrivals at 70% load)",
        "value": m["value"],
        "unit": "Gpixel-steps/s",
        "n_gpus": 1,
        "steps": m["server_steps"],
        "warmup": 8,
        "ms_per_step": round(m["span_s"] / max(m["server_steps"], 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": m["workload"], "requests": m["requests"],
                   "rate_per_s": m["rate_per_s"], "budget_bytes": m["budget_bytes"],
                   "capacity_gpixel_steps_per_s": m["capacity_gpixel_steps_per_s"],
                   "followups": m["followups"]},
        "latency_ms": m["latency_ms"],
        "evictions": m["evictions"],
        "readmissions": m["readmissions"],
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def single_call_measure(wl, dev, steps: int, warmup: int) -> dict:
    """One configuration as repeated convolve calls on device-resident
    inputs: ms per call and Gpixel-steps/s (synchronized start and end),
    and per-kernel times from the launch profiler in a separate, untimed
    call."""
    import torch

    from rlic_amd import _core, _native
    from rlic_amd._boundaries import BoundarySet

    def dev_tensor(a):
        if hasattr(a, "data_ptr"):
            return a
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    t = [dev_tensor(a) for a in (wl.texture, wl.u, wl.v, wl.kernel)]
    fn = _core.convolve_f32 if t[0].dtype == torch.float32 else _core.convolve_f64
    bs = BoundarySet.from_spec(wl.boundaries)

    def step():
        return fn(t[0], (t[1], t[2], wl.uv_mode), t[3], (bs.x, bs.y), wl.iterations)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    _native.profiler_enable(True)
    _native.profiler_read()
    step()
    torch.cuda.synchronize()
    prof = _native.profiler_read()
    _native.profiler_enable(False)
    ny, nx = wl.shape
    return {"workload": wl.name, "shape": [ny, nx], "kernel_len": wl.kernel_len,
            "iterations": wl.iterations, "uv_mode": wl.uv_mode,
            "ms_per_call": round(ms, 4),
            "value": round(wl.pixel_steps / (ms / 1e3) / 1e9, 3), "unit": "Gpixel-steps/s",
            "kernel_ms_per_call": {k: round(v["total_ms"], 4) for k, v in prof.items()},
            "steps": steps, "warmup": warmup,
            "inputs": wl.meta.get("generated_on", "host") + "-generated"}


def graph_call_measure(wl, dev, steps: int, warmup: int) -> dict:
    """One configuration as a whole allocation-free call
    (rlic_convolve_noalloc_*) captured in one HIP graph and replayed:
    microseconds per replay (synchronized start and end over `steps`
    replays), beside the same call launched directly; the replay's output
    is checked against the direct call's."""
    import ctypes

    import torch

    from rlic_amd import _native
    from rlic_amd._boundaries import BoundarySet

    lib = _native.load()
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev)
         for a in (wl.texture, wl.u, wl.v, wl.kernel)]
    f32 = t[0].dtype == torch.float32
    fn = lib.rlic_convolve_noalloc_f32 if f32 else lib.rlic_convolve_noalloc_f64
    bs = BoundarySet.from_spec(wl.boundaries)
    code = {"closed": 0, "periodic": 1}
    bounds = [code[bs.x[0]], code[bs.x[1]], code[bs.y[0]], code[bs.y[1]]]
    ny, nx = wl.shape
    out = torch.empty_like(t[0])
    direct = torch.empty_like(t[0])

    def call(dst, stream):
        st = fn(t[0].data_ptr(), nx, 1, t[1].data_ptr(), nx, 1, t[2].data_ptr(), nx, 1, ny, nx,
                t[3].data_ptr(), t[3].shape[0], {"velocity": 0, "polarization": 1}[wl.uv_mode],
                *bounds, wl.iterations, dst.data_ptr(), 3, ctypes.c_void_p(stream.cuda_stream))
        _native.raise_for_status(st)

    side = torch.cuda.Stream(dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(side):
        call(direct, side)
    side.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        call(out, torch.cuda.current_stream(dev))

    def timed(step):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps * 1e6

    stream = torch.cuda.current_stream(dev)
    us_direct = timed(lambda: call(direct, stream))
    us_graph = timed(graph.replay)
    same = bool(torch.equal(out, direct))
    return {"workload": wl.name, "iterations": wl.iterations,
            "us_per_call_direct": round(us_direct, 2), "us_per_replay": round(us_graph, 2),
            "value": round(wl.pixel_steps / (us_graph / 1e6) / 1e9, 3),
            "unit": "Gpixel-steps/s", "same_bits_as_direct": same,
            "steps": steps, "warmup": warmup}


def extras_measure(dev, args) -> dict:
    """The other configurations, each on its own inputs (freed after)."""
    import torch

    from rlic_amd.workloads import polarization_field_device, readme_example

    out = {}
    log("bench: extra readme_example (256^2 float64, L=65)")
    out["readme_256_f64"] = single_call_measure(readme_example(), dev, steps=50, warmup=5)
    log("bench: extra readme_example as one captured graph per call")
    out["readme_256_f64_graph"] = graph_call_measure(readme_example(), dev, steps=200, warmup=10)
    for key, n, dtype in (("polarization_16384_f32", 16384, np.float32),
                          ("polarization_8192_f64", 8192, np.float64)):
        log(f"bench: extra {key} (inputs generated on the device)")
        wl = polarization_field_device(n, dtype, device=dev)
        out[key] = single_call_measure(wl, dev, steps=3, warmup=1)
        del wl
        torch.cuda.empty_cache()
    log("bench: extra batch_uneven_512 (batched vs per-call)")
    out["batch_uneven_512"] = batch_measure(dev, steps=3, warmup=1)
    log("bench: extra serving")
    out["serving_poisson"] = serve_measure(dev, args.serve_burst, args.serve_requests)
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plain4096")
    ap.add_argument("--cpu-rows", type=int, default=2048,
                    help="rows of pass 1 timed on the CPU oracle (0 = skip)")
    ap.add_argument("--serve-burst", type=int, default=96,
                    help="serving: requests of the capacity burst")
    ap.add_argument("--serve-requests", type=int, default=800,
                    help="serving: Poisson-arrival requests")
    ap.add_argument("--no-extras", action="store_true",
                    help="time only the main workload (no other configurations)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="N",
                    help="first run the library's copy kernel N times on a 1 GiB array "
                         "(known HBM bytes, for calibrating FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the LIC path has "
              "no multi-GPU implementation in this build; see DESIGN.md for how it would "
              "shard.", file=sys.stderr)
        return 2

    import torch

    import rlic_amd
    from rlic_amd import _core, _native
    from rlic_amd.workloads import WORKLOADS

    assert torch.cuda.is_available(), "no ROCm device visible"
    dev = torch.device("cuda:0")
    if args.calibrate_copy > 0:
        # lic_copy2d reads and writes 2^28 float32 = 1 GiB each way per call
        src = torch.rand((16384, 16384), device=dev)
        dst = torch.empty_like(src)
        for _ in range(args.calibrate_copy):
            _core.convolve_f32(src, (src, src, "velocity"), torch.ones(1, device=dev),
                               (("closed",) * 2,) * 2, 0)
        torch.cuda.synchronize()
        log(f"bench: {args.calibrate_copy} calibration copies of 1 GiB done")
        del src, dst
        torch.cuda.empty_cache()
    if args.workload in ("batch512", "batch512small"):
        return bench_batch(args, dev)
    if args.workload == "serve":
        return bench_serve(args, dev)
    log(f"bench: building workload {args.workload}")
    wl = WORKLOADS[args.workload]()
    fn = _core.convolve_f32 if wl.texture.dtype == np.float32 else _core.convolve_f64
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in
         (wl.texture, wl.u, wl.v, wl.kernel)]
    bs = rlic_amd._boundaries.BoundarySet.from_spec(wl.boundaries)
    bpair = (bs.x, bs.y)

    def step():
        return fn(t[0], (t[1], t[2], wl.uv_mode), t[3], bpair, wl.iterations)

    log(f"bench: {args.warmup} warmup steps")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    _native.profiler_enable(True)
    _native.profiler_read()
    log(f"bench: {args.steps} timed steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = _native.profiler_read()
    _native.profiler_enable(False)
    log(f"bench: kernel times {json.dumps(prof)}")

    ms_per_step = elapsed / args.steps * 1e3
    value = wl.pixel_steps / (ms_per_step / 1e3) / 1e9

    dominant = max(prof, key=lambda k: prof[k]["total_ms"]) if prof else None
    roofline = None
    if dominant:
        avg_s = prof[dominant]["total_ms"] / prof[dominant]["count"] / 1e3
        b_alg = wl.algorithmic_bytes_per_pass()
        traffic = traffic_from_profiles(dominant, args.workload)
        hbm = None
        if traffic is not None:
            hbm_gbs = traffic / avg_s / 1e9
            hbm = {"achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(hbm_gbs / HBM_PEAK_GBS, 4)}
        valu = valu_from_profiles(dominant, args.workload)
        ny, nx = wl.shape
        lane_steps = ny * nx * (wl.kernel_len - 1)  # pixel-steps of one launch (one pass)
        if valu is not None:
            lane_instr = valu["insts"] * 64
            achieved = lane_instr / avg_s / 1e12
            roofline = {
                "bound": "valu",
                "kernel": dominant,
                "achieved": round(achieved, 3),
                "peak": VALU_PEAK_TLANE,
                "unit": "T VALU lane-instructions/s",
                "frac": round(achieved / VALU_PEAK_TLANE, 4),
                "traffic": traffic,
                "valu_lane_instructions_per_launch": lane_instr,
                "valu_per_lane_step": round(lane_instr / lane_steps, 2),
                "hbm": hbm,
            }
        else:  # no counter summary for this kernel: the HBM view only
            hbm_gbs = (traffic or 0.0) / avg_s / 1e9
            roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(hbm_gbs, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(hbm_gbs / HBM_PEAK_GBS, 4), "traffic": traffic}
        roofline.update({
            "logical_gather_gbs": round(b_alg / avg_s / 1e9, 1),
            "algorithmic_bytes_per_launch": b_alg,
            "avg_launch_ms": round(avg_s * 1e3, 4),
            "kernel_ms_per_step": {k: round(v["total_ms"] / args.steps, 4)
                                   for k, v in prof.items()},
        })

    if roofline is not None:
        # untimed: what a plain copy reaches on this HBM, beside the nominal peak
        roofline["measured_copy_gbs"] = round(measure_copy_gbs(dev), 1)
        log(f"bench: plain copy {roofline['measured_copy_gbs']} GB/s")

    # divergence (untimed): lane work of one tracing pass, read back from its
    # move record (RLIC_TRACE_STATS=1)
    os.environ["RLIC_TRACE_STATS"] = "1"
    _native.trace_stats_read()
    step()
    torch.cuda.synchronize()
    divergence = _native.trace_stats_read()
    del os.environ["RLIC_TRACE_STATS"]
    log(f"bench: trace lane efficiency {divergence['lane_efficiency']}")

    cpu = None
    if args.cpu_rows > 0:
        log(f"bench: CPU oracle baseline on {args.cpu_rows} rows")
        cpu = cpu_baseline(wl, args.cpu_rows)

    extras = None
    if not args.no_extras and args.workload == "plain4096":
        del t
        torch.cuda.empty_cache()
        extras = extras_measure(dev, args)

    ny, nx = wl.shape
    line = {
        "metric": "Gpixel-steps/s",
        "value": round(value, 3),
        "unit": "Gpixel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if wl.texture.dtype == np.float32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": wl.name,
            "shape": [ny, nx],
            "kernel_len": wl.kernel_len,
            "iterations": wl.iterations,
            "boundaries": wl.boundaries,
            "uv_mode": wl.uv_mode,
            "pixel_steps_per_step": wl.pixel_steps,
        },
        "roofline": roofline,
        "divergence": divergence,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
