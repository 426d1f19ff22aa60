"""Host-side sanitizer runs of ``librlic_hip.so`` (no GPU needed).

    python -m rlic_amd._sanitize [address|thread|all]

Builds the library with the host code instrumented (every ``-fsanitize=``
is placed after ``-Xarch_host``: the device code is built as usual) into
``_build/`` and runs, against it, the host code paths that need no GPU:

* ``address``: AddressSanitizer + UndefinedBehaviorSanitizer (UB aborts)
  over the CPU test files that call the C ABI: argument validation and
  size limits (``tests/test_abi.py``), batch field checks
  (``tests/test_batch.py``), serving budget arithmetic and argument errors
  (``tests/test_serve.py``), the autotuner table file parser
  (``tests/test_autotune.py``), equalization arguments
  (``tests/test_equalize.py``) and the allocation-free entry point's limits
  (``tests/test_zz_no_alloc.py``);
* ``thread``: ThreadSanitizer over a stress of the host-side shared state
  from 8 threads at once (ctypes releases the GIL): the autotuner table
  (load, read, clear), the launch-profiler record, the thread-local error
  messages of every entry point's argument checks and the serving budget
  arithmetic.

The sanitizer runtimes come from ROCm's clang and are preloaded into the
Python process (the library itself does not link them).  Exit status 0
means no report.
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys
import tempfile
import textwrap
from pathlib import Path

from rlic_amd import _build

ROOT = _build.ROOT
OUT_DIR = ROOT / "_build"

TEST_FILES = ["tests/test_abi.py", "tests/test_batch.py", "tests/test_serve.py",
              "tests/test_autotune.py", "tests/test_equalize.py", "tests/test_zz_no_alloc.py"]


def runtime(name: str) -> str:
    pattern = f"/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.{name}-x86_64.so"
    found = sorted(glob.glob(pattern))
    if not found:
        raise SystemExit(f"no {name} runtime matches {pattern}")
    return found[-1]


def build(kind: str) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    out = OUT_DIR / f"librlic_hip_{kind}.so"
    flags = {"address": ["-fsanitize=address", "-fsanitize=undefined",
                         "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"],
             "thread": ["-fsanitize=thread"]}[kind]
    cmd = _build.command(out)
    host = []
    for f in flags:
        host += ["-Xarch_host", f]
    cmd[1:1] = host + ["-g"]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


STRESS = textwrap.dedent('''
    import ctypes, os, sys, tempfile, threading
    sys.path.insert(0, sys.argv[1])
    from rlic_amd import _native
    lib = _native.load()
    table = os.path.join(tempfile.mkdtemp(), "plans.txt")
    with open(table, "w") as f:
        for k in range(64):
            f.write(f"rlic-replay-plan 4 {512 + k} 512 129 3 1 64 64 0 0 0.{k + 1}\\n")
    a = (ctypes.c_double * 16)(*([1.0] * 16))
    p = ctypes.cast(a, ctypes.c_void_p).value
    errors = []

    def work(t):
        buf = ctypes.create_string_buffer(1 << 16)
        for it in range(200):
            op = (t + it) % 6
            if op == 0:
                assert lib.rlic_autotune_load(table.encode()) == 64
            elif op == 1:
                lib.rlic_autotune_read(buf, len(buf))
            elif op == 2:
                lib.rlic_autotune_clear()
            elif op == 3:
                lib.rlic_profiler_enable(it % 2)
                lib.rlic_profiler_read(buf, len(buf))
            elif op == 4:
                st = lib.rlic_convolve_f64(p, 4, 1, p, 4, 1, p, 4, 1, 4, 4, p, 3, 9, 0, 0, 0, 0,
                                           1, p, 0, 3, None)
                msg = _native.last_error()
                if st >= 0 or "unknown uv_mode" not in msg:
                    errors.append((t, it, st, msg))
                st = lib.rlic_equalize_histogram_f64(p, 4, 1, 4, 4, 0, p, 0, None)
                if st >= 0 or "nbins" not in _native.last_error():
                    errors.append((t, it, "eq"))
            else:
                assert lib.rlic_server_field_bytes_f32(64 + t, 64, 33, it % 2, 2) > 0
                assert lib.rlic_server_field_bytes_f32(1, 1431655764, 3, 0, 1) == -1

    threads = [threading.Thread(target=work, args=(t,)) for t in range(8)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[:3]
    print("stress: 8 threads x 200 calls, no error")
''')


def run(kind: str) -> int:
    lib = build(kind)
    env = dict(os.environ)
    env["RLIC_AMD_LIB"] = str(lib)
    env["RLIC_AMD_SYSTEM_HIP"] = "1"
    env["PYTHONPATH"] = str(ROOT) + os.pathsep + env.get("PYTHONPATH", "")
    if kind == "address":
        env["LD_PRELOAD"] = runtime("asan")
        env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
        env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
        cmd = [sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
               *TEST_FILES]
    else:
        env["LD_PRELOAD"] = runtime("tsan")
        env["TSAN_OPTIONS"] = "halt_on_error=1:report_signal_unsafe=0"
        with tempfile.NamedTemporaryFile("w", suffix=".py", delete=False) as f:
            f.write(STRESS)
        cmd = [sys.executable, f.name, str(ROOT)]
    print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, env=env, cwd=ROOT).returncode


def main(argv: list[str]) -> int:
    which = argv[0] if argv else "all"
    kinds = ["address", "thread"] if which == "all" else [which]
    for kind in kinds:
        rc = run(kind)
        print(f"sanitize {kind}: exit status {rc}", flush=True)
        if rc != 0:
            return rc
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
