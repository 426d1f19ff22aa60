"""Generate the golden fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

The reference implementation (Rust, ``src/lib.rs``) cannot be built in this
environment (no Rust toolchain) and its regression oracle ``vectorplot`` is
not installed, so these vectors are "restatement-derived": outputs of the
C++ oracle, each tiny case cross-checked bit for bit against the independent
exact-arithmetic restatement (oracle/pyref.py) before it is written.

Fixtures:
  combos.npz      every dtype x arithmetic profile x uv_mode x boundary set
                  on a 9 x 11 field with NaN holes and signed zeros, L=7,
                  2 iterations (48 cases)
  kernels.npz     kernel lengths 1, 2, 3, 4, 5, 11, 64, 65 on a 16 x 13
                  float64 field, polarization, periodic, 2 iterations
  regression.npz  the 8 input cases of the reference's
                  tests/test_regressions.py (31 x 33 float32 unit half-plane
                  fields, K0 = sin(pi * arange(31) / 31)), default settings
  readme.json     SHA-256 of the float64 output of the README example
                  (256 x 256, L=65, periodic) and 16 sampled pixel values
"""

from __future__ import annotations

import hashlib
import itertools
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

from oracle import cpu, pyref  # noqa: E402

BOUND_SETS = {
    "closed": "closed",
    "periodic": "periodic",
    "xper_yclosed": {"x": "periodic", "y": "closed"},
}


def _field(rng, shape, dtype):
    tex = rng.random(shape).astype(dtype)
    u = (rng.random(shape) - 0.5).astype(dtype)
    v = (rng.random(shape) - 0.5).astype(dtype)
    u[1, 2] = np.nan
    v[4, 4] = np.nan
    u[3, :] = 0.0
    v[3, ::2] = -0.0
    v[:, 5] = 0.0
    return tex, u, v


def _check(out, tex, u, v, **kw):
    ref = pyref.convolve(tex, u, v, **kw)
    if not np.array_equal(out.view(np.uint8), ref.view(np.uint8)):
        raise SystemExit(f"oracle and exact restatement disagree on {kw}")


def combos() -> dict:
    rng = np.random.default_rng(2024)
    arrays = {}
    for dtype, profile, mode, (bname, b) in itertools.product(
        (np.float32, np.float64), range(4), ("velocity", "polarization"), BOUND_SETS.items()
    ):
        key = f"{np.dtype(dtype).name}_p{profile}_{mode}_{bname}"
        tex, u, v = _field(rng, (9, 11), dtype)
        k = (rng.random(7) - 0.1).astype(dtype)
        kw = dict(kernel=k, uv_mode=mode, boundaries=b, iterations=2, profile=profile)
        out = cpu.convolve(tex, u, v, **kw)
        _check(out, tex, u, v, **kw)
        for name, a in (("texture", tex), ("u", u), ("v", v), ("kernel", k), ("out", out)):
            arrays[f"{key}__{name}"] = a
    return arrays


def kernels() -> dict:
    rng = np.random.default_rng(7)
    arrays = {}
    tex, u, v = _field(rng, (16, 13), np.float64)
    arrays["texture"], arrays["u"], arrays["v"] = tex, u, v
    for klen in (1, 2, 3, 4, 5, 11, 64, 65):
        k = rng.random(klen)
        kw = dict(kernel=k, uv_mode="polarization", boundaries="periodic", iterations=2)
        out = cpu.convolve(tex, u, v, **kw)
        if klen <= 11:
            _check(out, tex, u, v, **kw)
        arrays[f"L{klen}__kernel"] = k
        arrays[f"L{klen}__out"] = out
    return arrays


def regression() -> dict:
    """Inputs of the reference's tests/test_regressions.py:19-58."""
    shape = (31, 33)
    rng = np.random.default_rng(0)
    texture = rng.random(shape, dtype="float32")
    one = np.ones_like(texture)
    zero = np.zeros_like(texture)
    cols = np.broadcast_to(np.arange(shape[1]), shape)
    u1 = np.where(cols < shape[1] / 2, -one, one)
    rows = np.broadcast_to(np.atleast_2d(np.arange(shape[0])).T, shape)
    v1 = np.where(rows < shape[0] / 2, -one, one)
    k0 = np.sin(np.arange(min(shape)) * np.pi / min(shape), dtype="float32")
    arrays = {"texture": texture, "kernel": k0, "u1": u1, "v1": v1}
    for (uname, uu), (vname, vv), mode in itertools.product(
        (("zero", zero), ("u1", u1)), (("zero", zero), ("v1", v1)), ("velocity", "polarization")
    ):
        out = cpu.convolve(texture, uu, vv, kernel=k0, uv_mode=mode)
        arrays[f"{uname}_{vname}_{mode}__out"] = out
    return arrays


def readme() -> dict:
    shape = (256, 256)
    texture = np.random.default_rng(0).random(shape)
    x = np.linspace(0, np.pi, 256)
    U = np.ascontiguousarray(np.broadcast_to(np.cos(2 * x), shape))
    V = np.ascontiguousarray(np.broadcast_to(np.sin(x).T, shape))
    kernel = 1 - np.abs(np.linspace(-1, 1, 65))
    out = cpu.convolve(texture, U, V, kernel=kernel, boundaries="periodic")
    idx = np.random.default_rng(1).integers(0, 256, size=(16, 2))
    return {
        "sha256": hashlib.sha256(np.ascontiguousarray(out).tobytes()).hexdigest(),
        "samples": [[int(i), int(j), float(out[i, j]).hex()] for i, j in idx],
    }


def main() -> None:
    cpu.build()
    np.savez_compressed(HERE / "combos.npz", **combos())
    np.savez_compressed(HERE / "kernels.npz", **kernels())
    np.savez_compressed(HERE / "regression.npz", **regression())
    (HERE / "readme.json").write_text(json.dumps(readme(), indent=1) + "\n")
    print("wrote", sorted(p.name for p in HERE.iterdir() if p.suffix in (".npz", ".json")))


if __name__ == "__main__":
    main()
