"""Golden fixtures (tests/golden/, written by make_golden.py).

CPU: the oracle reproduces every fixture bit for bit (a regression pin on
the oracle), and the regression-suite outputs satisfy the closed forms the
unit half-plane fields admit.  GPU: the product path reproduces every
fixture bit for bit.
"""

from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import rlic_amd
from oracle import cpu
from rlic_amd import _core

GOLDEN = Path(__file__).resolve().parent / "golden"
BOUND_SETS = {
    "closed": "closed",
    "periodic": "periodic",
    "xper_yclosed": {"x": "periodic", "y": "closed"},
}


def _combos():
    data = np.load(GOLDEN / "combos.npz")
    keys = sorted({k.split("__")[0] for k in data.files})
    for key in keys:
        dtype, prof, mode, bname = key.split("_", 3)
        yield key, {
            "texture": data[f"{key}__texture"], "u": data[f"{key}__u"], "v": data[f"{key}__v"],
            "kernel": data[f"{key}__kernel"], "out": data[f"{key}__out"],
            "profile": int(prof[1:]), "uv_mode": mode, "boundaries": BOUND_SETS[bname],
        }


def _bpair(b):
    if isinstance(b, str):
        return ((b, b), (b, b))
    return ((b["x"],) * 2, (b["y"],) * 2)


def _same(a, b):
    return a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_fixture_inventory():
    assert len(list(_combos())) == 48
    assert {"texture", "u", "v"} <= set(np.load(GOLDEN / "kernels.npz").files)
    assert json.loads((GOLDEN / "readme.json").read_text())["sha256"]


def test_oracle_reproduces_combos():
    for key, c in _combos():
        out = cpu.convolve(c["texture"], c["u"], c["v"], kernel=c["kernel"], uv_mode=c["uv_mode"],
                           boundaries=c["boundaries"], iterations=2, profile=c["profile"])
        assert _same(out, c["out"]), key


def test_oracle_reproduces_kernel_lengths():
    d = np.load(GOLDEN / "kernels.npz")
    for klen in (1, 2, 3, 4, 5, 11, 64, 65):
        out = cpu.convolve(d["texture"], d["u"], d["v"], kernel=d[f"L{klen}__kernel"],
                           uv_mode="polarization", boundaries="periodic", iterations=2)
        assert _same(out, d[f"L{klen}__out"]), klen


def test_oracle_reproduces_readme_digest():
    ref = json.loads((GOLDEN / "readme.json").read_text())
    shape = (256, 256)
    texture = np.random.default_rng(0).random(shape)
    x = np.linspace(0, np.pi, 256)
    U = np.ascontiguousarray(np.broadcast_to(np.cos(2 * x), shape))
    V = np.ascontiguousarray(np.broadcast_to(np.sin(x).T, shape))
    kernel = 1 - np.abs(np.linspace(-1, 1, 65))
    out = cpu.convolve(texture, U, V, kernel=kernel, boundaries="periodic")
    assert hashlib.sha256(out.tobytes()).hexdigest() == ref["sha256"]
    for i, j, h in ref["samples"]:
        assert out[i, j] == float.fromhex(h)


def test_regression_cases_closed_forms():
    """The reference's regression inputs: with u = v = 0 every tap samples
    the start cell, so out is the float32 fma chain of the kernel in the
    reference's order times the texture value (checked exactly on a few
    pixels); with v = 0 and u = +-1, rows are independent."""
    from oracle import pyref

    d = np.load(GOLDEN / "regression.npz")
    tex, k0 = d["texture"], d["kernel"]
    kmid = k0.size // 2
    order = [kmid] + list(range(kmid + 1, k0.size)) + list(range(kmid - 1, -1, -1))
    f32 = pyref.BINARY32
    for mode in ("velocity", "polarization"):
        zero = d[f"zero_zero_{mode}__out"]
        for i, j in ((0, 0), (7, 19), (30, 32)):
            acc = 0.0
            for kk in order:  # src/lib.rs:375-403
                acc = f32.fma(float(k0[kk]), float(tex[i, j]), acc)
            assert zero[i, j] == np.float32(acc)
        out_u = d[f"u1_zero_{mode}__out"]
        tex2 = tex.copy()
        tex2[5] += 1
        out2 = cpu.convolve(tex2, d["u1"], np.zeros_like(tex), kernel=k0, uv_mode=mode)
        changed = np.any(out2 != out_u, axis=1)
        assert changed[5] and changed.sum() == 1


@pytest.mark.gpu
def test_product_reproduces_combos():
    for key, c in _combos():
        native = _core.convolve_f32 if c["texture"].dtype == np.float32 else _core.convolve_f64
        out = native(c["texture"], (c["u"], c["v"], c["uv_mode"]), c["kernel"],
                     _bpair(c["boundaries"]), 2, profile=c["profile"])
        assert _same(out, c["out"]), key


@pytest.mark.gpu
def test_product_reproduces_kernel_lengths_and_regression():
    d = np.load(GOLDEN / "kernels.npz")
    for klen in (1, 2, 3, 4, 5, 11, 64, 65):
        out = rlic_amd.convolve(d["texture"], d["u"], d["v"], kernel=d[f"L{klen}__kernel"],
                                uv_mode="polarization", boundaries="periodic", iterations=2)
        assert _same(out, d[f"L{klen}__out"]), klen
    r = np.load(GOLDEN / "regression.npz")
    zero = np.zeros_like(r["texture"])
    for uname in ("zero", "u1"):
        for vname in ("zero", "v1"):
            for mode in ("velocity", "polarization"):
                uu = zero if uname == "zero" else r["u1"]
                vv = zero if vname == "zero" else r["v1"]
                out = rlic_amd.convolve(r["texture"], uu, vv, kernel=r["kernel"], uv_mode=mode)
                assert _same(out, r[f"{uname}_{vname}_{mode}__out"]), (uname, vname, mode)


@pytest.mark.gpu
def test_product_reproduces_readme_digest():
    ref = json.loads((GOLDEN / "readme.json").read_text())
    shape = (256, 256)
    texture = np.random.default_rng(0).random(shape)
    x = np.linspace(0, np.pi, 256)
    U = np.broadcast_to(np.cos(2 * x), shape)
    V = np.broadcast_to(np.sin(x).T, shape)
    kernel = 1 - np.abs(np.linspace(-1, 1, 65))
    out = rlic_amd.convolve(texture, U, V, kernel=kernel, boundaries="periodic")
    assert hashlib.sha256(out.tobytes()).hexdigest() == ref["sha256"]
