"""Parity at the benchmark sizes (GPU).

The oracle cannot run these sizes in seconds, so each case is checked
through properties that do not depend on size:
  * a band of rows of pass 1 computed by the oracle matches bit for bit
    (the oracle computes only those rows);
  * n passes in one call (trace once, replay the rest) equal n chained
    single passes (each one traced), bit for bit;
  * transposing every input (and swapping u, v) transposes the output.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import cpu
from rlic_amd import _core
from rlic_amd.workloads import plain_at_scale, polarization_field

pytestmark = pytest.mark.gpu


def _device(wl, device):
    import torch

    return [torch.from_numpy(np.ascontiguousarray(a)).to(device)
            for a in (wl.texture, wl.u, wl.v, wl.kernel)]


def _bpair(b):
    if isinstance(b, str):
        return ((b, b), (b, b))
    return ((b["x"],) * 2, (b["y"],) * 2)


def _run(wl, t, iterations):
    fn = _core.convolve_f32 if wl.texture.dtype == np.float32 else _core.convolve_f64
    return fn(t[0], (t[1], t[2], wl.uv_mode), t[3], _bpair(wl.boundaries), iterations)


def _band_check(wl, got, rows):
    r0, r1 = rows
    exp = cpu.convolve(wl.texture, wl.u, wl.v, kernel=wl.kernel, uv_mode=wl.uv_mode,
                       boundaries=wl.boundaries, iterations=1, rows=rows, threads=8)
    band = got[r0:r1]
    assert np.array_equal(band, exp[r0:r1]), (
        f"{np.count_nonzero(band != exp[r0:r1])} pixels of rows {rows} differ from the oracle")


def test_plain_4096_f32_five_passes(gpu_device):
    import torch

    wl = plain_at_scale(4096, np.float32, iterations=5)
    t = _device(wl, gpu_device)
    one = _run(wl, t, 1)
    torch.cuda.synchronize()
    host_one = one.cpu().numpy()
    for rows in ((0, 8), (2044, 2052), (4088, 4096)):  # both closed edges and the middle
        _band_check(wl, host_one, rows)
    chained = one
    for _ in range(4):
        chained = _run(wl, [chained, t[1], t[2], t[3]], 1)
    five = _run(wl, t, 5)
    torch.cuda.synchronize()
    assert torch.equal(five, chained), "5 passes in one call differ from 5 chained passes"


def test_plain_4096_transpose_symmetry(gpu_device):
    """Transposing every input (u and v swapped, x and y boundaries swapped)
    transposes the output, except where a step hits an exact tie tx == ty:
    the reference then always crosses y (src/lib.rs:249-269), which breaks the
    symmetry.  Every asymmetric pixel must therefore match the oracle in both
    orientations."""
    import torch

    wl = plain_at_scale(4096, np.float32, iterations=1)
    t = _device(wl, gpu_device)
    out = _run(wl, t, 1)
    out_t = _core.convolve_f32(t[0].T, (t[2].T, t[1].T, wl.uv_mode), t[3],
                               (("closed", "closed"), ("periodic", "periodic")), 1)
    torch.cuda.synchronize()
    a = out.cpu().numpy()
    b = out_t.T.cpu().numpy()
    diff = np.argwhere(a != b)
    assert len(diff) < a.size * 1e-4, f"{len(diff)} asymmetric pixels"
    tex_t = np.ascontiguousarray(wl.texture.T)
    u_t = np.ascontiguousarray(wl.v.T)
    v_t = np.ascontiguousarray(wl.u.T)
    for i, j in diff[:3]:
        rows = (int(i), int(i) + 1)
        exp = cpu.convolve(wl.texture, wl.u, wl.v, kernel=wl.kernel, uv_mode=wl.uv_mode,
                           boundaries=wl.boundaries, rows=rows, threads=8)
        assert a[i, j] == exp[i, j]
        rows_t = (int(j), int(j) + 1)
        exp_t = cpu.convolve(tex_t, u_t, v_t, kernel=wl.kernel, uv_mode=wl.uv_mode,
                             boundaries={"x": "closed", "y": "periodic"}, rows=rows_t, threads=8)
        assert out_t[j, i].item() == exp_t[j, i]


@pytest.mark.parametrize("n, dtype", [(16384, np.float32), (8192, np.float64)])
def test_polarization_field_band(n, dtype, gpu_device):
    import torch

    wl = polarization_field(n, dtype)
    t = _device(wl, gpu_device)
    got = _run(wl, t, 1)
    torch.cuda.synchronize()
    host = got.cpu().numpy()
    del got, t
    torch.cuda.empty_cache()
    for rows in ((0, 2), (n // 2 - 1, n // 2 + 1), (n - 2, n)):
        _band_check(wl, host, rows)


@pytest.mark.parametrize("shape,dtype", [((1500, 1700), np.float64), ((2049, 2047), np.float32),
                                         ((3, 5), np.float64)])
def test_host_arrays_match_device_arrays(shape, dtype, gpu_device):
    """Host (NumPy) arrays, uploaded and downloaded by the call itself (up
    to 20 MB per array; a transposed texture view, whose memory extent is
    uploaded): the result equals the same call on device tensors bit for
    bit, twice in a row, and two rows of it equal the oracle's."""
    import torch

    rng = np.random.default_rng(shape[0])
    ny, nx = shape
    tex_t = rng.random((nx, ny)).astype(dtype)
    tex = tex_t.T  # a strided view
    u = (rng.random(shape) - 0.5).astype(dtype)
    v = (rng.random(shape) - 0.5).astype(dtype)
    k = rng.random(31).astype(dtype)
    fn = _core.convolve_f32 if dtype == np.float32 else _core.convolve_f64
    bounds = (("periodic", "periodic"), ("closed", "closed"))
    dev = [torch.from_numpy(np.ascontiguousarray(a)).to(gpu_device) for a in (tex, u, v, k)]
    want = fn(dev[0], (dev[1], dev[2], "velocity"), dev[3], bounds, 2).cpu().numpy()
    for rep in range(2):
        got = fn(tex, (u, v, "velocity"), k, bounds, 2)
        assert isinstance(got, np.ndarray) and got.shape == shape
        assert np.array_equal(got, want), f"{shape} rep {rep}"
    rows = (ny // 2, min(ny, ny // 2 + 2))
    exp = cpu.convolve(np.ascontiguousarray(tex), u, v, kernel=k, uv_mode="velocity",
                       boundaries={"x": "periodic", "y": "closed"}, iterations=1, rows=rows,
                       threads=8)
    one = fn(tex, (u, v, "velocity"), k, bounds, 1)
    assert np.array_equal(one[rows[0]:rows[1]], exp[rows[0]:rows[1]])
