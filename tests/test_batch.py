"""Batched convolution: one persistent launch for many uneven fields gives,
field by field, the same bits as the single calls (GPU), and rejects bad
fields before any work (CPU)."""

from __future__ import annotations

import numpy as np
import pytest

import rlic_amd
from oracle import cpu
from oracle_check import assert_matches_oracle, oracle_convolve, same_bits
from rlic_amd import _native
from rlic_amd.batch import Field, convolve_batch


def _mixed_fields(rng, n):
    fields = []
    for k in range(n):
        dtype = np.float32 if k % 3 else np.float64
        ny, nx = int(rng.integers(1, 90)), int(rng.integers(1, 140))
        klen = int(rng.integers(1, 70))
        tex = rng.random((ny, nx)).astype(dtype)
        u = (rng.random((ny, nx)) - 0.5).astype(dtype)
        v = (rng.random((ny, nx)) - 0.5).astype(dtype)
        if k % 4 == 1:
            u[rng.random((ny, nx)) < 0.03] = np.nan
            v[:, nx // 2] = 0.0
        if k % 5 == 2:
            u = np.asfortranarray(u)  # strided input
        fields.append(Field(
            tex, u, v, rng.random(klen).astype(dtype),
            uv_mode="polarization" if k % 2 else "velocity",
            boundaries={"x": ("periodic", "closed")[k % 2], "y": ("closed", "periodic")[(k // 2) % 2]},
            iterations=int(rng.integers(0, 5)),
        ))
    return fields


@pytest.mark.gpu
def test_batch_matches_single_calls():
    rng = np.random.default_rng(42)
    fields = _mixed_fields(rng, 40)
    outs = convolve_batch(fields)
    for k, (f, out) in enumerate(zip(fields, outs)):
        single = rlic_amd.convolve(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                                   boundaries=f.boundaries, iterations=f.iterations)
        assert out.dtype == single.dtype and out.shape == single.shape, k
        assert np.array_equal(out, single, equal_nan=True), f"field {k} differs"


@pytest.mark.gpu
@pytest.mark.parametrize("streams", ["1", "3", "8"])
@pytest.mark.parametrize("single_min", ["1", "3000"])
def test_batch_single_field_launches_match_single_calls(monkeypatch, single_min, streams):
    """Fields at or above RLIC_BATCH_SINGLE_MIN pixels run as single-field
    launches, spread over RLIC_BATCH_STREAMS side streams, next to the
    persistent launch of the rest: every field, and the mix, gives the
    single calls' bits."""
    monkeypatch.setenv("RLIC_BATCH_SINGLE_MIN", single_min)
    monkeypatch.setenv("RLIC_BATCH_STREAMS", streams)
    rng = np.random.default_rng(5)
    fields = _mixed_fields(rng, 24)
    sizes = [f.texture.size for f in fields]
    assert any(s >= int(single_min) for s in sizes)
    if single_min != "1":
        assert any(s < int(single_min) for s in sizes)
    outs = convolve_batch(fields)
    for k, (f, out) in enumerate(zip(fields, outs)):
        single = rlic_amd.convolve(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                                   boundaries=f.boundaries, iterations=f.iterations)
        assert np.array_equal(out, single, equal_nan=True), f"field {k} differs"


@pytest.mark.gpu
def test_batch_matches_oracle_on_device_tensors(gpu_device):
    import torch

    rng = np.random.default_rng(7)
    fields = _mixed_fields(rng, 12)
    dev_fields = [Field(*(torch.from_numpy(np.ascontiguousarray(a)).to(gpu_device)
                          for a in (f.texture, f.u, f.v, f.kernel)),
                        uv_mode=f.uv_mode, boundaries=f.boundaries, iterations=f.iterations)
                  for f in fields]
    outs = convolve_batch(dev_fields)
    torch.cuda.synchronize()
    for k, (f, out) in enumerate(zip(fields, outs)):
        if f.iterations == 0:
            exp = f.texture
        else:
            exp = cpu.convolve(np.ascontiguousarray(f.texture), np.ascontiguousarray(f.u),
                               np.ascontiguousarray(f.v), kernel=f.kernel, uv_mode=f.uv_mode,
                               boundaries=f.boundaries, iterations=f.iterations)
        assert np.array_equal(out.cpu().numpy(), exp, equal_nan=True), f"field {k}"


@pytest.mark.gpu
def test_batch_uneven_workload_sample():
    """A few fields of the uneven-batch distribution, checked against single calls."""
    import torch

    from rlic_amd.workloads import batch_uneven

    dev = torch.device("cuda:0")
    fields = batch_uneven(16, seed=11, device=dev, max_side=512)
    outs = convolve_batch(fields)
    for k, (f, out) in enumerate(zip(fields, outs)):
        single = rlic_amd.convolve(f["texture"], f["u"], f["v"], kernel=f["kernel"],
                                   uv_mode=f["uv_mode"], boundaries=f["boundaries"],
                                   iterations=f["iterations"])
        assert torch.equal(out, single), f"field {k}"


@pytest.mark.gpu
def test_batch_with_empty_fields(gpu_device):
    """Empty images ((0, N) and (N, 0)) at 0-2 iterations, as NumPy arrays and
    as zero-size device tensors (null data pointers), mixed with normal
    fields: empty outputs, the normal fields' bits unchanged (the reference
    loops zero times over an empty image, src/lib.rs:373-374)."""
    import torch

    rng = np.random.default_rng(21)
    normal = _mixed_fields(rng, 6)
    empties = []
    for k, shape in enumerate([(0, 5), (5, 0), (0, 0), (0, 37), (23, 0), (0, 1)]):
        empties.append(Field(np.ones(shape), np.ones(shape), np.ones(shape), rng.random(7),
                             uv_mode=("velocity", "polarization")[k % 2],
                             boundaries=("closed", "periodic")[k % 3 == 0], iterations=k % 3))
    fields = [f for pair in zip(normal, empties) for f in pair]
    outs = convolve_batch(fields)
    for k, (f, out) in enumerate(zip(fields, outs)):
        assert out.shape == f.texture.shape, k
        single = rlic_amd.convolve(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                                   boundaries=f.boundaries, iterations=f.iterations)
        assert np.array_equal(out, single, equal_nan=True), k

    def t(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(gpu_device)

    dev_fields = [Field(t(f.texture), t(f.u), t(f.v), t(f.kernel), uv_mode=f.uv_mode,
                        boundaries=f.boundaries, iterations=f.iterations) for f in fields]
    assert dev_fields[1].texture.numel() == 0
    outs = convolve_batch(dev_fields)
    torch.cuda.synchronize()
    for k, (f, out) in enumerate(zip(fields, outs)):
        single = rlic_amd.convolve(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                                   boundaries=f.boundaries, iterations=f.iterations)
        assert np.array_equal(out.cpu().numpy(), single, equal_nan=True), k
    # a batch of empty fields only
    only = convolve_batch([dev_fields[1], dev_fields[3]])
    assert [tuple(o.shape) for o in only] == [(0, 5), (5, 0)]


@pytest.mark.gpu
def test_uneven_batch_at_full_size(gpu_device):
    """The batched uneven workload at full size: 512 float32 fields, sides
    log-uniform in 128..2048 (aspect 1:4..4:1), odd kernel lengths 17..257
    (triangular or boxcar), random periodic/closed axes and uv_mode, 1..10
    iterations.  Every field equals its single call bit for bit; fields of
    at most 256^2 pixels equal the oracle whole, one-pass larger fields on a
    band of rows (the reference semantics, src/lib.rs:408-443)."""
    import torch

    from rlic_amd import _core
    from rlic_amd.workloads import batch_uneven

    fields = batch_uneven(512, seed=4, device=gpu_device)
    sides = [s for f in fields for s in f["texture"].shape]
    assert min(sides) >= 128 and max(sides) <= 2048
    outs = convolve_batch(fields)
    torch.cuda.synchronize()
    checked = {"whole": 0, "band": 0}
    for k, (f, out) in enumerate(zip(fields, outs)):
        bs = rlic_amd._boundaries.BoundarySet.from_spec(f["boundaries"])
        single = _core.convolve_f32(f["texture"], (f["u"], f["v"], f["uv_mode"]), f["kernel"],
                                    (bs.x, bs.y), f["iterations"])
        assert same_bits(out, single), f"field {k} differs from its single call"
        small = f["texture"].numel() <= 256 * 256
        if (small and checked["whole"] < 16) or (f["iterations"] == 1 and checked["band"] < 24):
            how = assert_matches_oracle(out, f["texture"], f["u"], f["v"], kernel=f["kernel"],
                                        uv_mode=f["uv_mode"], boundaries=f["boundaries"],
                                        iterations=f["iterations"], what=f"field {k}")
            checked[how] += 1
    assert checked["whole"] >= 8 and checked["band"] >= 8, checked


def test_batch_validation_runs_first():
    good = Field(np.ones((4, 4)), np.ones((4, 4)), np.ones((4, 4)), np.ones(3))
    bad = Field(np.ones((4, 4)), np.ones((4, 5)), np.ones((4, 4)), np.ones(3))
    with pytest.raises(ValueError, match="^Shape mismatch"):
        convolve_batch([good, bad])
    assert convolve_batch([]) == []


def test_batch_abi_rejects_bad_fields_without_a_gpu():
    lib = _native.load()
    arr = (_native.FieldF64 * 2)()
    a = np.ones((4, 4))
    import ctypes

    p = ctypes.cast(a.ctypes.data, ctypes.POINTER(ctypes.c_double))
    for slot in arr:
        slot.texture = slot.u = slot.v = slot.out = slot.kernel = p
        slot.ny = slot.nx = 4
        slot.tex_sy = slot.u_sy = slot.v_sy = 4
        slot.tex_sx = slot.u_sx = slot.v_sx = 1
        slot.kernel_len = 3
        slot.iterations = 1
    arr[1].uv_mode = 9
    assert lib.rlic_convolve_batch_f64(arr, 2, 3, None) < 0
    assert _native.last_error().startswith("field 1: unknown uv_mode")
    arr[1].uv_mode = 0
    arr[0].kernel_len = 0
    assert lib.rlic_convolve_batch_f64(arr, 2, 3, None) < 0
    assert "kernel is empty" in _native.last_error()
    assert lib.rlic_convolve_batch_f64(arr, 0, 3, None) == 0


def test_batch_raises_for_the_first_invalid_field():
    ok = Field(np.ones((4, 4)), np.ones((4, 4)), np.ones((4, 4)), np.ones(3))
    neg = Field(-np.ones((4, 4)), np.ones((4, 4)), np.ones((4, 4)), np.ones(3))
    nan_kernel = Field(np.ones((4, 4)), np.ones((4, 4)), np.ones((4, 4)), np.array([1.0, np.nan]))
    shape = Field(np.ones((4, 4)), np.ones((4, 5)), np.ones((4, 4)), np.ones(3))
    with pytest.raises(ValueError, match="^Found invalid texture element"):
        convolve_batch([ok, neg, shape])
    with pytest.raises(ValueError, match="^Shape mismatch"):
        convolve_batch([ok, shape, neg])
    with pytest.raises(ValueError, match="^Found non-finite value"):
        convolve_batch([ok, nan_kernel, neg])
    # several problems in one field: the same ExceptionGroup as convolve
    both = Field(-np.ones((4, 4)), np.ones((4, 5)), np.ones((4, 4)), np.ones(3))
    with pytest.raises(Exception) as info:
        convolve_batch([ok, both])
    assert [type(e) for e in info.value.exceptions] == [ValueError, ValueError]


@pytest.mark.gpu
def test_batch_device_validation(gpu_device):
    import torch

    def t(a):
        return torch.from_numpy(np.asarray(a, dtype=np.float32)).to(gpu_device)

    ok = Field(t(np.ones((4, 4))), t(np.ones((4, 4))), t(np.ones((4, 4))), t(np.ones(3)))
    neg = Field(t(-np.eye(4)), t(np.ones((4, 4))), t(np.ones((4, 4))), t(np.ones(3)))
    inf_kernel = Field(t(np.ones((4, 4))), t(np.ones((4, 4))), t(np.ones((4, 4))),
                       t([1.0, np.inf, 1.0]))
    with pytest.raises(ValueError, match="^Found invalid texture element"):
        convolve_batch([ok, neg])
    with pytest.raises(ValueError, match="^Found non-finite value"):
        convolve_batch([ok, inf_kernel, neg])
    out = convolve_batch([ok, ok])
    assert all(o.shape == (4, 4) for o in out)


def test_batch_value_check_abi_without_a_gpu():
    """The batched value check validates its arguments before touching a
    device: a null result pointer and negative sizes are errors, an empty
    list passes."""
    import ctypes

    lib = _native.load()
    bad, what = ctypes.c_int64(7), ctypes.c_int(7)
    assert lib.rlic_batch_check_values_f32(None, 0, ctypes.byref(bad), ctypes.byref(what),
                                           None) == 0
    assert bad.value == -1 and what.value == 0
    assert lib.rlic_batch_check_values_f64(None, 0, None, None, None) < 0
    assert "first_bad" in _native.last_error()
    arr = (_native.FieldF32 * 1)()
    arr[0].ny, arr[0].nx, arr[0].kernel_len = -1, 4, 3
    assert lib.rlic_batch_check_values_f32(arr, 1, ctypes.byref(bad), None, None) < 0
    assert "negative size" in _native.last_error()
    arr[0].ny, arr[0].kernel_len = 4, 0  # null texture pointer
    assert lib.rlic_batch_check_values_f32(arr, 1, ctypes.byref(bad), None, None) < 0
    assert "null texture" in _native.last_error()
    assert lib.rlic_check_values_f64(None, 4, 4, 4, 1, None, 0, ctypes.byref(what), None) < 0
    assert "null texture" in _native.last_error()
    assert lib.rlic_check_values_f32(None, 0, 4, 4, 1, None, 0, None, None) < 0
    assert "null result" in _native.last_error()
    assert lib.rlic_check_values_f32(None, 0, 4, 4, 1, None, 0, ctypes.byref(what), None) == 0
    assert what.value == 0


def _native_first_bad(fields):
    """The first field failing the library's batched value check, called
    through the C ABI once per dtype as the batched call does, or None."""
    import ctypes

    import torch

    lib = _native.load()
    first = None
    for struct, check, dt in ((_native.FieldF32, lib.rlic_batch_check_values_f32, torch.float32),
                              (_native.FieldF64, lib.rlic_batch_check_values_f64, torch.float64)):
        index = [k for k, f in enumerate(fields) if f.texture.dtype == dt]
        arr = np.zeros(len(index), dtype=_native.FIELD_DTYPE)
        for row, k in enumerate(index):
            t, kern = fields[k].texture, fields[k].kernel
            arr[row]["texture"], arr[row]["kernel"] = t.data_ptr(), kern.data_ptr()
            arr[row]["tex_sy"], arr[row]["tex_sx"] = t.stride(0), t.stride(1)
            arr[row]["ny"], arr[row]["nx"] = t.shape
            arr[row]["kernel_len"] = kern.shape[0]
        bad = ctypes.c_int64(-2)
        stream = torch.cuda.current_stream().cuda_stream
        what = ctypes.c_int(-1)
        assert check(arr.ctypes.data_as(ctypes.POINTER(struct)), len(index), ctypes.byref(bad),
                     ctypes.byref(what), ctypes.c_void_p(stream)) == 0, _native.last_error()
        assert (bad.value == -1) == (what.value == 0) and what.value in (0, 1, 2, 3)
        assert bad.value >= -1
        if bad.value >= 0 and (first is None or index[bad.value] < first):
            first = index[bad.value]
    return first


@pytest.mark.gpu
def test_batch_value_check_finds_every_position(gpu_device):
    """The device check of a batch's element values (one launch per dtype)
    finds a single negative texture element or non-finite kernel weight
    wherever it sits, in contiguous and strided (transposed) textures of
    several sizes, reports the first failing field, and passes NaN and
    signed-zero texture elements, as ``rlic.convolve``'s checks do
    (src/rlic/_lib.py:174, :194)."""
    import torch

    rng = np.random.default_rng(8)
    shapes = [(1, 1), (1, 700), (700, 1), (37, 53), (300, 257), (1024, 1100)]
    fields = []
    for ny, nx in shapes:
        for dt in (torch.float32, torch.float64):
            tex = torch.rand((ny, nx), dtype=dt, device=gpu_device)
            tex.view(-1)[:: 97] = float("nan")
            tex.view(-1)[1:: 89] = -0.0
            fields.append(Field(tex, tex, tex, torch.rand(17, dtype=dt, device=gpu_device)))
            trans = torch.rand((nx, ny), dtype=dt, device=gpu_device).t()
            fields.append(Field(trans, trans, trans, torch.rand(5, dtype=dt, device=gpu_device)))
    assert _native_first_bad(fields) is None
    for trial in range(24):
        k = int(rng.integers(len(fields)))
        f = fields[k]
        ny, nx = f.texture.shape
        i, j = int(rng.integers(ny)), int(rng.integers(nx))
        if trial % 4 == 3:  # a non-finite kernel weight instead
            w = int(rng.integers(len(f.kernel)))
            saved = f.kernel[w].item()
            f.kernel[w] = float("inf") if trial % 8 == 3 else float("nan")
            assert _native_first_bad(fields) == k, (k, w)
            f.kernel[w] = saved
            continue
        saved = f.texture[i, j].item()
        f.texture[i, j] = -1e-30
        # a later field of the other dtype fails too: the first one is named
        later = min(len(fields) - 1, k + 1 + int(rng.integers(3)))
        g = fields[later]
        g_saved = g.texture[0, 0].item()
        if later != k:
            g.texture[0, 0] = -1.0
        assert _native_first_bad(fields) == k, (k, i, j, later)
        if later != k:
            g.texture[0, 0] = g_saved
        f.texture[i, j] = saved
    assert _native_first_bad(fields) is None
    # through the batched call: the first failing field's error, before any compute
    small = [Field(f.texture, f.u, f.v, f.kernel) for f in fields if f.texture.numel() < 5000]
    small[-1].texture[0, 0] = -2.0
    small[3].kernel[1] = float("nan")
    with pytest.raises(ValueError, match="^Found non-finite value"):
        convolve_batch(small)
    small[3].kernel[1] = 0.5
    with pytest.raises(ValueError, match="^Found invalid texture element"):
        convolve_batch(small)
    small[-1].texture[0, 0] = 0.0
    assert len(convolve_batch(small)) == len(small)


@pytest.mark.gpu
def test_batch_value_check_through_a_table(gpu_device):
    """More than 32 fields (the table path, not the kernel arguments): a
    bad field among 70 is found wherever it sits, and so is the first of
    two."""
    import torch

    rng = np.random.default_rng(70)
    fields = []
    for k in range(70):
        ny, nx = (int(s) for s in rng.integers(1, 90, size=2))
        tex = torch.rand((ny, nx), dtype=torch.float32, device=gpu_device)
        fields.append(Field(tex, tex, tex, torch.rand(9, device=gpu_device)))
    assert _native_first_bad(fields) is None
    for k in (0, 33, 37, 69):
        fields[k].texture[-1, -1] = -1.0
        assert _native_first_bad(fields) == k
        fields[50].kernel[3] = float("nan")
        assert _native_first_bad(fields) == min(k, 50)
        fields[50].kernel[3] = 0.5
        fields[k].texture[-1, -1] = 0.5
    assert _native_first_bad(fields) is None


@pytest.mark.gpu
def test_single_field_value_check(gpu_device):
    """``rlic_check_values_*`` (what validation runs for device tensors)
    agrees with NumPy's checks on every probe: one texture element below 0
    anywhere in contiguous, transposed and sliced textures, -inf, NaN and
    -0.0 elements (NaN and -0.0 pass), non-finite weights, empty textures
    and kernels."""
    import torch

    from rlic_amd._lib import _device_value_flags, _value_flags

    rng = np.random.default_rng(21)
    for dt in (torch.float32, torch.float64):
        base = torch.rand((300, 517), dtype=dt, device=gpu_device)
        kern = torch.rand(33, dtype=dt, device=gpu_device)
        views = {"contiguous": base, "transposed": base.t(), "sliced": base[3:250:3, 1::2],
                 "row": base[7:8], "column": base[:, 9:10]}
        for name, tex in views.items():
            assert _device_value_flags(tex, kern) == (False, False), name
            for _ in range(6):
                i, j = int(rng.integers(tex.shape[0])), int(rng.integers(tex.shape[1]))
                for val, neg in ((-1e-30, True), (float("-inf"), True), (float("nan"), False),
                                 (-0.0, False)):
                    saved = tex[i, j].item()
                    tex[i, j] = val
                    got = _device_value_flags(tex, kern)
                    want = (bool(np.any(tex.cpu().numpy() < 0)), False)
                    assert got == want == (neg, False), (name, i, j, val)
                    tex[i, j] = saved
        for val in (float("inf"), float("-inf"), float("nan")):
            w = int(rng.integers(33))
            k2 = kern.clone()
            k2[w] = val
            assert _device_value_flags(base, k2) == (False, True)
            base[5, 5] = -2.0
            assert _device_value_flags(base, k2) == (True, True)
            assert _device_value_flags(base, kern[::2]) == (True, False)  # strided kernel
            base[5, 5] = 0.5
        empty = torch.empty((0, 5), dtype=dt, device=gpu_device)
        assert _device_value_flags(empty, kern) == (False, False)
        assert _device_value_flags(base, kern[:0]) == (False, False)
        # other dtypes fall back to torch's reductions
        assert _device_value_flags(base.half(), kern.half()) is None
        assert _value_flags(-base.half(), kern.half()) == (True, False)


@pytest.mark.gpu
@pytest.mark.parametrize("profile", [0, 1, 2, 3])
def test_batch_and_serving_follow_each_arithmetic_profile(gpu_device, profile):
    """Fields with zero and negative velocity components, where the profiles
    differ (src/lib.rs:157-180), batched and served under each profile, equal
    the single calls under that profile."""
    from rlic_amd import _core
    from rlic_amd.serve import Server

    rng = np.random.default_rng(100 + profile)
    fields = []
    for k in range(6):
        dtype = np.float32 if k % 2 else np.float64
        ny, nx = 30 + 7 * k, 41 + 3 * k
        u = np.round(rng.random((ny, nx)) * 4 - 2).astype(dtype)  # many exact zeros and ties
        v = (rng.random((ny, nx)) - 0.5).astype(dtype)
        v[:, ::5] = -0.0
        fields.append(Field(rng.random((ny, nx)).astype(dtype), u, v,
                            rng.random(int(rng.integers(3, 40))).astype(dtype),
                            uv_mode=("velocity", "polarization")[k % 2],
                            boundaries=("closed", "periodic")[k % 3 == 0], iterations=1 + k % 3))
    outs = convolve_batch(fields, profile=profile)
    with Server(1 << 28, device=gpu_device, profile=profile) as srv:
        served = [srv.submit(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                             boundaries=f.boundaries, iterations=f.iterations) for f in fields]
        srv.run()
    for k, f in enumerate(fields):
        bs = rlic_amd._boundaries.BoundarySet.from_spec(f.boundaries)
        fn = _core.convolve_f32 if f.texture.dtype == np.float32 else _core.convolve_f64
        single = fn(f.texture, (f.u, f.v, f.uv_mode), f.kernel, (bs.x, bs.y), f.iterations,
                    profile=profile)
        exp = oracle_convolve(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                              boundaries=f.boundaries, iterations=f.iterations, profile=profile)
        assert np.array_equal(single, exp, equal_nan=True), k
        assert np.array_equal(outs[k], exp, equal_nan=True), k
        assert np.array_equal(served[k].result, exp, equal_nan=True), k


def test_quick_structure_check_agrees_with_validate():
    """The batched call's fast structural check of tensor fields accepts a
    field only when ``validate`` accepts it, with the same boundaries and
    dtype, and leaves every invalid field to ``validate`` (CPU tensors)."""
    import torch

    from rlic_amd._lib import validate
    from rlic_amd.batch import _quick_structure_check

    rng = np.random.default_rng(0)
    specs = [{"x": "periodic", "y": "closed"}, "closed", {"x": ["closed", "closed"], "y": "periodic"},
             {"x": "periodic", "y": ("periodic", "closed")}, "bogus", {"x": "closed"}]
    accepted = rejected = 0
    for k in range(240):
        ny, nx = int(rng.integers(1, 12)), int(rng.integers(1, 12))
        dt = torch.float32 if k % 2 else torch.float64
        t, u, v = (torch.rand(ny, nx, dtype=dt) for _ in range(3))
        kern = torch.rand(5, dtype=dt)
        if k % 7 == 0:
            u = torch.rand(ny, nx + 1, dtype=dt)
        if k % 11 == 0:
            kern = torch.rand(2, 2, dtype=dt)
        if k % 13 == 0:
            v = v.to(torch.float16)
        f = Field(t, u, v, kern, uv_mode=["velocity", "polarization", "sideways"][k % 3],
                  boundaries=specs[k % len(specs)], iterations=[0, 1, 3, -1][k % 4])
        quick = _quick_structure_check(f)
        try:
            full = validate(f.texture, f.u, f.v, kernel=f.kernel, uv_mode=f.uv_mode,
                            boundaries=f.boundaries, iterations=f.iterations, check_values=False)
        except Exception:
            full = None
        if quick is None:
            rejected += 1
            continue
        accepted += 1
        assert full is not None, k
        assert quick[1] == full[1] and (quick[0].x, quick[0].y) == (full[0].x, full[0].y), k
    assert accepted > 10 and rejected > 10
