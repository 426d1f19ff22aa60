"""The C ABI library loads and exports exactly what include/rlic_hip.h
declares; argument errors come back as status codes (CPU only: no call here
reaches a GPU)."""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from rlic_amd import _core, _native

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "rlic_hip.h"


def _declared_functions() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(rlic_[a-z0-9_]+)\s*\(", text))


def test_library_loads():
    lib = _native.load()
    assert isinstance(lib, ctypes.CDLL)
    assert lib.rlic_version().decode()


def test_every_declared_symbol_is_exported():
    declared = _declared_functions()
    assert {"rlic_convolve_f32", "rlic_convolve_f64", "rlic_last_error"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    assert set(_native.SIGNATURES) <= exported


def _call(**over):
    a = np.ones((4, 4))
    args = dict(texture=a.ctypes.data, tex_sy=4, tex_sx=1, u=a.ctypes.data, u_sy=4, u_sx=1,
                v=a.ctypes.data, v_sy=4, v_sx=1, ny=4, nx=4, kernel=a.ctypes.data,
                kernel_len=3, uv_mode=0, x_left=0, x_right=0, y_left=0, y_right=0,
                iterations=1, out=a.ctypes.data, location=0, profile=3, stream=None)
    args.update(over)
    return _native.load().rlic_convolve_f64(*args.values())


@pytest.mark.parametrize(
    "over, message",
    [
        (dict(ny=-1), "negative image size"),
        (dict(kernel_len=-2), "negative kernel length"),
        (dict(iterations=-1), "Invalid number of iterations"),
        (dict(uv_mode=7), "unknown uv_mode"),
        (dict(x_right=3), "unknown boundary"),
        (dict(location=5), "unknown memory location"),
        (dict(profile=9), "unknown arithmetic profile"),
        (dict(texture=None), "null array pointer"),
        (dict(kernel_len=0), "kernel is empty"),
    ],
)
def test_argument_errors_are_status_codes(over, message):
    status = _call(**over)
    assert status < 0
    assert message in _native.last_error()


def test_empty_image_needs_no_device():
    # nothing to compute: returns before touching the GPU, like the reference's
    # zero-trip pixel loop
    assert _call(ny=0) == 0
    assert _call(nx=0, kernel_len=0) == 0


@pytest.mark.parametrize(
    "ny, nx, message",
    [
        # more than 2^31-1 pixels
        (1 << 16, 1 << 16, "exceeds the supported 2^31-1 pixels"),
        # a side beyond 2^31-1: rejected before ny * nx is formed
        (1 << 40, 4, "exceeds the supported 2^31-1 pixels"),
        # ny * nx would wrap to 0 in 64-bit arithmetic: still rejected, not empty
        (1 << 32, 1 << 32, "exceeds the supported 2^31-1 pixels"),
        # thin images under the pixel limit whose (ny+2) x (nx+2) guarded
        # cell array exceeds 32-bit cell indices
        (1, 1_431_655_764, "guarded array of 3 x 1431655766 cells"),
        (2_000_000_000, 1, "guarded array of 2000000002 x 3 cells"),
    ],
)
def test_size_limits_are_status_codes(ny, nx, message):
    # pointers are never dereferenced: the checks run before any HIP call
    assert _call(ny=ny, nx=nx) == _native.RLIC_ERR_INVALID_ARGUMENT
    assert message in _native.last_error()


def test_largest_thin_image_passes_the_size_checks():
    # 1 x 1,431,655,763 has 3 x 1,431,655,765 = 4,294,967,295 guarded cells:
    # accepted by the size checks, so the call gets as far as its null
    # pointer check (no GPU work is started)
    assert _call(ny=1, nx=1_431_655_763, out=None) == _native.RLIC_ERR_INVALID_ARGUMENT
    assert "null array pointer" in _native.last_error()


def test_empty_image_with_a_huge_side_is_empty():
    # the reference's zero-trip loop: an empty image returns at once, whatever
    # its other side
    assert _call(ny=0, nx=1 << 62) == 0
    assert _call(ny=1 << 62, nx=0, kernel=None) == 0


def test_server_field_bytes_rejects_oversized_images():
    lib = _native.load()
    assert lib.rlic_server_field_bytes_f32(1, 1_431_655_764, 3, 0, 1) == -1
    assert lib.rlic_server_field_bytes_f32(1 << 40, 4, 3, 0, 1) == -1
    # empty images are held as 0 x 0 with no weights
    assert lib.rlic_server_field_bytes_f32(0, 1 << 40, 1 << 40, 1, 5) == \
        lib.rlic_server_field_bytes_f32(0, 0, 0, 1, 5)


def test_python_binding_argument_errors():
    a = np.ones((4, 4))
    k = np.ones(3)
    with pytest.raises(TypeError):
        _core.convolve_f64(a.astype(np.float32), (a, a, "velocity"), k, (("closed",) * 2,) * 2)
    with pytest.raises(TypeError):
        _core.convolve_f64(a, (a, a, "velocity"), np.ones((3, 3)), (("closed",) * 2,) * 2)
    with pytest.raises(OverflowError):
        _core.convolve_f64(a, (a, a, "velocity"), k, (("closed",) * 2,) * 2, 2**70)
    with pytest.raises(ValueError):
        _core.convolve_f64(a, (a, a, "sideways"), k, (("closed",) * 2,) * 2)
    with pytest.raises(ValueError):
        _core.convolve_f64(a, (a, a, "velocity"), k, (("open", "closed"), ("closed",) * 2))
