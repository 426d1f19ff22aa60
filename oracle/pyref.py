"""Pure-Python restatement of the reference convolution with exact arithmetic.

TEST INFRASTRUCTURE ONLY (tiny inputs: every operation is done on exact
rationals and rounded once to the target format, so it is slow).

Written independently of ``lic_oracle.cpp`` from the step rule of the
reference (``src/lib.rs:157-443``): it shares no code with the C++ oracle and
does not use the machine's floating-point unit for any rounded operation.
Each IEEE operation (+, -, *, /, fused multiply-add) is computed exactly
with :class:`fractions.Fraction` and rounded to nearest-even into binary32
or binary64, with IEEE rules for signed zeros, infinities and NaN.  Where
the reference calls ``mul_add`` the emulator rounds once; everywhere else
it rounds after every operation.  A disagreement with the C++ oracle
therefore points at a transcription or arithmetic mistake in one of the two.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

INF = math.inf
NAN = math.nan


class IEEEFormat:
    """Exact-rational emulation of one binary floating-point format."""

    def __init__(self, precision: int, emin: int, emax: int) -> None:
        self.p = precision
        self.emin = emin
        self.emax = emax

    # -- rounding ---------------------------------------------------------
    def round(self, q: Fraction) -> float:
        """Round an exact nonzero-or-zero rational to the nearest format value."""
        if q == 0:
            return 0.0
        negative = q < 0
        a = -q if negative else q
        e = a.numerator.bit_length() - a.denominator.bit_length()
        if Fraction(2) ** e > a:
            e -= 1
        # now 2**e <= a < 2**(e+1)
        e = max(e, self.emin)
        quantum = Fraction(2) ** (e - self.p + 1)
        scaled = a / quantum
        n = scaled.numerator // scaled.denominator
        rem = scaled - n
        if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
            n += 1
        value = n * quantum
        if value >= Fraction(2) ** (self.emax + 1):
            r = INF
        else:
            r = float(value)  # exactly representable
        return -r if negative else r

    # -- operations (arguments are format values stored as Python floats) --
    @staticmethod
    def signbit(x: float) -> bool:
        return math.copysign(1.0, x) < 0

    def add(self, a: float, b: float) -> float:
        if math.isnan(a) or math.isnan(b):
            return NAN
        if math.isinf(a) or math.isinf(b):
            if math.isinf(a) and math.isinf(b) and (a > 0) != (b > 0):
                return NAN
            return a if math.isinf(a) else b
        s = Fraction(a) + Fraction(b)
        if s == 0:
            # exact zero: -0 only when both addends are -0 (round to nearest)
            if a == 0 and b == 0 and self.signbit(a) and self.signbit(b):
                return -0.0
            return 0.0
        return self.round(s)

    def sub(self, a: float, b: float) -> float:
        return self.add(a, -b)

    def mul(self, a: float, b: float) -> float:
        if math.isnan(a) or math.isnan(b):
            return NAN
        neg = self.signbit(a) != self.signbit(b)
        if math.isinf(a) or math.isinf(b):
            if a == 0 or b == 0:
                return NAN
            return -INF if neg else INF
        p = Fraction(a) * Fraction(b)
        if p == 0:
            return -0.0 if neg else 0.0
        r = self.round(p)
        return math.copysign(r, -1.0 if neg else 1.0) if r == 0 else r

    def div(self, a: float, b: float) -> float:
        if math.isnan(a) or math.isnan(b):
            return NAN
        neg = self.signbit(a) != self.signbit(b)
        if math.isinf(a):
            if math.isinf(b):
                return NAN
            return -INF if neg else INF
        if math.isinf(b):
            return -0.0 if neg else 0.0
        if b == 0:
            if a == 0:
                return NAN
            return -INF if neg else INF
        if a == 0:
            return -0.0 if neg else 0.0
        r = self.round(Fraction(a) / Fraction(b))
        return math.copysign(r, -1.0 if neg else 1.0) if r == 0 else r

    def fma(self, a: float, b: float, c: float) -> float:
        if math.isnan(a) or math.isnan(b) or math.isnan(c):
            return NAN
        pneg = self.signbit(a) != self.signbit(b)
        if math.isinf(a) or math.isinf(b):
            if a == 0 or b == 0:
                return NAN
            pinf = -INF if pneg else INF
            if math.isinf(c) and (c > 0) != (pinf > 0):
                return NAN
            return pinf
        if math.isinf(c):
            return c
        exact = Fraction(a) * Fraction(b) + Fraction(c)
        if exact == 0:
            prod_is_zero = a == 0 or b == 0
            if prod_is_zero and c == 0 and pneg and self.signbit(c):
                return -0.0
            return 0.0
        r = self.round(exact)
        if r == 0:
            return -0.0 if exact < 0 else 0.0
        return r


BINARY64 = IEEEFormat(53, -1022, 1023)
BINARY32 = IEEEFormat(24, -126, 127)


def _fmt(dtype) -> IEEEFormat:
    return BINARY32 if np.dtype(dtype) == np.float32 else BINARY64


def time_to_next_pixel(velocity: float, frac: float, fma: bool, branchless: bool,
                       fmt: IEEEFormat) -> float:
    """Time to the next cell edge (reference ``src/lib.rs:157-180``)."""
    if branchless:
        if math.isnan(velocity):
            sig = NAN
        else:
            sig = -1.0 if fmt.signbit(velocity) else 1.0
        factor = fmt.add(1.0, sig)
        half_minus = fmt.sub(0.5, frac)
        if fma:
            remaining = fmt.fma(factor, half_minus, frac)
        else:
            remaining = fmt.add(fmt.mul(factor, half_minus), frac)
        q = fmt.div(remaining, velocity)
        return math.fabs(q)
    if velocity > 0:
        return fmt.div(fmt.sub(1.0, frac), velocity)
    if velocity < 0:
        return -fmt.div(frac, velocity)
    return INF


def _wrap(c: int, size: int, low: str, high: str) -> int:
    """Boundary rule after a step (``src/lib.rs:83-95``)."""
    if c < 0:
        return 0 if low == "closed" else size - 1
    if c >= size:
        return size - 1 if high == "closed" else 0
    return c


def step(state: dict, mu: float, mv: float, fma: bool, branchless: bool, fmt: IEEEFormat,
         shape: tuple[int, int], bounds: tuple[tuple[str, str], tuple[str, str]]) -> None:
    """Advance ``state`` = {i, j, fx, fy} by one cell (``src/lib.rs:235-273``)."""
    if mu == 0 and mv == 0:
        return
    tx = time_to_next_pixel(mu, state["fx"], fma, branchless, fmt)
    ty = time_to_next_pixel(mv, state["fy"], fma, branchless, fmt)
    # axis crossed first; a tie or an unordered pair goes to y
    axis = "x" if tx < ty else "y"
    vel, orth, t = (mu, mv, tx) if axis == "x" else (mv, mu, ty)
    coord, fpar, forth = ("j", "fx", "fy") if axis == "x" else ("i", "fy", "fx")
    if vel >= 0:
        state[coord] += 1
        state[fpar] = 0.0
    else:
        state[coord] -= 1
        state[fpar] = 1.0
    if fma:
        state[forth] = fmt.fma(t, orth, state[forth])
    else:
        state[forth] = fmt.add(state[forth], fmt.mul(t, orth))
    (xl, xr), (yl, yr) = bounds
    state["j"] = _wrap(state["j"], shape[1], xl, xr)
    state["i"] = _wrap(state["i"], shape[0], yl, yr)


def _norm_bounds(boundaries):
    if isinstance(boundaries, str):
        return ((boundaries, boundaries), (boundaries, boundaries))
    if isinstance(boundaries, dict):
        bx, by = boundaries["x"], boundaries["y"]
        bx = (bx, bx) if isinstance(bx, str) else tuple(bx)
        by = (by, by) if isinstance(by, str) else tuple(by)
        return (bx, by)
    return tuple(tuple(b) for b in boundaries)


def convolve(texture, u, v, *, kernel, uv_mode="velocity", boundaries="closed",
             iterations=1, profile=3):
    """Exact-arithmetic LIC of small float32/float64 arrays (returns ndarray)."""
    dtype = np.asarray(texture).dtype
    fmt = _fmt(dtype)
    fma = bool(profile & 1)
    branchless = bool(profile & 2)
    bounds = _norm_bounds(boundaries)
    U = [[float(x) for x in row] for row in np.asarray(u)]
    V = [[float(x) for x in row] for row in np.asarray(v)]
    K = [float(x) for x in np.asarray(kernel)]
    X = [[float(x) for x in row] for row in np.asarray(texture)]
    ny, nx = len(X), len(X[0]) if X else 0
    L = len(K)
    mid = L // 2

    def mac(acc, w, x):
        return fmt.fma(w, x, acc) if fma else fmt.add(acc, fmt.mul(w, x))

    for _ in range(iterations):
        Y = [[0.0] * nx for _ in range(ny)]
        for i0 in range(ny):
            for j0 in range(nx):
                acc = mac(0.0, K[mid], X[i0][j0])
                for sense, taps in ((+1, range(mid + 1, L)), (-1, range(mid - 1, -1, -1))):
                    st = {"i": i0, "j": j0, "fx": 0.5, "fy": 0.5}
                    prev = (0.0, 0.0)
                    for k in taps:
                        pu, pv = U[st["i"]][st["j"]], V[st["i"]][st["j"]]
                        if math.isnan(pu) or math.isnan(pv):
                            break
                        if uv_mode == "polarization":
                            dot = fmt.add(fmt.mul(pu, prev[0]), fmt.mul(pv, prev[1]))
                            if dot < 0:
                                pu, pv = -pu, -pv
                            prev = (pu, pv)
                        mu, mv = (pu, pv) if sense > 0 else (-pu, -pv)
                        step(st, mu, mv, fma, branchless, fmt, (ny, nx), bounds)
                        acc = mac(acc, K[k], X[st["i"]][st["j"]])
                Y[i0][j0] = acc
        X = Y
    return np.array(X, dtype=dtype).reshape(ny, nx)
