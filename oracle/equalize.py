"""CPU restatement of global histogram equalization (TEST INFRASTRUCTURE
ONLY: tests and the smoke check use it as the checker; the product path
never imports it).

The reference declares ``equalize_histogram_f32/_f64(image, nbins)`` in
``src/rlic/_core.pyi:30-37`` with no implementation anywhere in the
repository (``README.md:19-23`` sends users to the sister project ``ahe``,
which is not installed here).  So this is not a restatement of reference
code: it fixes the semantics the GPU kernels implement, from the textbook
definition of histogram equalization (the output of a pixel is the
empirical cumulative distribution of the image at that pixel's bin):

* finite pixels take part; NaN and +-inf pixels are left out of the range
  and the histogram and are copied to the output unchanged;
* vmin, vmax = the smallest and largest finite value; N = finite pixels;
* bin(x) = min(floor((x - vmin) * (nbins / (vmax - vmin))), nbins - 1),
  every operation in IEEE double (float32 pixels are widened exactly);
  when vmax - vmin overflows, both are halved first:
  bin(x) = min(floor((x/2 - vmin/2) * (nbins / (vmax/2 - vmin/2))), nbins - 1);
  when vmax == vmin every finite pixel is in bin 0;
* out(x) = cdf[bin(x)] / N in double, rounded once to the image dtype,
  cdf[b] = the number of finite pixels in bins 0..b.

Parity against the reference is unpinned (it has no implementation).  What
pins this restatement (tests/test_equalize.py): monotonicity, permutation
invariance, the closed form of ramp and constant images, the value set of
two-level images, and a brute-force O(N^2) count on small images.
"""

from __future__ import annotations

import numpy as np


def bins(image: np.ndarray, nbins: int) -> tuple[np.ndarray, np.ndarray, int]:
    """(bin index of every pixel, -1 where not finite; finite mask; N)."""
    if nbins < 1:
        raise ValueError("nbins must be at least 1")
    x = np.asarray(image).astype(np.float64)
    finite = np.isfinite(x)
    n = int(finite.sum())
    b = np.full(x.shape, -1, dtype=np.int64)
    if n == 0:
        return b, finite, 0
    vmin = x[finite].min()
    vmax = x[finite].max()
    xf = x[finite]
    if vmax == vmin:
        b[finite] = 0
        return b, finite, n
    with np.errstate(over="ignore"):
        d = vmax - vmin
    if np.isfinite(d):
        t = (xf - vmin) * (np.float64(nbins) / d)
    else:
        t = (xf * 0.5 - vmin * 0.5) * (np.float64(nbins) / (vmax * 0.5 - vmin * 0.5))
    b[finite] = np.minimum(np.floor(t), nbins - 1).astype(np.int64)
    return b, finite, n


def equalize_histogram(image: np.ndarray, nbins: int) -> np.ndarray:
    """Global histogram equalization of a 2-D float32/float64 image."""
    image = np.asarray(image)
    assert image.dtype in (np.float32, np.float64) and image.ndim == 2
    b, finite, n = bins(image, nbins)
    out = image.copy()
    if n == 0:
        return out
    hist = np.bincount(b[finite], minlength=nbins)
    cdf = np.cumsum(hist)
    out[finite] = (cdf[b[finite]].astype(np.float64) / np.float64(n)).astype(image.dtype)
    return out


def equalize_histogram_bruteforce(image: np.ndarray, nbins: int) -> np.ndarray:
    """The same mapping by direct counting: out(x) = #{finite y with
    bin(y) <= bin(x)} / N, pixel by pixel (small images only)."""
    image = np.asarray(image)
    b, finite, n = bins(image, nbins)
    out = image.copy()
    flat_b = b.ravel()
    fin = finite.ravel()
    res = out.ravel()
    for p in range(flat_b.size):
        if fin[p]:
            count = sum(1 for q in range(flat_b.size) if fin[q] and flat_b[q] <= flat_b[p])
            res[p] = np.float64(count) / np.float64(n)
    return res.reshape(image.shape)


# ---------------------------------------------------------------------------
# Adaptive (tile-based) equalization: the two other equalization stubs of the
# reference, ``equalize_histogram_sliding_tile_f32/_f64(image, nbins,
# tile_shape)`` and ``equalize_histogram_tile_interpolation_f32/_f64(image,
# nbins, tile_shape)`` (src/rlic/_core.pyi:38-57), likewise without any
# implementation in the reference.  The semantics fixed here (parity with
# the reference unpinned):
#
# * bins are the global ones above (the finite range of the whole image);
#   NaN and +-inf pixels are copied unchanged and never counted;
# * sliding tile (th, tw): the window of pixel (i, j) is rows
#   [i - (th - 1) // 2, i - (th - 1) // 2 + th) and columns likewise,
#   clipped to the image; out = #{finite q in window: bin(q) <= bin(p)} /
#   #{finite q in window}, in double, rounded once;
# * tile interpolation (th, tw): tiles of th x tw pixels from (0, 0) (the
#   last row and column of tiles may be partial); tile t maps bin b to
#   M_t(b) = cdf_t(b) / N_t in double (the global mapping when the tile has
#   no finite pixel); tile centres c(k) = k * t + (h_k - 1) / 2 (h_k the
#   tile's extent); a row i below the first centre or above the last uses
#   that tile alone, otherwise the two tiles k, k + 1 with
#   c(k) <= i < c(k + 1) and w = (i - c(k)) / (c(k + 1) - c(k)); columns
#   alike; out = lerp(lerp(M00, M01, wx), lerp(M10, M11, wx), wy) with
#   lerp(a, b, w) = a + w * (b - a), each operation in double, rounded once
#   to the image dtype.
# ---------------------------------------------------------------------------


def _check_tile(tile_shape) -> tuple[int, int]:
    th, tw = (int(t) for t in tile_shape)
    if th < 1 or tw < 1:
        raise ValueError("tile_shape must be positive")
    return th, tw


def _window_counts_shifted(b: np.ndarray, th: int, tw: int) -> tuple[np.ndarray, np.ndarray]:
    """(#{q in window: bin(q) <= bin(p)}, #{finite q in window}) per pixel,
    by one array pass per window offset."""
    h, w = b.shape
    up, left = (th - 1) // 2, (tw - 1) // 2
    padded = np.full((h + th - 1, w + tw - 1), -1, dtype=np.int64)
    padded[up:up + h, left:left + w] = b
    below = np.zeros((h, w), dtype=np.int64)
    count = np.zeros((h, w), dtype=np.int64)
    for dy in range(th):
        for dx in range(tw):
            q = padded[dy:dy + h, dx:dx + w]
            ok = q >= 0
            count += ok
            below += ok & (q <= b)
    return below, count


def _window_sums(ind: np.ndarray, th: int, tw: int) -> np.ndarray:
    """Per pixel, the sum of `ind` over its clipped window (summed-area table)."""
    h, w = ind.shape
    s = np.zeros((h + 1, w + 1), dtype=np.int64)
    s[1:, 1:] = ind.astype(np.int64).cumsum(0).cumsum(1)
    i, j = np.arange(h), np.arange(w)
    r0 = np.clip(i - (th - 1) // 2, 0, h)
    r1 = np.clip(i - (th - 1) // 2 + th, 0, h)
    c0 = np.clip(j - (tw - 1) // 2, 0, w)
    c1 = np.clip(j - (tw - 1) // 2 + tw, 0, w)
    return (s[np.ix_(r1, c1)] - s[np.ix_(r0, c1)] - s[np.ix_(r1, c0)] + s[np.ix_(r0, c0)])


def _window_counts_integral(b: np.ndarray, th: int, tw: int) -> tuple[np.ndarray, np.ndarray]:
    """The same counts by summed-area tables: one per distinct bin."""
    finite = b >= 0
    count = _window_sums(finite, th, tw)
    below = np.zeros(b.shape, dtype=np.int64)
    for k in np.unique(b[finite]):
        at = b == k
        below[at] = _window_sums(finite & (b <= k), th, tw)[at]
    return below, count


def equalize_histogram_sliding_tile(image: np.ndarray, nbins: int, tile_shape,
                                    method: str | None = None) -> np.ndarray:
    """Sliding-window histogram equalization.  The window counts come from
    shifted comparisons (one pass per window offset) or from summed-area
    tables (one per distinct bin), whichever takes fewer passes (`method`:
    "shifted" or "integral" to force one); both count exactly."""
    image = np.asarray(image)
    assert image.dtype in (np.float32, np.float64) and image.ndim == 2
    th, tw = _check_tile(tile_shape)
    b, finite, n = bins(image, nbins)
    out = image.copy()
    if n == 0:
        return out
    h, w = image.shape
    # a window of 2n - 1 covers the image from every pixel, as any larger one
    th, tw = min(th, 2 * h - 1), min(tw, 2 * w - 1)
    if method is None:
        method = "shifted" if th * tw <= np.unique(b[finite]).size else "integral"
    counts = _window_counts_shifted if method == "shifted" else _window_counts_integral
    below, count = counts(b, th, tw)
    out[finite] = (below[finite].astype(np.float64) / count[finite].astype(np.float64)
                   ).astype(image.dtype)
    return out


def _centres(n: int, t: int) -> np.ndarray:
    k = np.arange(-(-n // t))
    extent = np.minimum(t, n - k * t)
    return (k * t).astype(np.float64) + (extent - 1).astype(np.float64) * 0.5


def _axis_weights(n: int, t: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Per index along an axis: the two tiles and the weight of the second."""
    c = _centres(n, t)
    a0 = np.zeros(n, dtype=np.int64)
    a1 = np.zeros(n, dtype=np.int64)
    wt = np.zeros(n, dtype=np.float64)
    for i in range(n):
        x = np.float64(i)
        if len(c) == 1 or x <= c[0]:
            continue
        if x >= c[-1]:
            a0[i] = a1[i] = len(c) - 1
            continue
        k = int(np.searchsorted(c, x, side="right")) - 1  # c[k] <= x < c[k + 1]
        a0[i], a1[i] = k, k + 1
        wt[i] = (x - c[k]) / (c[k + 1] - c[k])
    return a0, a1, wt


def _lerp(a, b, w):
    return a + w * (b - a)


def equalize_histogram_tile_interpolation(image: np.ndarray, nbins: int,
                                          tile_shape) -> np.ndarray:
    """Tile-interpolated adaptive histogram equalization."""
    image = np.asarray(image)
    assert image.dtype in (np.float32, np.float64) and image.ndim == 2
    th, tw = _check_tile(tile_shape)
    b, finite, n = bins(image, nbins)
    out = image.copy()
    if n == 0:
        return out
    h, w = image.shape
    ty_n, tx_n = -(-h // th), -(-w // tw)
    gmap = np.cumsum(np.bincount(b[finite], minlength=nbins)).astype(np.float64) / np.float64(n)
    maps = np.empty((ty_n, tx_n, nbins), dtype=np.float64)
    for ty in range(ty_n):
        for tx in range(tx_n):
            tb = b[ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw]
            tb = tb[tb >= 0]
            if tb.size == 0:
                maps[ty, tx] = gmap
            else:
                maps[ty, tx] = (np.cumsum(np.bincount(tb, minlength=nbins)).astype(np.float64)
                                / np.float64(tb.size))
    y0, y1, wy = _axis_weights(h, th)
    x0, x1, wx = _axis_weights(w, tw)
    ii, jj = np.nonzero(finite)
    bb = b[ii, jj]
    top = _lerp(maps[y0[ii], x0[jj], bb], maps[y0[ii], x1[jj], bb], wx[jj])
    bot = _lerp(maps[y1[ii], x0[jj], bb], maps[y1[ii], x1[jj], bb], wx[jj])
    out[ii, jj] = _lerp(top, bot, wy[ii]).astype(image.dtype)
    return out


def equalize_histogram_sliding_tile_bruteforce(image: np.ndarray, nbins: int,
                                               tile_shape) -> np.ndarray:
    """The sliding-tile mapping pixel by pixel with explicit window loops
    (small images only)."""
    image = np.asarray(image)
    th, tw = _check_tile(tile_shape)
    b, finite, n = bins(image, nbins)
    out = image.copy()
    h, w = image.shape
    for i in range(h):
        for j in range(w):
            if not finite[i, j]:
                continue
            below = count = 0
            for r in range(max(0, i - (th - 1) // 2), min(h, i - (th - 1) // 2 + th)):
                for c in range(max(0, j - (tw - 1) // 2), min(w, j - (tw - 1) // 2 + tw)):
                    if finite[r, c]:
                        count += 1
                        below += b[r, c] <= b[i, j]
            out[i, j] = np.float64(below) / np.float64(count)
    return out
