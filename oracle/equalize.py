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
