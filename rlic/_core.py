"""``rlic._core``: the native entry points (reference PyO3 module
``src/lib.rs:445-485``, stub ``src/rlic/_core.pyi``), here over
``librlic_hip.so``."""

from rlic_amd._core import (  # noqa: F401
    convolve_f32,
    convolve_f64,
    equalize_histogram_f32,
    equalize_histogram_f64,
    equalize_histogram_sliding_tile_f32,
    equalize_histogram_sliding_tile_f64,
    equalize_histogram_tile_interpolation_f32,
    equalize_histogram_tile_interpolation_f64,
)
