"""What the trained pre-sizer gains over the method without it.

The first 65,536 paths of the pre-sizer's evaluation scene (outlines of the
uneven family on 8192^2) are drawn three ways: by the exact two-pass method
(count passes, read-back, write passes), by the learned single-pass method
(the MLP's capacities, fallback to the exact method on under-prediction) and
by the same single-pass method with capacities from a bound linear in the
segment count and the bounding box (no network).  The three frames must be
identical; the learned form must be faster than the exact form; the linear
row is reported beside it so the network's own contribution shows.
"""
import numpy as np
import pytest

from skity_amd import device, presizer

pytestmark = pytest.mark.gpu


def test_learned_presizing_gain_at_identical_output():
    ctx = device.Context(0)
    try:
        presizer.install(ctx, *presizer.load())
        # one warm-up slice so first-use allocations are not timed
        presizer.evaluate_large(ctx, n_paths=4096, slice_paths=4096, log=lambda m: None)
        rep = presizer.evaluate_large(ctx, n_paths=65536, slice_paths=65536,
                                      log=lambda m: print(m, flush=True))
    finally:
        ctx.close()
    for k in ("exact", "learned", "linear_estimate"):
        r = rep[k]
        print(f"{k}: {r['ms']:.1f} ms, {r['paths_per_s']:.0f} paths/s, fallback rate "
              f"{r['fallback_rate']:.4f}, under-estimate rate {r['underestimate_rate']:.4f}, "
              f"identical {r['identical_to_exact']}", flush=True)
    assert rep["learned"]["identical_to_exact"]
    assert rep["linear_estimate"]["identical_to_exact"]
    assert rep["learned"]["jobs"] == rep["exact"]["jobs"] > 60000
    gain = rep["exact"]["ms"] / rep["learned"]["ms"]
    print(f"learned vs exact: {gain:.2f}x; linear estimate vs exact: "
          f"{rep['exact']['ms'] / rep['linear_estimate']['ms']:.2f}x", flush=True)
    assert gain > 1.0
    assert rep["learned"]["fallback_rate"] < 0.05
