"""Shared helpers for the parity tests (inputs and comparisons)."""
import numpy as np

from skity_amd import abi, scenes


def fill_only(sc):
    """Turn every draw of a generated scene into a plain fill (in place)."""
    s = sc.scene.contents
    for p in range(s.n_paths):
        s.paints[p].style = abi.STYLE_FILL
    return sc


def no_blur(sc):
    s = sc.scene.contents
    for p in range(s.n_paths):
        s.paints[p].blur_radius = 0.0
    return sc


def diff_report(got, want):
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    return int(d.max()), int((d.max(-1) > 0).sum())


def workload(k, n=0, seed=None):
    return scenes.generate(k, n_paths=n, seed=seed)
