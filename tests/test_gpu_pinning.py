"""GPU parity of what the north star calls bit-exact, beyond pixels:

* the analytic-AA integer arithmetic of the walk against closed forms
  (tests/aaa_forms.py: rectangles at 1/256-px offsets, slivers within one
  row, 45-degree half-planes), so that the device and the oracle are both
  pinned to the same independent prediction;
* tile bins: per raster job, the 16x16 tiles and planes the walk allocated
  (hipr_surface_get_tile_bins) equal the oracle's integer definition of bins
  over the reference's spans (orc_path_tile_bins), in the exact two-pass form,
  the learned single-pass form and with row-banded walks;
* edge records of stroke outlines: the edges the device built from its own
  stroke outlines (hipr_surface_get_job_edges) equal the oracle's edges of
  PathStroker's outline, record for record, in both stroke forms (the exact
  count + write passes and the single pass into estimated ranges).
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer

import aaa_forms as F
from helpers import workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    prm, margin = presizer.load()
    presizer.install(c, prm, margin)
    yield c
    c.close()


@pytest.mark.parametrize("make,seed", [(F.rect_scene, 7), (F.rect_scene, 8),
                                       (F.parallelogram_scene, 9),
                                       (F.parallelogram_scene, 10)])
def test_aaa_closed_forms(ctx, make, seed):
    sc, want = make(seed)
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(sc.scene)
    got = surf.read()
    surf.close()
    exp = F.pixels_of(want, sc.width, sc.height)
    bad = np.argwhere((got != exp).any(-1))
    print(f"{make.__name__}({seed}): {len(want)} covered pixels, {len(bad)} differ", flush=True)
    assert bad.size == 0, (f"first differing pixel (y, x) = {tuple(bad[0])}: device "
                           f"{got[tuple(bad[0])]} vs closed form {exp[tuple(bad[0])]}")


def _draw(ctx, sc, **kw):
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(sc.scene, **kw)
    return surf


FORMS = {
    "exact": dict(mode=abi.MODE_EXACT),
    "exact_unbanded": dict(mode=abi.MODE_EXACT, band_rows=-1),
    "learned": dict(mode=abi.MODE_LEARNED),
    "learned_all_banded": dict(mode=abi.MODE_LEARNED, band_min_edges=-1),
}


@pytest.mark.parametrize("form", sorted(FORMS))
@pytest.mark.parametrize("wl,n", [(1, 0), (2, 40)])
def test_tile_bins_match_oracle(ctx, wl, n, form):
    sc = workload(wl, n)
    surf = _draw(ctx, sc, **FORMS[form])
    path, kind, canvas = surf.jobs()
    bins = surf.tile_bins()
    n_bins = n_plane1 = 0
    for j in range(len(path)):
        assert kind[j] in (0, 1) and canvas[j] == 0
        want, stacked = O.path_tile_bins(sc.scene, int(path[j]), int(kind[j]), sc.width,
                                         sc.height)
        assert stacked == 0
        got = bins[j]
        n_bins += len(want)
        n_plane1 += int((want & 1).sum())
        if not np.array_equal(got, want):
            extra = np.setdiff1d(got, want)[:4]
            missing = np.setdiff1d(want, got)[:4]
            pytest.fail(f"job {j} (path {path[j]}, kind {kind[j]}): {len(got)} bins vs "
                        f"{len(want)}; extra {extra}, missing {missing}")
    print(f"workload {wl} ({len(path)} jobs, {form}): {n_bins} bins equal, "
          f"{n_plane1} of them plane 1", flush=True)
    surf.close()


@pytest.mark.parametrize("form", ["exact", "learned"])
def test_stroke_outline_edges_match_oracle(ctx, form):
    sc = workload(2, 160)
    s = sc.scene.contents
    n_stroke = sum(1 for p in range(s.n_paths) if s.paints[p].style != abi.STYLE_FILL)
    assert n_stroke >= 20
    surf = _draw(ctx, sc, **FORMS[form])
    path, kind, _ = surf.jobs()
    checked = 0
    for j in range(len(path)):
        got = surf.job_edges(j)
        want = O.path_edges(sc.scene, int(path[j]), int(kind[j]))
        assert got.shape == want.shape, (f"job {j} (path {path[j]}, kind {kind[j]}): "
                                         f"{got.shape[0]} edges vs {want.shape[0]}")
        bad = np.nonzero((got != want).any(1))[0]
        assert bad.size == 0, (f"job {j} (path {path[j]}, kind {kind[j]}): edge {bad[0]} "
                               f"{got[bad[0]]} vs {want[bad[0]]}")
        checked += int(kind[j] == 1)
    print(f"{form}: {checked} stroke outlines and {len(path) - checked} fills, every edge "
          f"record equal", flush=True)
    assert checked == n_stroke
    surf.close()


def test_clear_blend_keeps_zero_cover_spans(ctx):
    """The device keeps the reference's zero-cover spans for a blend mode
    that changes pixels under a transparent source: kClear parallelograms
    clear exactly the pixels the closed form names."""
    sc, cleared = F.clear_parallelogram_scene(12)
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(sc.scene)
    got = surf.read()
    surf.close()
    backdrop = got[0, 0].copy()
    want = np.empty_like(got)
    want[...] = backdrop
    for (x, y) in cleared:
        want[y, x] = 0
    bad = np.argwhere((got != want).any(-1))
    print(f"kClear parallelograms: {len(cleared)} pixels cleared, {len(bad)} differ", flush=True)
    assert bad.size == 0
