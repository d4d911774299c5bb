"""GPU parity of stroked draws, prepared scenes and the pre-sized mode.

Strokes are outlined on the device (PathStroker restated in
skity_amd/csrc/hipr_stroke.hpp) and then rasterised like fills, so the bar is
the same as for fills: the outline is float code evaluated in the same order
as the reference, the rasteriser is integer code, and the expected pixel
difference is 0 (tolerance 1/255 per channel).
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer
from skity_amd.canvas import Canvas, Paint, Path

from helpers import diff_report, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    yield c
    c.close()


def render_gpu(ctx, scene, w, h, **kw):
    surf = device.Surface(ctx, w, h)
    surf.draw(scene, **kw)
    img = surf.read()
    st = surf.stats()
    surf.close()
    return img, st


def test_stroke_then_fill_kat(ctx):
    """SWCanvas.StrokeThenFillDrawsFillAfterStroke (test/ut/render/sw_canvas_test.cc:67-85)."""
    c = Canvas(48, 48)
    paint = Paint()
    paint.set_style(abi.STYLE_STROKE_THEN_FILL).set_stroke_width(10.0)
    paint.set_stroke_color((1, 0, 0, 1)).set_fill_color((1, 1, 1, 1))
    c.draw_path(Path().add_rect(10, 10, 34, 34), paint)
    sc = c.scene()
    got, _ = render_gpu(ctx, sc.scene, 48, 48)
    assert list(got[20, 12]) == [255, 255, 255, 255]
    assert list(got[20, 7]) == [255, 0, 0, 255]
    want = O.render(sc.scene, 48, 48)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("join", [abi.JOIN_MITER, abi.JOIN_ROUND, abi.JOIN_BEVEL])
@pytest.mark.parametrize("cap", [abi.CAP_BUTT, abi.CAP_ROUND, abi.CAP_SQUARE])
def test_stroke_joins_caps(ctx, join, cap):
    c = Canvas(160, 120)
    p = Paint((0.1, 0.6, 0.9, 0.8)).set_style(abi.STYLE_STROKE).set_stroke_width(9.0)
    p.set_stroke_join(join).set_stroke_cap(cap)
    open_path = Path().move_to(12, 100).line_to(40, 20).line_to(70, 95).quad_to(100, 10, 140, 60)
    c.draw_path(open_path, p)
    closed = Path().move_to(20, 20).cubic_to(60, -10, 120, 50, 60, 70).line_to(25, 60).close()
    c.draw_path(closed, p)
    c.draw_path(Path().add_circle(120, 90, 18), p)
    sc = c.scene()
    got, st = render_gpu(ctx, sc.scene, 160, 120)
    want = O.render(sc.scene, 160, 120)
    mx, cnt = diff_report(got, want)
    assert mx <= 1 and cnt == 0, f"max diff {mx}, {cnt} pixels"


@pytest.mark.parametrize("seed", [21, 22])
def test_curves_and_strokes_2048(ctx, seed):
    sc = workload(2, 40, seed=seed)
    got, st = render_gpu(ctx, sc.scene, sc.width, sc.height)
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"curves+strokes seed {seed}: max diff {mx}, differing pixels {cnt}, stats {st}",
          flush=True)
    assert mx <= 1
    assert cnt == 0


def test_all_strokes_2048(ctx):
    sc = workload(2, 24, seed=23)
    s = sc.scene.contents
    for p in range(s.n_paths):
        s.paints[p].style = [abi.STYLE_STROKE, abi.STYLE_STROKE_AND_FILL,
                             abi.STYLE_STROKE_THEN_FILL][p % 3]
        s.paints[p].stroke_width = [1.5, 4.0, 11.0, 23.0][p % 4]
        s.paints[p].join = p % 3
        s.paints[p].cap = (p // 3) % 3
        s.paints[p].miter_limit = 4.0
        s.paints[p].stroke_color[0] = 1.0 - s.paints[p].color[0]
    got, st = render_gpu(ctx, sc.scene, sc.width, sc.height)
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"all strokes: max diff {mx}, differing pixels {cnt}, stats {st}", flush=True)
    assert mx <= 1
    assert cnt == 0


def test_prepared_scene_redraw(ctx):
    """A prepared scene drawn twice gives the same pixels as a one-shot draw."""
    sc = workload(2, 32, seed=24)
    prep = device.Prepared(ctx, sc.scene)
    a, _ = render_gpu(ctx, prep, sc.width, sc.height)
    b, _ = render_gpu(ctx, prep, sc.width, sc.height)
    c, _ = render_gpu(ctx, sc.scene, sc.width, sc.height)
    prep.close()
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)


def test_presized_mode_matches_exact(ctx):
    """Mode 1 (capacities from estimates, exact fallback on overflow) draws the
    same pixels as the exact two-pass mode, whether the estimates are short
    (every draw falls back) or ample (none does: a 2048^2 canvas has 16384
    tiles and a tile holds at most two planes)."""
    sc = workload(2, 32, seed=25)
    n = sc.n_paths
    ref, _ = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=0)
    short, st_s = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=1,
                             edge_estimates=np.ones(n, np.int32),
                             tile_estimates=np.ones(n, np.int32))
    ample, st_a = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=1,
                             edge_estimates=np.full(n, 20000, np.int32),
                             tile_estimates=np.full(n, 2 * 16384, np.int32))
    print(f"fallbacks: short {st_s['fallbacks']}, ample {st_a['fallbacks']}", flush=True)
    assert np.array_equal(short, ref)
    assert np.array_equal(ample, ref)
    assert st_s["fallbacks"] > 0 and st_s["underestimates"] > 0
    assert st_a["fallbacks"] == 0 and st_a["underestimates"] == 0


def test_overflow_pool_and_exact_fallback(ctx):
    """Plane estimates of 1 per path: the shared overflow pool serves the
    first under-estimated paths and the exact path re-walks those that find
    it spent; edge estimates are ample, so no flatten fallback occurs.  A
    mild under-estimate is absorbed by the pool alone."""
    sc = workload(2, 64, seed=27)
    n = sc.n_paths
    ref, st_ref = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=0)
    edges = np.full(n, 20000, np.int32)
    spent, st1 = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=1, edge_estimates=edges,
                            tile_estimates=np.ones(n, np.int32))
    print("pool spent:", st1["underestimates"], "under-estimated,", st1["fallbacks"], "re-walked",
          flush=True)
    assert np.array_equal(spent, ref)
    assert st1["underestimates"] > st1["raster_paths"] // 2
    assert st1["fallbacks"] > 0
    assert st1["tile_planes"] == st_ref["tile_planes"]
    # a uniform estimate of 80% of the mean plane count: the pool holds the rest
    mild, st2 = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=1, edge_estimates=edges,
                           tile_estimates=np.full(n, max(1, st_ref["tile_planes"] * 4 // (5 * n)),
                                                  np.int32))
    print("mild:", st2["underestimates"], "under-estimated,", st2["fallbacks"], "re-walked",
          flush=True)
    assert np.array_equal(mild, ref)


@pytest.mark.parametrize("band_rows", [16, 48, 0])
def test_row_bands_match_single_band(ctx, band_rows):
    """The write pass restarts each row band from the snapshot the count pass
    took at a tile-row boundary; the bands together draw exactly what one
    sequential walk per path draws, with the same plane and tile counts."""
    sc = workload(2, 48, seed=31)
    ref, st_ref = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=0, band_rows=-1)
    got, st = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=0, band_rows=band_rows)
    print(f"band_rows {band_rows}: planes {st['tile_planes']} vs {st_ref['tile_planes']}",
          flush=True)
    assert np.array_equal(got, ref)
    assert st["tile_planes"] == st_ref["tile_planes"]
    assert st["tiles"] == st_ref["tiles"]
    want = O.render(sc.scene, sc.width, sc.height)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("band_rows,band_min_edges", [(16, -1), (0, -1), (32, 100)])
def test_row_bands_with_estimates(ctx, band_rows, band_min_edges):
    """Estimates mode with banding on: the schedule pass records the
    snapshots, the bands of a job share its estimated plane range, and short
    estimates fall back."""
    sc = workload(2, 40, seed=32)
    n = sc.n_paths
    kw = dict(band_rows=band_rows, band_min_edges=band_min_edges)
    ref, _ = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=0, band_rows=-1)
    short, st_s = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=1,
                             edge_estimates=np.full(n, 3000, np.int32),
                             tile_estimates=np.full(n, 40, np.int32), **kw)
    ample, st_a = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=1,
                             edge_estimates=np.full(n, 20000, np.int32),
                             tile_estimates=np.full(n, 2 * 16384, np.int32), **kw)
    print(f"fallbacks: short {st_s['fallbacks']}, ample {st_a['fallbacks']}", flush=True)
    assert np.array_equal(short, ref)
    assert np.array_equal(ample, ref)
    assert st_s["underestimates"] > 0
    assert st_a["fallbacks"] == 0 and st_a["underestimates"] == 0


@pytest.mark.parametrize("flags", [abi.FLAG_NO_GRAPH, abi.FLAG_PERSISTENT,
                                   abi.FLAG_PERSISTENT | abi.FLAG_NO_GRAPH])
def test_walk_launch_forms_match(ctx, flags):
    """The walk passes drawn as hipGraphs (default), as direct launches, and
    as persistent kernels on an atomic item queue give the same pixels."""
    sc = workload(2, 40, seed=33)
    n = sc.n_paths
    ref, st_ref = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=0)
    got, st = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=0, flags=flags)
    assert np.array_equal(got, ref)
    assert st["tile_planes"] == st_ref["tile_planes"]
    est, _ = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=1, flags=flags,
                        band_min_edges=-1,
                        edge_estimates=np.full(n, 3000, np.int32),
                        tile_estimates=np.full(n, 60, np.int32))
    assert np.array_equal(est, ref)


def test_single_pass_stroke_outgrows_its_range(ctx):
    """Estimate forms stroke in one pass into ranges sized from the quad-path
    segment count.  A zigzag with round joins needs 238 conic weights where
    its range holds 138: the job reports its exact sizes, is written again
    into an exact range (a fallback), and the pixels equal the exact form's
    and the oracle's."""
    c = Canvas(512, 400)
    p = Path().move_to(10, 20)
    for i in range(120):
        p.line_to(12 + 4 * i, 20 + (340 if i % 2 == 0 else 0))
    zig = Paint((0.8, 0.2, 0.1, 0.9)).set_style(abi.STYLE_STROKE).set_stroke_width(9.0)
    zig.set_stroke_join(abi.JOIN_ROUND)
    c.draw_path(p, zig)
    easy = Paint((0.1, 0.5, 0.9, 0.7)).set_style(abi.STYLE_STROKE).set_stroke_width(5.0)
    c.draw_path(Path().add_circle(250, 200, 90), easy)
    sc = c.scene()
    n = sc.scene.contents.n_paths
    ref, _ = render_gpu(ctx, sc.scene, 512, 400, mode=0)
    got, st = render_gpu(ctx, sc.scene, 512, 400, mode=1,
                         edge_estimates=np.full(n, 50000, np.int32),
                         tile_estimates=np.full(n, 2 * 1024, np.int32))
    print("fallbacks", st["fallbacks"], flush=True)
    assert st["fallbacks"] == 1
    assert np.array_equal(got, ref)
    want = O.render(sc.scene, 512, 400)
    mx, cnt = diff_report(got, want)
    assert mx <= 1 and cnt == 0, f"max diff {mx}, {cnt} pixels"


@pytest.mark.parametrize("cap", [abi.CAP_BUTT, abi.CAP_ROUND, abi.CAP_SQUARE])
@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_ESTIMATES])
def test_stroke_degenerate_segments(ctx, cap, mode):
    """Segments the stroker skips or reduces (near-zero lines, repeated
    points, a folded quad, a cusp cubic, a width far above the curve size):
    the stroke kernel's lanes stroke segments ahead from the quad path's
    points, and the stroker must fall back where its own point differs.
    Exact form (count + write) and the single-pass form of the estimate
    modes."""
    c = Canvas(200, 160)
    for k, width in enumerate((3.0, 40.0)):
        p = Paint((0.9, 0.3, 0.1, 0.7)).set_style(abi.STYLE_STROKE).set_stroke_width(width)
        p.set_stroke_join(k % 3).set_stroke_cap(cap)
        path = (Path().move_to(20, 20).line_to(20.00001, 20.0).quad_to(60, 5, 90, 40)
                .line_to(90, 40).line_to(90.000003, 40.000002).quad_to(95, 45, 90, 40)
                .quad_to(130, 80, 60, 60).cubic_to(140, 140, 20, 140, 100, 60)
                .quad_to(100.0001, 60.0001, 100.0002, 60.0).line_to(150, 150)
                .conic_to(180, 150, 180, 120, 0.3).close())
        c.draw_path(path, p)
        c.draw_path(Path().move_to(5, 150).quad_to(6, 149, 5, 150).line_to(5, 150), p)
    sc = c.scene()
    kw = {}
    if mode == abi.MODE_ESTIMATES:
        n = sc.n_paths
        kw = dict(mode=mode, edge_estimates=np.full(n, 4096, np.int32),
                  tile_estimates=np.full(n, 256, np.int32))
    got, _ = render_gpu(ctx, sc.scene, 200, 160, **kw)
    want = O.render(sc.scene, 200, 160)
    mx, cnt = diff_report(got, want)
    assert mx <= 1 and cnt == 0, f"max diff {mx}, {cnt} pixels"


def test_width_zero_strokes_match_oracle(ctx):
    """Width 0 is stroked as any width (PathStroker at radius 0,
    sw_canvas.cc:388-405, stroke.cc:865-866): open paths draw nothing, closed
    ones leave the slivers of their near-coincident outer and inner
    contours; the device draws what the oracle draws."""
    c = Canvas(200, 160)
    for join in (abi.JOIN_MITER, abi.JOIN_ROUND, abi.JOIN_BEVEL):
        for cap in (abi.CAP_BUTT, abi.CAP_ROUND, abi.CAP_SQUARE):
            p = Paint((0.9, 0.2, 0.1, 1.0)).set_style(abi.STYLE_STROKE).set_stroke_width(0.0)
            p.set_stroke_join(join).set_stroke_cap(cap)
            c.draw_path(Path().move_to(12, 100).line_to(40, 20).line_to(70, 95)
                        .quad_to(100, 10, 140, 60), p)
            c.draw_path(Path().move_to(20, 20).cubic_to(60, -10, 120, 50, 60, 70)
                        .line_to(25, 60).close(), p)
            c.draw_path(Path().add_circle(150, 110, 30), p)
    sc = c.scene()
    presizer.install(ctx, *presizer.load())
    for mode in (abi.MODE_EXACT, abi.MODE_LEARNED):
        got, st = render_gpu(ctx, sc.scene, sc.width, sc.height, mode=mode)
        want = O.render(sc.scene, sc.width, sc.height)
        mx, cnt = diff_report(got, want)
        print(f"width 0, mode {mode}: {int((want[..., 3] > 0).sum())} pixels drawn, max diff "
              f"{mx}, {cnt} differ", flush=True)
        assert mx == 0 and cnt == 0
