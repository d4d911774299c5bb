"""GPU: glyph runs (f4, text as paths; SWCanvas::DrawGlyphsInternal,
sw_canvas.cc:542-639) drawn by the device equal the oracle bit for bit: a
2048^2 page of 14 px text with larger, rotated, stroked, stroke-and-fill and
stroke-then-fill headings and other blend modes, in the exact and learned
forms, and small runs at 6-96 px."""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer
from skity_amd.canvas import Canvas, Paint

from helpers import diff_report
import text_kit
import ttf_kit

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(ttf_kit.find_font() is None, reason="no TrueType font")]


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    presizer.install(c, *presizer.load())
    yield c
    c.close()


@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
def test_text_page_matches_oracle(ctx, mode):
    sc, glyphs = text_kit.text_page(3)
    surf = device.Surface(ctx, 2048, 2048)
    surf.clear(0)
    surf.draw(sc.scene, mode=mode)
    got = surf.read()
    st = surf.stats()
    surf.close()
    want = O.render_mt(sc.scene, 2048, 2048)
    mx, cnt = diff_report(got, want)
    print(f"text page, mode {mode}: {glyphs} glyphs, {sc.scene.contents.n_paths} draws, "
          f"max diff {mx}, {cnt} pixels differ, fallbacks {st['fallbacks']}", flush=True)
    assert mx == 0


def test_glyph_sizes_and_matrices(ctx):
    font = ttf_kit.TrueTypeFont(ttf_kit.find_font())
    cv = Canvas(640, 480)
    y = 10.0
    rng = np.random.default_rng(11)
    for k, size in enumerate([6, 9, 12, 17, 24, 33, 48, 70, 96]):
        paths, xs, ys = font.layout("Ag8@&Wq", float(size), 8.0, y + size)
        p = Paint((rng.random(), rng.random(), rng.random(), 1.0))
        if k % 3 == 1:
            p.set_style(abi.STYLE_STROKE_AND_FILL)
            p.set_stroke_width(1.5)
        cv.save()
        if k % 2:
            cv.translate(300, y)
            cv.rotate(float(rng.uniform(-30, 30)))
            cv.skew(0.2, 0.0)
            cv.translate(-300, -y)
        cv.draw_glyphs(paths, xs, ys, p)
        cv.restore()
        y += size * 0.9
    sc = cv.scene()
    surf = device.Surface(ctx, 640, 480)
    surf.clear(0)
    surf.draw(sc.scene)
    got = surf.read()
    surf.close()
    want = O.render(sc.scene, 640, 480)
    mx, cnt = diff_report(got, want)
    print(f"glyph sizes 6-96 px: max diff {mx}, {cnt} pixels differ", flush=True)
    assert mx == 0
