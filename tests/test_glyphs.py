"""Glyph runs (f4, text as paths) on the canvas mirror and the C++ front
end's rules, on the CPU: the draw list DrawGlyphsInternal makes
(sw_canvas.cc:542-639) and the oracle's drawing of it."""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi
from skity_amd.canvas import Canvas, Paint, Path

import ttf_kit

FONT = ttf_kit.find_font()
pytestmark = pytest.mark.skipif(FONT is None, reason="no TrueType font installed")


def _two_glyphs():
    a = Path().add_rect(0, -10, 8, 0)
    b = Path().move_to(0, 0).line_to(6, -12).line_to(12, 0).close()
    return [a, b], [5.0, 9.0], [20.0, 21.5]


def test_glyph_run_order_and_matrices():
    paths, xs, ys = _two_glyphs()
    cv = Canvas(64, 64)
    cv.translate(3, 2)
    cv.rotate(10)
    p = Paint((0.2, 0.4, 0.6, 1)).set_style(abi.STYLE_STROKE_AND_FILL)
    cv.draw_glyphs(paths, xs, ys, p)
    styles = [d[1].p.style for d in cv.draws]
    # all fills, then all strokes (paint_order.hpp:12-31)
    assert styles == [abi.STYLE_FILL, abi.STYLE_FILL, abi.STYLE_STROKE, abi.STYLE_STROKE]
    assert cv.draws[0][0] is paths[0] and cv.draws[2][0] is paths[0]
    # each glyph under the canvas matrix times Translate(x, y)
    ref = Canvas(64, 64)
    ref.translate(3, 2)
    ref.rotate(10)
    ref.translate(xs[1], ys[1])
    assert cv.draws[1][2] == ref.get_total_matrix()
    cv2 = Canvas(64, 64)
    cv2.draw_glyphs(paths, xs, ys, Paint((0, 0, 0, 1)).set_style(abi.STYLE_STROKE_THEN_FILL))
    assert [d[1].p.style for d in cv2.draws] == [abi.STYLE_STROKE] * 2 + [abi.STYLE_FILL] * 2
    # the caller's paint is not changed
    assert p.p.style == abi.STYLE_STROKE_AND_FILL


def test_glyph_runs_with_filters_or_shaders_are_refused():
    paths, xs, ys = _two_glyphs()
    cv = Canvas(64, 64)
    with pytest.raises(NotImplementedError):
        cv.draw_glyphs(paths, xs, ys, Paint((0, 0, 0, 1)).set_blur(2.0))


def test_text_coverage_matches_outline_area():
    """A filled text line: the oracle's total coverage equals the area the
    glyph outlines enclose (quads flattened finely, nonzero winding), within
    the antialiasing error along the outline."""
    font = ttf_kit.TrueTypeFont(FONT)
    paths, xs, ys = font.layout("Bold glyph 8 @", 48.0, 10.0, 70.0)
    cv = Canvas(420, 100)
    cv.draw_glyphs(paths, xs, ys, Paint((0, 0, 0, 1)))
    sc = cv.scene()
    img = O.render(sc.scene, 420, 100)
    cov = img[..., 3].astype(np.float64).sum() / 255.0
    # exact area by supersampling the nonzero winding of the flattened outline
    area, perim = 0.0, 0.0
    for p, x0, y0 in zip(paths, xs, ys):
        polys, cur, i = [], [], 0
        for v in p.verbs:
            if v == abi.VERB_MOVE:
                if cur:
                    polys.append(cur)
                cur = [p.points[i]]
                i += 1
            elif v == abi.VERB_LINE:
                cur.append(p.points[i])
                i += 1
            elif v == abi.VERB_QUAD:
                a, c, b = np.array(cur[-1]), np.array(p.points[i]), np.array(p.points[i + 1])
                for t in np.linspace(0, 1, 17)[1:]:
                    cur.append(tuple((1 - t) ** 2 * a + 2 * t * (1 - t) * c + t * t * b))
                i += 2
        if cur:
            polys.append(cur)
        for poly in polys:
            q = np.array(poly) + (x0, y0)
            area += 0.5 * np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1])
            perim += np.sum(np.hypot(*(np.roll(q, -1, 0) - q).T))
    # TrueType outer contours run clockwise in y-up units: positive here (y down)
    assert abs(cov - abs(area)) < 0.03 * perim, (cov, area, perim)
