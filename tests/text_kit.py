"""Test scenes of glyph runs (text as paths) on the canvas mirror."""
import numpy as np

from skity_amd import abi
from skity_amd.canvas import Canvas, Paint

import ttf_kit

WORDS = ("the quick brown fox jumps over a lazy dog while glyph runs fill and stroke "
         "every outline 0123456789 ABCDEFGHIJKLMNOPQRSTUVWXYZ @#&%?!").split()


def text_page(seed, w=2048, h=2048, size=14.0, lines=None, styles=True):
    """A page of text lines (size px, DejaVu Sans) with headings that are
    larger, rotated, stroked, stroke-and-fill and stroke-then-fill runs, some
    under a colour filter or another blend mode.  None when no font is
    installed."""
    path = ttf_kit.find_font()
    if path is None:
        return None
    font = ttf_kit.TrueTypeFont(path)
    rng = np.random.default_rng(seed)
    cv = Canvas(w, h)
    cache = {}
    n_lines = lines if lines is not None else int((h - 40) / (1.4 * size))
    y = 20 + size
    glyphs = 0
    for i in range(n_lines):
        text = " ".join(rng.choice(WORDS, int(w / (size * 4))).tolist())
        sz = size if (i % 9 or not styles) else size * float(rng.uniform(2, 3.5))
        key = (text, sz)
        if key not in cache:
            cache[key] = font.layout(text, sz, 0.0, 0.0)
        paths, xs, ys = cache[key]
        xs = [x + 12.0 for x in xs]
        ys = [yy + y for yy in ys]
        paint = Paint((rng.random() * 0.6, rng.random() * 0.6, rng.random() * 0.6,
                       float(rng.uniform(0.6, 1.0))))
        if styles and i % 9 == 0:
            paint.set_style(int(rng.choice([abi.STYLE_STROKE, abi.STYLE_STROKE_AND_FILL,
                                            abi.STYLE_STROKE_THEN_FILL])))
            paint.set_stroke_width(float(rng.uniform(0.5, 2.5)))
            paint.set_stroke_join(int(rng.integers(0, 3)))
            paint.set_fill_color((0.9, 0.8, 0.2, 0.9))
        if styles and i % 7 == 3:
            paint.set_blend_mode(int(rng.choice([abi.BLEND_MULTIPLY, abi.BLEND_SCREEN,
                                                 abi.BLEND_XOR])))
        cv.save()
        if styles and i % 11 == 5:
            cv.translate(w / 2, y)
            cv.rotate(float(rng.uniform(-8, 8)))
            cv.scale(float(rng.uniform(0.8, 1.2)), 1.0)
            cv.translate(-w / 2, -y)
        cv.draw_glyphs(paths, xs, ys, paint)
        cv.restore()
        glyphs += len(paths)
        y += 1.4 * sz
        if y > h - 10:
            break
    return cv.scene(), glyphs
