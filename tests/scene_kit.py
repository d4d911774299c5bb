"""Small seeded scenes exercising the canvas matrix, DrawPaint, blend modes,
target formats and blur styles (shared by the oracle pins and the GPU
parity tests)."""
import math

import numpy as np

from skity_amd import abi, canvas
from skity_amd.canvas import Canvas, Paint, Path

BLEND_MODES = list(range(abi.BLEND_SCREEN + 1)) + [abi.BLEND_SOFT_LIGHT, 16, 28]


def random_path(rng, cx, cy, size, kind="mixed"):
    p = Path()
    p.move_to(cx + rng.uniform(-size, size), cy + rng.uniform(-size, size))
    for _ in range(int(rng.integers(3, 9))):
        u = rng.random() if kind == "mixed" else 0.0
        pt = lambda: (cx + rng.uniform(-size, size), cy + rng.uniform(-size, size))
        if u < 0.35:
            p.line_to(*pt())
        elif u < 0.6:
            p.quad_to(*pt(), *pt())
        elif u < 0.8:
            p.cubic_to(*pt(), *pt(), *pt())
        else:
            p.conic_to(*pt(), *pt(), rng.uniform(0.3, 3.0))
    p.close()
    if rng.random() < 0.5:
        p.fill_type = abi.FILL_EVEN_ODD
    return p


def random_gradient(rng, cx, cy, size):
    g = abi.Gradient()
    g.type = int(rng.choice([abi.GRADIENT_LINEAR, abi.GRADIENT_RADIAL, abi.GRADIENT_SWEEP]))
    g.tile_mode = int(rng.integers(0, 3))
    g.color_count = int(rng.integers(2, 6))
    g.has_offsets = 1
    offs = np.sort(rng.random(g.color_count))
    for i in range(g.color_count):
        g.offsets[i] = float(offs[i])
        for c in range(4):
            g.colors[i][c] = float(rng.random())
    g.points[0], g.points[1] = cx - size / 2, cy - size / 3
    g.points[2], g.points[3] = cx + size / 2, cy + size / 3
    if g.type == abi.GRADIENT_RADIAL:
        g.radius[0] = size
    elif g.type == abi.GRADIENT_SWEEP:
        start, end = 0.0, float(rng.uniform(90, 360))
        g.radius[0] = -start / 360.0
        g.radius[1] = 1.0 / ((end - start) / 360.0)
    g.local_matrix[0] = 1.0
    g.local_matrix[4] = 1.0
    return g


def random_paint(rng, size, cx, cy, allow_blur=True, blend=None, gradient=0.3):
    paint = Paint((rng.random(), rng.random(), rng.random(), rng.uniform(0.3, 1.0)))
    r = rng.random()
    if r < 0.25:
        paint.set_style(abi.STYLE_STROKE)
    elif r < 0.35:
        paint.set_style(abi.STYLE_STROKE_AND_FILL)
    paint.set_stroke_width(float(rng.uniform(1, 8)))
    paint.set_stroke_join(int(rng.integers(0, 3)))
    paint.set_stroke_cap(int(rng.integers(0, 3)))
    if rng.random() < gradient:
        paint.set_gradient(random_gradient(rng, cx, cy, size))
    if allow_blur and rng.random() < 0.2:
        paint.set_blur(float(rng.uniform(1.5, 8.0)), int(rng.integers(1, 5)))
    if blend is not None:
        paint.set_blend_mode(blend)
    return paint


def matrix_scene(seed, w=320, h=240, n=24):
    """Draws under rotations, scales, skews and translations (Save/Restore
    around each), with fills, strokes, gradients and blurred draws."""
    cv = Canvas(w, h)
    for path, paint, m, kind in matrix_scene_draws(seed, w, h, n):
        cv.save()
        cv.concat(m)
        cv.draw_path(path, paint)
        cv.restore()
    return cv.scene()


def matrix_scene_draws(seed, w=320, h=240, n=24):
    """(path, paint, matrix, kind) of matrix_scene's draws."""
    rng = np.random.default_rng(seed)
    cv = Canvas(w, h)
    for i in range(n):
        cv.save()
        cx, cy = rng.uniform(40, w - 40), rng.uniform(40, h - 40)
        cv.translate(cx, cy)
        k = i % 4
        if k == 0:
            cv.rotate(float(rng.uniform(0, 360)))
        elif k == 1:
            cv.scale(float(rng.uniform(0.4, 2.5)), float(rng.uniform(0.4, 2.5)))
        elif k == 2:
            cv.skew(float(rng.uniform(-0.6, 0.6)), float(rng.uniform(-0.6, 0.6)))
        else:
            cv.rotate(float(rng.uniform(0, 360)))
            cv.scale(float(rng.uniform(0.5, 2.0)), float(rng.uniform(0.5, 2.0)))
        size = float(rng.uniform(10, 40))
        cv.draw_path(random_path(rng, 0.0, 0.0, size), random_paint(rng, size, 0.0, 0.0))
        cv.restore()
    return cv.draws


def paint_blend_scene(seed, w=96, h=72):
    """A backdrop, then DrawPaint and DrawPath calls in every blend mode the
    software backend blends (and modes that fall back to kSrcOver)."""
    rng = np.random.default_rng(seed)
    cv = Canvas(w, h)
    cv.draw_paint(Paint((0.2, 0.4, 0.6, 0.5)))
    for mode in BLEND_MODES:
        size = float(rng.uniform(8, 30))
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        paint = random_paint(rng, size, cx, cy, allow_blur=False, blend=mode)
        cv.draw_path(random_path(rng, cx, cy, size), paint)
    cv.save()
    cv.rotate(20, w / 2, h / 2)
    g = Paint((1, 1, 1, 1)).set_gradient(random_gradient(rng, w / 2, h / 2, 40))
    g.set_blend_mode(abi.BLEND_SRC_ATOP)
    cv.draw_paint(g)
    cv.restore()
    cv.draw_paint(Paint((0.9, 0.1, 0.1, 0.25)).set_blend_mode(abi.BLEND_DST_OVER))
    return cv.scene()


def conical_gradient(rng, cx, cy, size, case):
    """A two-point conical gradient around (cx, cy): nested circles, a cone of
    separate circles, an end radius of 0, equal radii (strip) or concentric
    circles, with random stops, tiling and local rotation."""
    g = abi.Gradient()
    g.type = abi.GRADIENT_CONICAL
    g.tile_mode = int(rng.integers(0, 4))
    g.color_count = int(rng.integers(2, 6))
    g.has_offsets = 1
    offs = np.sort(rng.random(g.color_count))
    for i in range(g.color_count):
        g.offsets[i] = float(offs[i])
        for c in range(4):
            g.colors[i][c] = float(rng.uniform(0.2, 1.0))
    s = size
    c0 = (cx - 0.3 * s, cy - 0.1 * s)
    if case == "nested":
        c1, r0, r1 = (cx + 0.2 * s, cy + 0.1 * s), 0.2 * s, 0.9 * s
    elif case == "cone":
        c0, c1, r0, r1 = (cx - 0.6 * s, cy), (cx + 0.6 * s, cy + 0.2 * s), 0.1 * s, 0.35 * s
    elif case == "to_point":
        c1, r0, r1 = (cx + 0.25 * s, cy + 0.15 * s), 0.8 * s, 0.0
    elif case == "strip":
        c1, r0, r1 = (cx + 0.5 * s, cy + 0.2 * s), 0.3 * s, 0.3 * s
    else:
        c1, r0, r1 = c0, 0.15 * s, 0.8 * s
    g.points[0], g.points[1], g.points[2], g.points[3] = c0[0], c0[1], c1[0], c1[1]
    g.radius[0], g.radius[1] = r0, r1
    a = float(rng.uniform(-0.4, 0.4))
    ca, sa = float(np.cos(a)), float(np.sin(a))
    # rotation about (cx, cy) as the shader's local matrix
    m = (ca, -sa, cx - ca * cx + sa * cy, sa, ca, cy - sa * cx - ca * cy)
    for i, v in enumerate(m):
        g.local_matrix[i] = v
    return g


CONICAL_CASES = ("nested", "cone", "to_point", "strip", "concentric")


def conical_scene(seed, w=256, h=192, n=30):
    """Fills and strokes with two-point conical gradients of every case,
    some under a canvas matrix."""
    rng = np.random.default_rng(seed)
    cv = Canvas(w, h)
    for i in range(n):
        cx, cy = float(rng.uniform(30, w - 30)), float(rng.uniform(30, h - 30))
        size = float(rng.uniform(20, 70))
        paint = Paint((1, 1, 1, float(rng.uniform(0.5, 1.0))))
        if i % 5 == 4:
            paint.set_style(abi.STYLE_STROKE)
            paint.set_stroke_width(float(rng.uniform(2, 9)))
        paint.set_gradient(conical_gradient(rng, cx, cy, size, CONICAL_CASES[i % 5]))
        cv.save()
        if i % 3 == 2:
            cv.translate(cx, cy)
            cv.rotate(float(rng.uniform(-30, 30)))
            cv.translate(-cx, -cy)
        cv.draw_path(random_path(rng, cx, cy, size), paint)
        cv.restore()
    return cv.scene()


def random_color_filter(rng):
    """One of the ColorFilters factories with random arguments (never None)."""
    k = int(rng.integers(0, 5))
    if k == 0:
        argb = (int(rng.integers(40, 256)) << 24) | int(rng.integers(0, 1 << 24))
        mode = int(rng.choice([abi.BLEND_SRC, abi.BLEND_SRC_OVER, abi.BLEND_SRC_IN,
                               abi.BLEND_DST_OVER, abi.BLEND_MODULATE, abi.BLEND_SCREEN,
                               abi.BLEND_XOR, abi.BLEND_PLUS, abi.BLEND_SRC_ATOP]))
        return canvas.blend_color_filter(argb, mode)
    if k == 1:
        m = rng.uniform(-0.5, 1.2, 20)
        m[[4, 9, 14, 19]] = rng.uniform(-0.3, 0.3, 4)
        m[18] = rng.uniform(0.6, 1.0)
        return canvas.matrix_color_filter([float(v) for v in m])
    if k == 2:
        return canvas.gamma_color_filter(True)
    if k == 3:
        return canvas.gamma_color_filter(False)
    return canvas.compose_color_filter(canvas.gamma_color_filter(True),
                                       canvas.gamma_color_filter(False))


def filter_scene(seed, w=200, h=150, n=36):
    """Fills, strokes, gradients, blurred draws and DrawPaint, each with a
    random colour filter and some with other blend modes."""
    rng = np.random.default_rng(seed)
    cv = Canvas(w, h)
    cv.draw_paint(Paint((0.3, 0.6, 0.9, 0.8)))
    for i in range(n):
        cx, cy = float(rng.uniform(20, w - 20)), float(rng.uniform(20, h - 20))
        size = float(rng.uniform(8, 40))
        blend = None
        if i % 4 == 3:
            blend = int(rng.choice([abi.BLEND_SRC, abi.BLEND_DST_IN, abi.BLEND_XOR,
                                    24, abi.BLEND_SOFT_LIGHT]))
        paint = random_paint(rng, size, cx, cy, allow_blur=True, blend=blend)
        if i % 6 != 5:
            paint.set_color_filter(random_color_filter(rng))
        cv.draw_path(random_path(rng, cx, cy, size), paint)
    return cv.scene()


def random_image(rng, w, h, alpha_type):
    """An RGBA8 image: premultiplied (each channel at most alpha) or
    unpremultiplied, with some opaque and some clear texels."""
    img = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    a = img[..., 3]
    a[rng.random((h, w)) < 0.3] = 255
    a[rng.random((h, w)) < 0.1] = 0
    if alpha_type == abi.ALPHA_PREMUL:
        img[..., :3] = (img[..., :3].astype(np.uint32) * a[..., None] // 255).astype(np.uint8)
    return img


def random_local_matrix(rng, cx, cy):
    """A local matrix that scales (0.25-4x), rotates and translates the
    image to near (cx, cy)."""
    s = float(rng.uniform(0.25, 4.0))
    th = float(rng.uniform(-np.pi, np.pi)) if rng.random() < 0.5 else 0.0
    c, sn = s * np.cos(th), s * np.sin(th)
    return (c, -sn, cx + float(rng.uniform(-20, 20)), sn, c, cy + float(rng.uniform(-20, 20)))


def pixmap_scene(seed, w=240, h=180, n=30):
    """Fills, strokes, blurred draws and a DrawPaint with pixmap shaders:
    nearest and linear sampling, every pair of tile modes, local matrices
    that scale and rotate, some draws under a canvas matrix, premultiplied
    and unpremultiplied images (a few shared between draws), some with a
    colour filter or another blend mode."""
    rng = np.random.default_rng(seed)
    cv = Canvas(w, h)
    pool = []
    for k in range(4):
        at = abi.ALPHA_PREMUL if k % 2 == 0 else abi.ALPHA_UNPREMUL
        iw, ih = int(rng.integers(1, 40)), int(rng.integers(1, 40))
        pool.append((random_image(rng, iw, ih, at), at))
    bg = Paint((1, 1, 1, 0.7)).set_image_shader(
        pool[0][0], pool[0][1], abi.FILTER_LINEAR, abi.TILE_REPEAT, abi.TILE_MIRROR,
        (2.0, 0.3, 5.0, -0.2, 1.5, 3.0))
    cv.draw_paint(bg)
    for i in range(n):
        cx, cy = float(rng.uniform(20, w - 20)), float(rng.uniform(20, h - 20))
        size = float(rng.uniform(10, 50))
        blend = None
        if i % 7 == 6:
            blend = int(rng.choice([abi.BLEND_SRC, abi.BLEND_DST_IN, abi.BLEND_XOR, 24]))
        paint = random_paint(rng, size, cx, cy, allow_blur=True, blend=blend, gradient=0.0)
        if i % 3 == 0:
            img, at = pool[int(rng.integers(0, len(pool)))]
        else:
            at = abi.ALPHA_PREMUL if rng.random() < 0.5 else abi.ALPHA_UNPREMUL
            img = random_image(rng, int(rng.integers(1, 48)), int(rng.integers(1, 48)), at)
        paint.set_image_shader(img, at, int(i % 2), int(rng.integers(0, 4)),
                               int(rng.integers(0, 4)), random_local_matrix(rng, cx, cy))
        if i % 5 == 4:
            paint.set_color_filter(random_color_filter(rng))
        cv.save()
        if i % 4 == 1:
            cv.translate(cx, cy)
            cv.rotate(float(rng.uniform(-40, 40)))
            cv.scale(float(rng.uniform(0.6, 1.4)), float(rng.uniform(0.6, 1.4)))
            cv.translate(-cx, -cy)
        cv.draw_path(random_path(rng, cx, cy, size), paint)
        cv.restore()
    return cv.scene()


def clip_scene(seed, w=256, h=192, n=40):
    """Draws under nested Save / ClipRect / Restore: integer, fractional,
    partly off-canvas and empty clips, under scale + translate matrices;
    fills, strokes, gradients, blurred draws, pixmap shaders, DrawPaint
    (which ignores the clip) and other blend modes."""
    rng = np.random.default_rng(seed)
    cv = Canvas(w, h)
    depth = 0
    img = random_image(rng, 17, 11, abi.ALPHA_PREMUL)
    for i in range(n):
        r = rng.random()
        if r < 0.2 and depth < 4:
            cv.save()
            depth += 1
            if rng.random() < 0.5:
                cv.translate(float(rng.uniform(-20, 20)), float(rng.uniform(-20, 20)))
                cv.scale(float(rng.uniform(0.6, 1.6)), float(rng.uniform(0.6, 1.6)))
            l, t = float(rng.uniform(-30, w - 20)), float(rng.uniform(-30, h - 20))
            if rng.random() < 0.5:
                l, t = float(round(l)), float(round(t))
            cv.clip_rect(l, t, l + float(rng.uniform(0, 160)), t + float(rng.uniform(0, 120)))
            continue
        if r < 0.3 and depth > 0:
            cv.restore()
            depth -= 1
            continue
        cx, cy = float(rng.uniform(0, w)), float(rng.uniform(0, h))
        size = float(rng.uniform(10, 60))
        blend = None
        if i % 9 == 8:
            blend = int(rng.choice([abi.BLEND_SRC, abi.BLEND_XOR, abi.BLEND_MULTIPLY]))
        paint = random_paint(rng, size, cx, cy, allow_blur=True, blend=blend)
        if i % 7 == 3 and paint.gradient is None:
            paint.set_image_shader(img, abi.ALPHA_PREMUL, abi.FILTER_LINEAR, abi.TILE_REPEAT,
                                   abi.TILE_MIRROR, (1.5, 0.0, cx, 0.0, 1.5, cy))
        if i % 13 == 12:
            cv.draw_paint(Paint((rng.random(), rng.random(), rng.random(), 0.3)))
        else:
            cv.draw_path(random_path(rng, cx, cy, size), paint)
    return cv.scene()
