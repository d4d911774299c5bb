"""Python mirror of skity's Path / Paint / Canvas::DrawPath recording API.

A ``Canvas`` records DrawPath calls (include/skity/render/canvas.hpp:
Canvas::DrawPath) into the hipr_scene layout of include/skity_hip.h and hands
them to the HIP library in one call.  ``Path`` stores verbs and points the way
skity's Path does (src/graphic/path.cc:494-560, MoveTo/LineTo/QuadTo/ConicTo/
CubicTo with InjectMoveToIfNeed, :837-859 Close, :815-835 AddRect,
:1001-1035 AddCircle/AddOval), so the arrays handed over are exactly the ones
skity would iterate.
"""
import ctypes
import copy
import math

import numpy as np

from . import abi

FLOAT_ROOT2_OVER2 = 0.707106781


class Path:
    """Verb / point / conic-weight storage with skity's builder rules."""

    def __init__(self):
        self.verbs = []
        self.points = []
        self.weights = []
        self.fill_type = abi.FILL_WINDING
        self._last_move = -1  # index of the last move point; < 0 after Close

    # -- builder (src/graphic/path.cc) --
    def move_to(self, x, y):
        if self.verbs and self.verbs[-1] == abi.VERB_MOVE:
            self.points[-1] = (float(x), float(y))
        else:
            self._last_move = len(self.points)
            self.verbs.append(abi.VERB_MOVE)
            self.points.append((float(x), float(y)))
        return self

    def _inject_move(self):
        if self._last_move < 0:
            if not self.points:
                x, y = 0.0, 0.0
            else:
                x, y = self.points[~self._last_move]
            self.move_to(x, y)

    def line_to(self, x, y):
        self._inject_move()
        self.verbs.append(abi.VERB_LINE)
        self.points.append((float(x), float(y)))
        return self

    def quad_to(self, x1, y1, x2, y2):
        self._inject_move()
        self.verbs.append(abi.VERB_QUAD)
        self.points += [(float(x1), float(y1)), (float(x2), float(y2))]
        return self

    def conic_to(self, x1, y1, x2, y2, w):
        w = np.float32(w)
        if not (w > 0):
            return self.line_to(x2, y2)
        if math.isinf(w):
            self.line_to(x1, y1)
            return self.line_to(x2, y2)
        if w == 1:
            return self.quad_to(x1, y1, x2, y2)
        self._inject_move()
        self.verbs.append(abi.VERB_CONIC)
        self.weights.append(float(w))
        self.points += [(float(x1), float(y1)), (float(x2), float(y2))]
        return self

    def cubic_to(self, x1, y1, x2, y2, x3, y3):
        self._inject_move()
        self.verbs.append(abi.VERB_CUBIC)
        self.points += [(float(x1), float(y1)), (float(x2), float(y2)),
                        (float(x3), float(y3))]
        return self

    def close(self):
        if self.verbs and self.verbs[-1] != abi.VERB_CLOSE:
            self.verbs.append(abi.VERB_CLOSE)
        if self._last_move >= 0:
            self._last_move = ~self._last_move
        return self

    def add_rect(self, l, t, r, b, ccw=False, start=0):
        pts = [(l, t), (r, t), (r, b), (l, b)]
        adv = 3 if ccw else 1
        i = start % 4
        self.move_to(*pts[i])
        for _ in range(3):
            i = (i + adv) % 4
            self.line_to(*pts[i])
        return self.close()

    def add_oval(self, l, t, r, b, ccw=False, start=1):
        l, t, r, b = (np.float32(v) for v in (l, t, r, b))
        cx = np.float32(np.float32(0.5) * np.float32(r + l))  # Rect::CenterX, rect.cc:52
        cy = np.float32(np.float32(0.5) * np.float32(t + b))
        oval = [(cx, t), (r, cy), (cx, b), (l, cy)]
        rect = [(l, t), (r, t), (r, b), (l, b)]
        adv = 3 if ccw else 1
        oi = start % 4
        ri = (start + (0 if not ccw else 1)) % 4
        self.move_to(*oval[oi])
        for _ in range(4):
            ri = (ri + adv) % 4
            oi = (oi + adv) % 4
            self.conic_to(*rect[ri], *oval[oi], FLOAT_ROOT2_OVER2)
        return self.close()

    def add_circle(self, x, y, radius, ccw=False):
        if radius > 0:
            self.add_oval(x - radius, y - radius, x + radius, y + radius, ccw)
        return self


F32 = np.float32
IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0)  # scale_x, skew_x, trans_x, skew_y, scale_y, trans_y


def mat_mul(a, b):
    """Matrix a * b (src/geometry/matrix.cc:450-458): glm's 4x4 product
    restricted to affine operands, every float operation rounded to f32."""
    if tuple(a) == IDENTITY:
        return tuple(b)
    if tuple(b) == IDENTITY:
        return tuple(a)
    asx, akx, atx, aky, asy, aty = (F32(v) for v in a)
    bsx, bkx, btx, bky, bsy, bty = (F32(v) for v in b)
    return (float(asx * bsx + akx * bky), float(asx * bkx + akx * bsy),
            float((asx * btx + akx * bty) + atx), float(aky * bsx + asy * bky),
            float(aky * bkx + asy * bsy), float((aky * btx + asy * bty) + aty))


def map_rect(m, rect):
    """Matrix::MapRect (matrix.cc:306-337) for an affine m: the corners
    mapped as glm's mat4 * vec4 sums them, (sx x + kx y) + (0 + tx), and
    their min / max."""
    sx, kx, tx, ky, sy, ty = (F32(v) for v in m)
    l, t, r, b = (F32(v) for v in rect)
    xs, ys = [], []
    for x, y in ((l, t), (r, t), (r, b), (l, b)):
        xs.append((sx * x + kx * y) + (F32(0) + tx))
        ys.append((ky * x + sy * y) + (F32(0) + ty))
    return min(xs), min(ys), max(xs), max(ys)


def mat_rotate(degrees, px=0.0, py=0.0):
    """Rotation about (px, py) (Matrix::RotateDeg)."""
    r = math.radians(degrees)
    c, s = float(F32(math.cos(r))), float(F32(math.sin(r)))
    rot = (c, -s, 0.0, s, c, 0.0)
    if px == 0 and py == 0:
        return rot
    return mat_mul(mat_mul((1.0, 0.0, px, 0.0, 1.0, py), rot), (1.0, 0.0, -px, 0.0, 1.0, -py))


class Paint:
    """The Paint fields the DrawPath path reads (include/skity/graphic/paint.hpp)."""

    def __init__(self, color=(0, 0, 0, 1)):
        self.p = abi.Paint()
        self.set_color(color)
        self.p.style = abi.STYLE_FILL
        self.p.stroke_width = 1.0
        self.p.miter_limit = 4.0
        self.p.cap = abi.CAP_BUTT
        self.p.join = abi.JOIN_MITER
        self.p.gradient = -1
        self.p.blend = abi.BLEND_SRC_OVER
        self.gradient = None
        self.color_filter = None
        self.image_shader = None

    def set_color(self, rgba):
        # Paint::SetColor sets both the fill and the stroke colour
        for i in range(4):
            self.p.color[i] = rgba[i]
            self.p.stroke_color[i] = rgba[i]
        return self

    def set_fill_color(self, rgba):
        for i in range(4):
            self.p.color[i] = rgba[i]
        return self

    def set_stroke_color(self, rgba):
        for i in range(4):
            self.p.stroke_color[i] = rgba[i]
        return self

    def set_style(self, style):
        self.p.style = style
        return self

    def set_stroke_width(self, w):
        self.p.stroke_width = w
        return self

    def set_stroke_cap(self, cap):
        self.p.cap = cap
        return self

    def set_stroke_join(self, join):
        self.p.join = join
        return self

    def set_stroke_miter(self, m):
        self.p.miter_limit = m
        return self

    def set_blur(self, radius, style=abi.BLUR_NORMAL):
        """MaskFilter::MakeBlur(style, radius)."""
        self.p.blur_radius = radius
        self.p.blur_style = style if radius > 0 else 0
        return self

    def set_blend_mode(self, mode):
        self.p.blend = mode
        return self

    def set_gradient(self, g):
        """A Shader::MakeLinear/Radial/Sweep/TwoPointConical gradient
        (abi.Gradient)."""
        self.gradient = g
        return self

    def set_image_shader(self, pixels, alpha_type=abi.ALPHA_PREMUL, filter=abi.FILTER_NEAREST,
                         tile_x=abi.TILE_CLAMP, tile_y=abi.TILE_CLAMP,
                        local_matrix=(1, 0, 0, 0, 1, 0)):
        """Shader::MakeShader(Image::MakeImage(pixmap), sampling, tile_x,
        tile_y, local_matrix): pixels is an (H, W, 4) uint8 RGBA array."""
        self.image_shader = (np.ascontiguousarray(pixels, np.uint8), alpha_type, filter,
                             tile_x, tile_y, tuple(float(v) for v in local_matrix))
        return self

    def set_color_filter(self, cf):
        """Paint::SetColorFilter: an abi.ColorFilter from the ColorFilters
        factories below, or None."""
        self.color_filter = cf
        return self


# ColorFilters (include/skity/effect/color_filter.hpp:26-37,
# src/effect/color_filter.cc:64-106): the factories return None where the
# reference returns no filter.
def blend_color_filter(argb, mode):
    a = (argb >> 24) & 255
    if (mode == abi.BLEND_DST or (mode == abi.BLEND_DST_IN and a == 255)
            or (mode == abi.BLEND_DST_OUT and a == 0)):
        return None
    cf = abi.ColorFilter()
    cf.type, cf.color, cf.blend = abi.CF_BLEND, argb, mode
    return cf


def matrix_color_filter(row_major):
    ident = (1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0)
    if row_major is None or tuple(float(v) for v in row_major) == ident:
        return None
    cf = abi.ColorFilter()
    cf.type = abi.CF_MATRIX
    for i, v in enumerate(row_major):
        cf.matrix[i] = v
    return cf


def gamma_color_filter(linear_to_srgb):
    cf = abi.ColorFilter()
    cf.type = abi.CF_LINEAR_TO_SRGB if linear_to_srgb else abi.CF_SRGB_TO_LINEAR
    return cf


def compose_color_filter(outer, inner):
    if inner is None:
        return outer
    if outer is None:
        return inner
    cf = abi.ColorFilter()
    cf.type = abi.CF_COMPOSE
    return cf


class Canvas:
    """Records DrawPath / DrawPaint calls with the canvas matrix of each
    (Canvas::Save/Restore/Translate/Scale/Rotate/Skew/Concat/SetMatrix,
    src/render/canvas.cc:70-125, pre-concatenated as in
    src/render/canvas_state.cc:18-45); ``scene()`` packs them as a
    hipr_scene."""

    def __init__(self, width, height):
        self.width, self.height = width, height
        self.draws = []
        self._stack = [IDENTITY]
        self._clips = [abi.CULL_RECT]  # global clip bounds per save level

    # -- matrix and clip state --
    def save(self):
        self._stack.append(self._stack[-1])
        self._clips.append(self._clips[-1])
        return len(self._stack) - 1

    def restore(self):
        if len(self._stack) > 1:
            self._stack.pop()
            self._clips.pop()

    def clip_rect(self, l, t, r, b, difference=False):
        """Canvas::ClipRect (canvas.cc:130-135,463-476, sw_canvas.cc:297-313):
        an intersect clip under a scale + translate matrix narrows the global
        clip bounds, the scan clip of later path draws; a difference clip or
        one under a rotating / skewing matrix clips span lists, which the
        backend does not do."""
        m = self._stack[-1]
        if difference or m[1] != 0 or m[3] != 0:
            raise NotImplementedError("span-list clips (difference, or under skew/rotation)")
        ml, mt, mr, mb = map_rect(m, (l, t, r, b))
        cl, ct, cr, cb = self._clips[-1]
        il, ir = max(F32(cl), F32(ml)), min(F32(cr), F32(mr))
        it, ib = max(F32(ct), F32(mt)), min(F32(cb), F32(mb))
        if il < ir and it < ib:
            self._clips[-1] = (float(il), float(it), float(ir), float(ib))
        else:
            self._clips[-1] = (0.0, 0.0, 0.0, 0.0)  # Rect::SetEmpty

    def get_clip_bounds(self):
        return self._clips[-1]

    def get_total_matrix(self):
        return self._stack[-1]

    def concat(self, m):
        self._stack[-1] = mat_mul(self._stack[-1], m)

    def set_matrix(self, m):
        self._stack[-1] = tuple(float(F32(v)) for v in m)

    def reset_matrix(self):
        self._stack[-1] = IDENTITY

    def translate(self, dx, dy):
        self.concat((1.0, 0.0, dx, 0.0, 1.0, dy))

    def scale(self, sx, sy):
        self.concat((sx, 0.0, 0.0, 0.0, sy, 0.0))

    def rotate(self, degrees, px=0.0, py=0.0):
        self.concat(mat_rotate(degrees, px, py))

    def skew(self, kx, ky):
        self.concat((1.0, kx, 0.0, ky, 1.0, 0.0))

    # -- draws --
    def draw_path(self, path, paint):
        # (path, paint, matrix, kind[, clip bounds when something clips])
        clip = self._clips[-1]
        d = (path, paint, self._stack[-1], abi.DRAW_PATH)
        self.draws.append(d if clip == abi.CULL_RECT else d + (clip,))

    def draw_paint(self, paint):
        self.draws.append((Path(), paint, self._stack[-1], abi.DRAW_PAINT))

    def draw_glyphs(self, glyph_paths, xs, ys, paint):
        """Canvas::DrawGlyphs as the software canvas draws a glyph run
        (sw_canvas.cc:542-639): glyph k's outline under the canvas matrix
        times Translate(xs[k], ys[k]), all fills then all strokes (the
        reverse for kStrokeThenFill).  Runs with a mask filter or a shader
        are not drawn by the backend."""
        if paint.p.blur_radius > 0 or paint.gradient is not None or paint.image_shader:
            raise NotImplementedError("glyph runs with a mask filter or a shader")
        style = paint.p.style
        need_fill = style != abi.STYLE_STROKE
        need_stroke = style != abi.STYLE_FILL

        def run(st):
            p = copy.copy(paint)
            p.p = abi.Paint.from_buffer_copy(paint.p)
            p.p.style = st
            for path, x, y in zip(glyph_paths, xs, ys):
                m = mat_mul(self._stack[-1], (1.0, 0.0, float(F32(x)), 0.0, 1.0, float(F32(y))))
                self.draws.append((path, p, m, abi.DRAW_PATH))

        order = [abi.STYLE_FILL, abi.STYLE_STROKE]
        if style == abi.STYLE_STROKE_THEN_FILL:
            order.reverse()
        for st in order:
            if (st == abi.STYLE_FILL and need_fill) or (st == abi.STYLE_STROKE and need_stroke):
                run(st)

    def scene(self):
        return PackedScene(self.draws, self.width, self.height)


class PackedScene:
    """Owns the flat arrays of a hipr_scene; ``.scene`` points at it."""

    def __init__(self, draws, width, height):
        self.width, self.height = width, height
        n = len(draws)
        vo = np.zeros(n + 1, np.int32)
        po = np.zeros(n + 1, np.int32)
        wo = np.zeros(n + 1, np.int32)
        verbs, pts, ws, fills = [], [], [], []
        grads = []
        self.paints = (abi.Paint * max(n, 1))()
        draws = [tuple(d) + (None,) if len(d) == 4 else
                 (tuple(d) if len(d) == 5 else (d[0], d[1], IDENTITY, abi.DRAW_PATH, None))
                 for d in draws]
        mats = np.array([d[2] for d in draws] or [IDENTITY], np.float32).reshape(-1)
        kinds = np.array([d[3] for d in draws] or [0], np.uint8)
        self.cfs = (abi.ColorFilter * max(n, 1))()
        any_cf = False
        self.shaders = (abi.ImageShader * max(n, 1))()
        images, image_index = [], {}
        for i, d in enumerate(draws):
            paint = d[1]
            self.shaders[i].image = -1
            ish = getattr(paint, "image_shader", None)
            if ish is None:
                continue
            px, alpha_type, filt, tx, ty, local = ish
            key = (id(px), alpha_type)
            if key not in image_index:
                image_index[key] = len(images)
                images.append((px, alpha_type))
            sh = self.shaders[i]
            sh.image, sh.filter, sh.tile_x, sh.tile_y = image_index[key], filt, tx, ty
            for k, v in enumerate(local):
                sh.local_matrix[k] = v
        self.image_arrays = [px for px, _ in images]
        self.images = (abi.Image * max(len(images), 1))()
        for k, (px, alpha_type) in enumerate(images):
            self.images[k].height, self.images[k].width = px.shape[0], px.shape[1]
            self.images[k].alpha_type = alpha_type
            self.images[k].pixels = px.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        for i, d in enumerate(draws):
            path, paint = d[0], d[1]
            if getattr(paint, "color_filter", None) is not None:
                ctypes.memmove(ctypes.byref(self.cfs[i]), ctypes.byref(paint.color_filter),
                               ctypes.sizeof(abi.ColorFilter))
                any_cf = True
            verbs += path.verbs
            pts += path.points
            ws += path.weights
            fills.append(path.fill_type)
            vo[i + 1] = len(verbs)
            po[i + 1] = len(pts)
            wo[i + 1] = len(ws)
            ctypes.memmove(ctypes.byref(self.paints[i]), ctypes.byref(paint.p),
                           ctypes.sizeof(abi.Paint))
            if paint.gradient is not None:
                self.paints[i].gradient = len(grads)
                grads.append(paint.gradient)
            else:
                self.paints[i].gradient = -1
        self.verbs = np.array(verbs or [0], np.uint8)
        self.points = np.array(pts or [(0.0, 0.0)], np.float32).reshape(-1)
        self.weights = np.array(ws or [0.0], np.float32)
        self.fills = np.array(fills or [0], np.uint8)
        self.vo, self.po, self.wo = vo, po, wo
        self.grads = (abi.Gradient * max(len(grads), 1))(*grads)
        s = abi.Scene()
        s.n_paths = n
        s.verbs = self.verbs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        s.verb_offsets = vo.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        s.points = self.points.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        s.point_offsets = po.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        s.weights = self.weights.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        s.weight_offsets = wo.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        s.fill_types = self.fills.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        s.paints = ctypes.cast(self.paints, ctypes.POINTER(abi.Paint))
        s.gradients = ctypes.cast(self.grads, ctypes.POINTER(abi.Gradient))
        s.n_gradients = len(grads)
        self.mats, self.kinds = mats, kinds
        if any(tuple(d[2]) != IDENTITY for d in draws):
            s.matrices = mats.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        if any(d[3] != abi.DRAW_PATH for d in draws):
            s.draw_kinds = kinds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        if any_cf:
            s.color_filters = ctypes.cast(self.cfs, ctypes.POINTER(abi.ColorFilter))
        self.clips = np.array([d[4] if d[4] is not None else abi.CULL_RECT for d in draws]
                              or [abi.CULL_RECT], np.float32).reshape(-1)
        if any(d[4] is not None for d in draws):
            s.clip_rects = self.clips.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        if images:
            s.images = ctypes.cast(self.images, ctypes.POINTER(abi.Image))
            s.n_images = len(images)
            s.image_shaders = ctypes.cast(self.shaders, ctypes.POINTER(abi.ImageShader))
        self._s = s
        self.scene = ctypes.pointer(s)

    @property
    def n_paths(self):
        return self._s.n_paths
