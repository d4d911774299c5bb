"""ctypes mirror of include/skity_hip.h (the C-ABI boundary).

Only plain data layouts and the library loader live here; the rasteriser
itself is the HIP library ``libskity_hip.so`` built from ``skity_amd/csrc``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

MAX_STOPS = 16

VERB_MOVE, VERB_LINE, VERB_QUAD, VERB_CONIC, VERB_CUBIC, VERB_CLOSE = range(6)
FILL_WINDING, FILL_EVEN_ODD = 0, 1
STYLE_FILL, STYLE_STROKE, STYLE_STROKE_AND_FILL, STYLE_STROKE_THEN_FILL = range(4)
CAP_BUTT, CAP_ROUND, CAP_SQUARE = range(3)
JOIN_MITER, JOIN_ROUND, JOIN_BEVEL = range(3)
GRADIENT_LINEAR, GRADIENT_RADIAL, GRADIENT_SWEEP, GRADIENT_CONICAL = 2, 3, 4, 5
TILE_CLAMP, TILE_REPEAT, TILE_MIRROR, TILE_DECAL = range(4)
BLUR_NORMAL, BLUR_SOLID, BLUR_OUTER, BLUR_INNER = 1, 2, 3, 4
(BLEND_CLEAR, BLEND_SRC, BLEND_DST, BLEND_SRC_OVER, BLEND_DST_OVER, BLEND_SRC_IN,
 BLEND_DST_IN, BLEND_SRC_OUT, BLEND_DST_OUT, BLEND_SRC_ATOP, BLEND_DST_ATOP, BLEND_XOR,
 BLEND_PLUS, BLEND_MODULATE, BLEND_SCREEN) = range(15)
BLEND_SOFT_LIGHT = 21
BLEND_MULTIPLY = 24
BLEND_LAST = 28
ALPHA_OPAQUE, ALPHA_PREMUL, ALPHA_UNPREMUL = 1, 2, 3
COLOR_RGBA, COLOR_BGRA = 1, 2
DRAW_PATH, DRAW_PAINT = 0, 1
CULL_RECT = (-1e9, -1e9, 1e9, 1e9)  # kMaxCullRect: a draw without a clip

# hipr_stage, in enum order
STAGES = ("stroke", "flatten_count", "flatten", "sort", "walk_count", "walk", "layers",
          "blur", "composite", "presize", "walk_schedule")
MODE_EXACT, MODE_ESTIMATES, MODE_LEARNED = 0, 1, 2
PRESIZER_FEATURES, PRESIZER_TRAINED, PRESIZER_PARAMS = 16, 5378, 5414

SUCCESS = 0
ERROR_INVALID_HANDLE = -1
ERROR_INVALID_ARGUMENT = -2
ERROR_NOT_SUPPORTED = -5
ERROR_CAPACITY = -8


class Paint(ctypes.Structure):
    _fields_ = [
        ("color", ctypes.c_float * 4),
        ("stroke_color", ctypes.c_float * 4),
        ("style", ctypes.c_int32),
        ("stroke_width", ctypes.c_float),
        ("miter_limit", ctypes.c_float),
        ("cap", ctypes.c_int32),
        ("join", ctypes.c_int32),
        ("gradient", ctypes.c_int32),
        ("blur_radius", ctypes.c_float),
        ("blur_style", ctypes.c_int32),
        ("blend", ctypes.c_int32),
    ]


class Gradient(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("tile_mode", ctypes.c_int32),
        ("color_count", ctypes.c_int32),
        ("has_offsets", ctypes.c_int32),
        ("points", ctypes.c_float * 4),
        ("radius", ctypes.c_float * 2),
        ("local_matrix", ctypes.c_float * 6),
        ("colors", (ctypes.c_float * 4) * MAX_STOPS),
        ("offsets", ctypes.c_float * MAX_STOPS),
    ]


CF_NONE, CF_BLEND, CF_MATRIX, CF_LINEAR_TO_SRGB, CF_SRGB_TO_LINEAR, CF_COMPOSE = range(6)


class ColorFilter(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32),
        ("color", ctypes.c_uint32),
        ("blend", ctypes.c_int32),
        ("matrix", ctypes.c_float * 20),
    ]


FILTER_NEAREST, FILTER_LINEAR = 0, 1


class Image(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("alpha_type", ctypes.c_int32),
        ("pixels", ctypes.POINTER(ctypes.c_uint8)),
    ]


class ImageShader(ctypes.Structure):
    _fields_ = [
        ("image", ctypes.c_int32),
        ("filter", ctypes.c_int32),
        ("tile_x", ctypes.c_int32),
        ("tile_y", ctypes.c_int32),
        ("local_matrix", ctypes.c_float * 6),
    ]


class Scene(ctypes.Structure):
    _fields_ = [
        ("n_paths", ctypes.c_int32),
        ("verbs", ctypes.POINTER(ctypes.c_uint8)),
        ("verb_offsets", ctypes.POINTER(ctypes.c_int32)),
        ("points", ctypes.POINTER(ctypes.c_float)),
        ("point_offsets", ctypes.POINTER(ctypes.c_int32)),
        ("weights", ctypes.POINTER(ctypes.c_float)),
        ("weight_offsets", ctypes.POINTER(ctypes.c_int32)),
        ("fill_types", ctypes.POINTER(ctypes.c_uint8)),
        ("paints", ctypes.POINTER(Paint)),
        ("gradients", ctypes.POINTER(Gradient)),
        ("n_gradients", ctypes.c_int32),
        ("matrices", ctypes.POINTER(ctypes.c_float)),
        ("draw_kinds", ctypes.POINTER(ctypes.c_uint8)),
        ("color_filters", ctypes.POINTER(ColorFilter)),
        ("images", ctypes.POINTER(Image)),
        ("n_images", ctypes.c_int32),
        ("image_shaders", ctypes.POINTER(ImageShader)),
        ("clip_rects", ctypes.POINTER(ctypes.c_float)),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("paths", ctypes.c_int64),
        ("raster_paths", ctypes.c_int64),
        ("edges", ctypes.c_int64),
        ("tiles", ctypes.c_int64),
        ("tile_planes", ctypes.c_int64),
        ("fallbacks", ctypes.c_int64),
        ("max_active", ctypes.c_int64),
        ("ms_total", ctypes.c_double),
        ("blends", ctypes.c_int64),
        ("pixels", ctypes.c_int64),
        ("underestimates", ctypes.c_int64),
    ]


BLUR_IMPL_MFMA, BLUR_IMPL_LINES = 0, 1
FLAG_NO_GRAPH, FLAG_PERSISTENT = 1, 2


class Options(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("edge_estimates", ctypes.POINTER(ctypes.c_int32)),
        ("tile_estimates", ctypes.POINTER(ctypes.c_int32)),
        ("blur_impl", ctypes.c_int32),
        ("band_rows", ctypes.c_int32),
        ("band_min_edges", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("chunk_points", ctypes.c_int32),
    ]


def load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run __graft_entry__.build() first")
    return ctypes.CDLL(path)
