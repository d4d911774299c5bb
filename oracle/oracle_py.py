"""ctypes loader of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product package never imports this module.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from skity_amd import abi  # noqa: E402

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        SP = ctypes.POINTER(abi.Scene)
        L.orc_render_scene.argtypes = [SP, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
        L.orc_render_scene_mt.argtypes = [SP, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32]
        L.orc_render_scene_ex.argtypes = [SP, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
        L.orc_path_edges.restype = ctypes.c_int64
        L.orc_path_edges.argtypes = [SP, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                     ctypes.c_int64]
        L.orc_path_spans.restype = ctypes.c_int64
        L.orc_path_spans.argtypes = [SP, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_void_p]
        L.orc_path_tile_bins.restype = ctypes.c_int64
        L.orc_path_tile_bins.argtypes = [SP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.POINTER(ctypes.c_int64)]
        L.orc_cubic_quad_count.restype = ctypes.c_int32
        L.orc_cubic_quad_count.argtypes = [ctypes.c_void_p]
        L.orc_stackblur.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_int32]
        L.orc_stackblur_tables.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.POINTER(ctypes.c_uint32)]
        L.orc_gradient_color.restype = ctypes.c_uint32
        L.orc_gradient_color.argtypes = [ctypes.POINTER(abi.Gradient), ctypes.c_int32,
                                         ctypes.c_int32]
        L.orc_blur_layer.restype = ctypes.c_int64
        L.orc_blur_layer.argtypes = [SP, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                     ctypes.c_void_p]
        L.orc_quad_path.restype = ctypes.c_int64
        L.orc_quad_path.argtypes = [SP, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.POINTER(ctypes.c_int64)]
        _lib = L
    return _lib


def render(scene_ptr, width, height, base=None):
    img = np.zeros((height, width, 4), np.uint8) if base is None else base.copy()
    rc = lib().orc_render_scene(scene_ptr, img.ctypes.data, width, height)
    assert rc == 0
    return img


def render_mt(scene_ptr, width, height, threads=None, base=None):
    """render() on several threads (same image)."""
    if threads is None:
        threads = max(1, min(16, os.cpu_count() or 1))
    img = np.zeros((height, width, 4), np.uint8) if base is None else base.copy()
    rc = lib().orc_render_scene_mt(scene_ptr, img.ctypes.data, width, height, threads)
    assert rc == 0
    return img


def render_ex(scene_ptr, width, height, alpha_type, color_type, base=None, threads=None):
    """render() onto a target of the given abi.ALPHA_* / abi.COLOR_* types."""
    if threads is None:
        threads = max(1, min(16, os.cpu_count() or 1))
    img = np.zeros((height, width, 4), np.uint8) if base is None else base.copy()
    rc = lib().orc_render_scene_ex(scene_ptr, img.ctypes.data, width, height, alpha_type,
                                   color_type, threads)
    assert rc == 0
    return img


def path_edges(scene_ptr, p, stroke=0):
    n = lib().orc_path_edges(scene_ptr, p, stroke, None, 0)
    out = np.zeros((max(n, 1), 16), np.int32)
    lib().orc_path_edges(scene_ptr, p, stroke, out.ctypes.data, n)
    return out[:n]


def path_spans(scene_ptr, p, stroke=0):
    b = np.zeros(4, np.float32)
    n = lib().orc_path_spans(scene_ptr, p, stroke, None, 0, b.ctypes.data)
    out = np.zeros((max(n, 1), 4), np.int32)
    lib().orc_path_spans(scene_ptr, p, stroke, out.ctypes.data, n, b.ctypes.data)
    return out[:n], b


def path_tile_bins(scene_ptr, p, stroke, width, height):
    """Tile-bin codes of one raster of path p (see orc_path_tile_bins) and the
    number of pixels with two direct covers."""
    st = ctypes.c_int64()
    n = lib().orc_path_tile_bins(scene_ptr, p, stroke, width, height, None, 0, ctypes.byref(st))
    out = np.zeros(max(n, 1), np.int32)
    lib().orc_path_tile_bins(scene_ptr, p, stroke, width, height, out.ctypes.data, n,
                             ctypes.byref(st))
    return out[:n], st.value


def stackblur(img, radius):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    out = np.zeros_like(img)
    lib().orc_stackblur(img.ctypes.data, out.ctypes.data, w, h, radius)
    return out


def stackblur_tables(r):
    m, s = ctypes.c_uint32(), ctypes.c_uint32()
    lib().orc_stackblur_tables(r, ctypes.byref(m), ctypes.byref(s))
    return m.value, s.value


def cubic_quad_count(pts8):
    a = np.ascontiguousarray(pts8, np.float32)
    return lib().orc_cubic_quad_count(a.ctypes.data)


def gradient_color(grad, x, y):
    return lib().orc_gradient_color(ctypes.byref(grad), x, y)


def blur_layer(scene_ptr, p):
    b = np.zeros(4, np.int32)
    n = lib().orc_blur_layer(scene_ptr, p, None, 0, b.ctypes.data)
    w, h = b[2] - b[0], b[3] - b[1]
    out = np.zeros((h, w, 4), np.uint8)
    lib().orc_blur_layer(scene_ptr, p, out.ctypes.data, n, b.ctypes.data)
    return out, b


def quad_path(scene_ptr, p):
    """Stroke::QuadPath of path p: (verbs uint8[n], points float32[m, 2])."""
    npts = ctypes.c_int64()
    nv = lib().orc_quad_path(scene_ptr, p, None, 0, None, 0, ctypes.byref(npts))
    verbs = np.zeros(max(nv, 1), np.uint8)
    pts = np.zeros((max(npts.value, 1), 2), np.float32)
    lib().orc_quad_path(scene_ptr, p, verbs.ctypes.data, nv, pts.ctypes.data, npts.value,
                        ctypes.byref(npts))
    return verbs[:nv], pts[:npts.value]
