"""Seeded synthetic DrawPath scenes (inputs shared by tests and bench.py).

Wraps ``libskity_scene.so`` (skity_amd/csrc/scene_gen.cpp).  A scene is the
ordered list of DrawPath calls a skity application would issue, packed as
the hipr_scene structure of include/skity_hip.h.
"""
import ctypes

import numpy as np

from . import abi

_lib = None

SEED_BASE = 0x5EED0000


def _load():
    global _lib
    if _lib is None:
        lib = abi.load("libskity_scene.so")
        lib.sg_generate.restype = ctypes.c_void_p
        lib.sg_generate.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32]
        lib.sg_destroy.argtypes = [ctypes.c_void_p]
        lib.sg_view.restype = ctypes.POINTER(abi.Scene)
        lib.sg_view.argtypes = [ctypes.c_void_p]
        lib.sg_canvas_size.argtypes = [ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int32)]
        lib.sg_generate_outlines.restype = ctypes.c_void_p
        lib.sg_generate_outlines.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_float, ctypes.c_float]
        lib.sg_slice.restype = ctypes.c_void_p
        lib.sg_slice.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
        lib.sg_struct_sizes.argtypes = [ctypes.POINTER(ctypes.c_int32)] * 3
        p, g, sc = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        lib.sg_struct_sizes(ctypes.byref(p), ctypes.byref(g), ctypes.byref(sc))
        if (p.value, g.value, sc.value) != (ctypes.sizeof(abi.Paint),
                                            ctypes.sizeof(abi.Gradient),
                                            ctypes.sizeof(abi.Scene)):
            raise RuntimeError(
                "libskity_scene.so was built from another include/skity_hip.h "
                f"(paint {p.value} B, gradient {g.value} B, scene {sc.value} B); rebuild it")
        _lib = lib
    return _lib


class SceneHandle:
    """Owns a generated scene; ``.scene`` is a ctypes pointer to hipr_scene."""

    def __init__(self, handle):
        self._lib = _load()
        self._h = handle
        self.scene = self._lib.sg_view(handle)
        w, h = ctypes.c_int32(), ctypes.c_int32()
        self._lib.sg_canvas_size(handle, ctypes.byref(w), ctypes.byref(h))
        self.width, self.height = w.value, h.value

    @property
    def n_paths(self):
        return self.scene.contents.n_paths

    def slice(self, begin, end):
        return SceneHandle(self._lib.sg_slice(self._h, begin, end))

    def paint(self, p):
        return self.scene.contents.paints[p]

    def verb_counts(self):
        s = self.scene.contents
        off = np.ctypeslib.as_array(s.verb_offsets, shape=(s.n_paths + 1,))
        return np.diff(off)

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._lib is not None:
            self._lib.sg_destroy(self._h)
            self._h = None


def generate(workload, n_paths=0, seed=None):
    """Workload 1: plain fill (800x600, 64 polygons); 2: curves + strokes
    (2048^2, 4096 paths); 3: gradients + blur (4096^2, 1024 paths); 4: one
    canvas of the batched uneven family (side 256-4096, 100-50,000 paths of
    4-10,000 line/quad/cubic segments, 30% gradients); 5: the same family
    scaled to 10-1,000 paths of 4-316 segments.  n_paths > 0 overrides the
    path count."""
    if seed is None:
        seed = SEED_BASE + workload
    h = _load().sg_generate(workload, seed, n_paths)
    if not h:
        raise ValueError(f"unknown workload {workload}")
    return SceneHandle(h)


def outlines(seed, side, first_path, n_paths, box_lo, box_hi):
    """Paths [first_path, first_path + n_paths) of the outline family on a
    side x side canvas, boxes of side side * 2^U[box_lo, box_hi] (see
    sg_generate_outlines)."""
    h = _load().sg_generate_outlines(seed, side, first_path, n_paths, box_lo, box_hi)
    if not h:
        raise ValueError("bad outline scene parameters")
    return SceneHandle(h)
