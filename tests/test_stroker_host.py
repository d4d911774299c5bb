"""The device stroker (skity_amd/csrc/hipr_stroke.hpp: StrokeOne, and the
stroke kernel's chunked feed) compiled for the host and compared with the
oracle's PathStroker restatement: the
outlines (verbs, points, conic weights) must be bit-identical.  This runs the
product's stroke code on the CPU, so a stroker bug shows up here without a
GPU; the GPU tests then check the same code on the device."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, scenes
from skity_amd.build import HIPCC
from skity_amd.canvas import Canvas, Paint, Path

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host_stroker", "host_stroker.hip")
LIB = os.path.join(HERE, "host_stroker", "libhost_stroker.so")
DEPS = [SRC] + [os.path.join(HERE, "..", "skity_amd", "csrc", f)
                for f in ("hipr_stroke.hpp", "hipr_geom.hpp", "hipr_common.hpp",
                          "hipr_flatten.hpp")] + [os.path.join(HERE, "..", "include",
                                                               "skity_hip.h")]


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(d) for d in DEPS):
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-fPIC", "-shared",
                        "-ffp-contract=off", "-Wno-unused-value", "-Wno-unused-result",
                        "-o", LIB + ".tmp", SRC], check=True)
        os.replace(LIB + ".tmp", LIB)
    lib = ctypes.CDLL(LIB)
    SP = ctypes.POINTER(abi.Scene)
    for L, name in ((lib, "hs_stroke_outline"), (lib, "hs_stroke_outline_chunked"),
                    (O.lib(), "orc_stroke_outline")):
        fn = getattr(L, name)
        fn.restype = ctypes.c_int64
        fn.argtypes = [SP, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                       ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    return lib


def _outline(fn, scene, p):
    c = np.zeros(3, np.int64)
    fn(scene, p, None, 0, None, 0, None, 0, c.ctypes.data)
    v = np.zeros(max(c[0], 1), np.uint8)
    pt = np.zeros((max(c[1], 1), 2), np.float32)
    w = np.zeros(max(c[2], 1), np.float32)
    fn(scene, p, v.ctypes.data, c[0], pt.ctypes.data, c[1], w.ctypes.data, c[2], c.ctypes.data)
    return v[:c[0]], pt[:c[1]], w[:c[2]]


def _same(host, scene, p):
    """The device stroker, run call by call and through the stroke kernel's
    chunked feed (segments stroked ahead by lanes), against the oracle."""
    b = _outline(O.lib().orc_stroke_outline, scene, p)
    for fn in (host.hs_stroke_outline, host.hs_stroke_outline_chunked):
        a = _outline(fn, scene, p)
        if not (np.array_equal(a[0], b[0]) and a[1].shape == b[1].shape
                and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
                and np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))):
            return False
    return True


@pytest.mark.parametrize("join", [abi.JOIN_MITER, abi.JOIN_ROUND, abi.JOIN_BEVEL])
@pytest.mark.parametrize("cap", [abi.CAP_BUTT, abi.CAP_ROUND, abi.CAP_SQUARE])
def test_joins_and_caps(host, join, cap):
    c = Canvas(160, 120)
    p = Paint().set_style(abi.STYLE_STROKE).set_stroke_width(9.0)
    p.set_stroke_join(join).set_stroke_cap(cap)
    c.draw_path(Path().move_to(12, 100).line_to(40, 20).line_to(70, 95).quad_to(100, 10, 140, 60), p)
    c.draw_path(Path().move_to(20, 20).cubic_to(60, -10, 120, 50, 60, 70).line_to(25, 60).close(), p)
    c.draw_path(Path().add_circle(120, 90, 18), p)
    c.draw_path(Path().move_to(5, 5).conic_to(80, 0, 60, 60, 3.0).line_to(5, 5), p)
    sc = c.scene()
    for i in range(sc.n_paths):
        assert _same(host, sc.scene, i), (join, cap, i)


def test_workload_strokes(host):
    """Every stroked draw of the curves + strokes workload (about 800)."""
    sc = scenes.generate(2)
    s = sc.scene.contents
    bad = [p for p in range(s.n_paths)
           if s.paints[p].style == abi.STYLE_STROKE and not _same(host, sc.scene, p)]
    assert not bad, bad[:10]


def test_forced_stroke_parameters(host):
    """Widths 0.3-23 px, all joins and caps, miter limits 1 (bevel), 4, 10."""
    sc = scenes.generate(2, 240, seed=23)
    s = sc.scene.contents
    for p in range(s.n_paths):
        s.paints[p].style = abi.STYLE_STROKE
        s.paints[p].stroke_width = [1.5, 4.0, 11.0, 23.0, 0.3][p % 5]
        s.paints[p].join = p % 3
        s.paints[p].cap = (p // 3) % 3
        s.paints[p].miter_limit = [4.0, 1.0, 10.0][p % 3]
    bad = [p for p in range(s.n_paths) if not _same(host, sc.scene, p)]
    assert not bad, bad[:10]


@pytest.mark.parametrize("cap", [abi.CAP_BUTT, abi.CAP_ROUND, abi.CAP_SQUARE])
def test_degenerate_segments(host, cap):
    """Segments the stroker skips or reduces: near-zero lines (skipped with a
    butt cap, so the next quad starts from an earlier point than the quad
    path's), repeated points, a quad folded back on itself, a cusp cubic and
    a width far above the curve size (deep quad-fit recursion)."""
    c = Canvas(200, 160)
    for k, width in enumerate((3.0, 40.0)):
        p = Paint().set_style(abi.STYLE_STROKE).set_stroke_width(width)
        p.set_stroke_join(k % 3).set_stroke_cap(cap)
        path = (Path().move_to(20, 20).line_to(20.00001, 20.0).quad_to(60, 5, 90, 40)
                .line_to(90, 40).line_to(90.000003, 40.000002).quad_to(95, 45, 90, 40)
                .quad_to(130, 80, 60, 60).cubic_to(140, 140, 20, 140, 100, 60)
                .quad_to(100.0001, 60.0001, 100.0002, 60.0).line_to(150, 150)
                .conic_to(180, 150, 180, 120, 0.3).close())
        c.draw_path(path, p)
        c.draw_path(Path().move_to(5, 150).quad_to(6, 149, 5, 150).line_to(5, 150), p)
    sc = c.scene()
    for i in range(sc.n_paths):
        assert _same(host, sc.scene, i), (cap, i)


def test_recursion_limit_carries_over(host):
    """Curves with coordinates near 1e6 and a 1 px width make PathStroker's
    quad fit give up after 33 levels; its depth then stays raised for the
    rest of the path (stroke.cc:552-564, reset only in the constructor,
    stroke.cc:849), which the chunked feed must not paper over."""
    import random
    rng = random.Random(5)
    c = Canvas(64, 64)
    for _ in range(24):
        p = Paint().set_style(abi.STYLE_STROKE).set_stroke_width(1.0)

        def r():
            return rng.uniform(-1e6, 1e6)
        c.draw_path(Path().move_to(r(), r()).quad_to(r(), r(), r(), r()).quad_to(r(), r(), r(), r())
                    .cubic_to(r(), r(), r(), r(), r(), r()).line_to(r(), r())
                    .quad_to(r(), r(), r(), r()), p)
    sc = c.scene()
    bad = [i for i in range(sc.n_paths) if not _same(host, sc.scene, i)]
    assert not bad, bad


def _hairline_scene(closed):
    c = Canvas(200, 160)
    for join in (abi.JOIN_MITER, abi.JOIN_ROUND, abi.JOIN_BEVEL):
        for cap in (abi.CAP_BUTT, abi.CAP_ROUND, abi.CAP_SQUARE):
            p = Paint().set_style(abi.STYLE_STROKE).set_stroke_width(0.0)
            p.set_stroke_join(join).set_stroke_cap(cap)
            if closed:
                c.draw_path(Path().move_to(20, 20).cubic_to(60, -10, 120, 50, 60, 70)
                            .line_to(25, 60).close(), p)
            else:
                c.draw_path(Path().move_to(12, 100).line_to(40, 20).line_to(70, 95)
                            .quad_to(100, 10, 140, 60), p)
    return c.scene()


@pytest.mark.parametrize("closed", [False, True])
def test_width_zero_strokes(host, closed):
    """Width 0: SWCanvas strokes it like any width, with PathStroker at radius
    0 (sw_canvas.cc:388-405, stroke.cc:865-866).  The host build of the
    device stroker gives the oracle's outlines bit for bit."""
    sc = _hairline_scene(closed)
    for i in range(sc.n_paths):
        assert _same(host, sc.scene, i), i


def test_width_zero_open_strokes_draw_nothing():
    """Closed form: an open path stroked at radius 0 has an outline whose
    forward and backward sides coincide, so every interval of the walk has
    winding 0 and no pixel is covered."""
    sc = _hairline_scene(False)
    img = O.render(sc.scene, sc.width, sc.height)
    assert not img.any()
