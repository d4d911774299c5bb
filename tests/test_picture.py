"""The SKP reader and display lists of the C++ front end (f1), on the CPU:
the oracle draws the scenes the recordings make.

* tiger.skp (the reference's own SKP) played back under Translate(-130, 20)
  and drawn by the oracle against the reference's golden image of the same
  recording (test/golden/cases/skp/skp.cc:50-68).  The golden comes from the
  reference's coverage-AA GPU backend, so edges differ by its antialiasing;
  interior pixels are exact.  The thresholds are shown to fail when the
  reader misreads the matrices or the stroke widths (planted in the SKP).
* Op semantics (picture.cc:312-1043) on synthetic SKPs: the same calls made
  on the Python canvas mirror draw identical pixels.
* Calls the backend does not draw make the recording NOT_SUPPORTED; ops on
  which the reference ends playback end it here; truncated and corrupted
  files are refused or end playback without a crash."""
import struct

import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi
from skity_amd.canvas import Canvas, Paint, Path

import skp_kit as K

W = H = 1000
TIGER_MATRIX = (1, 0, -130, 0, 1, 20)


def _oracle(rec):
    return O.render_ex(rec.scene, rec.width, rec.height, abi.ALPHA_PREMUL, abi.COLOR_RGBA)


def _golden_stats(img, golden):
    g = golden.astype(np.int32)
    d = np.abs(g - img.astype(np.int32)).max(-1)
    pad = np.pad(g, ((2, 2), (2, 2), (0, 0)), mode="edge")
    interior = np.ones(g.shape[:2], bool)
    for dy in range(5):
        for dx in range(5):
            interior &= (pad[dy:dy + g.shape[0], dx:dx + g.shape[1]] == g).all(-1)
    return dict(mean=float(d.mean()), over16=float((d > 16).mean()),
                over64=float((d > 64).mean()),
                interior_exact=float((d[interior] == 0).mean()),
                interior_max=int(d[interior].max()))


def _near_golden(st):
    return (st["mean"] < 1.0 and st["over16"] < 0.02 and st["over64"] < 0.001
            and st["interior_exact"] > 0.999 and st["interior_max"] <= 1)


@pytest.fixture(scope="module")
def golden():
    return K.tiger_golden()


def test_tiger_tables():
    rec = K.Recording(K.tiger_bytes(), W, H, TIGER_MATRIX)
    assert rec.rc == 0, rec.error
    cull, version, paints, paths = rec.info()
    assert cull == (0.0, 0.0, 2560.0, 1920.0) and version == 109
    assert (paints, paths) == (305, 240)
    sc = rec.scene.contents
    assert sc.n_paths == 305  # one DrawPath op per paint
    m = np.ctypeslib.as_array(sc.matrices, (305 * 6,)).reshape(305, 6)
    # every draw is under Translate(-130, 20) * the picture's CONCAT44s
    assert np.all(m[:, 2] != 0) and len(np.unique(m, axis=0)) <= 3
    rec.close()


def test_tiger_matches_reference_golden(golden):
    rec = K.Recording(K.tiger_bytes(), W, H, TIGER_MATRIX)
    st = _golden_stats(_oracle(rec), golden)
    print("tiger vs the reference's golden:", st, flush=True)
    assert _near_golden(st)
    rec.close()


def _tiger_ops_offset(data):
    # header 29 bytes, then the "read" chunk: tag, size, op stream
    assert data[29:33] == b"daer"
    size, = struct.unpack_from("<I", data, 33)
    return 37, size


def test_golden_check_catches_misread_matrices_and_widths(golden):
    data = bytearray(K.tiger_bytes())
    off, size = _tiger_ops_offset(data)
    # CONCAT44 read as identity
    bad = bytearray(data)
    o, n_concat = off, 0
    while o < off + size:
        bits, = struct.unpack_from("<I", bad, o)
        op, sz = bits >> 24, bits & 0xFFFFFF
        if op == K.OPS["CONCAT44"]:
            struct.pack_into("<16f", bad, o + 4, *np.eye(4, dtype=np.float32).ravel())
            n_concat += 1
        o += sz
    assert n_concat == 2
    rec = K.Recording(bytes(bad), W, H, TIGER_MATRIX)
    st = _golden_stats(_oracle(rec), golden)
    print("matrices dropped:", st, flush=True)
    assert not _near_golden(st)
    rec.close()
    # stroke widths read 1.5x
    bad = bytearray(data)
    p = bad.index(b" tnp")
    count, = struct.unpack_from("<I", bad, p + 4)
    q = p + 8
    for _ in range(count):
        w, = struct.unpack_from("<f", bad, q)
        struct.pack_into("<f", bad, q, 1.5 * w)
        q += 28
    rec = K.Recording(bytes(bad), W, H, TIGER_MATRIX)
    st = _golden_stats(_oracle(rec), golden)
    print("widths 1.5x:", st, flush=True)
    assert not _near_golden(st)
    rec.close()


def _q(c):
    """Color4fToColor then Color4fFromColor: 8-bit truncation, / 255."""
    return tuple(float(np.float32(int(min(max(np.float32(v) * np.float32(255), 0), 255)))
                       / np.float32(255)) for v in c)


def test_op_semantics_match_canvas_calls():
    red, blue, green = (0.9, 0.1, 0.2, 1.0), (0.1, 0.3, 0.8, 0.6), (0.2, 0.7, 0.3, 0.8)
    paints = [K.paint_bytes(red),
              K.paint_bytes(blue, style=abi.STYLE_STROKE, width=3.0, join=abi.JOIN_ROUND,
                            cap=abi.CAP_ROUND),
              K.paint_bytes(green, blend=abi.BLEND_MULTIPLY)]
    tri = ([0, 1, 1, 4, 1], [(10, 10), (60, 15), (30, 50), (70, 60), (90, 20), (95, 70), (40, 90)])
    paths = [K.path_bytes(*tri), K.path_bytes([0, 1, 1, 1, 5], [(5, 5), (40, 5), (40, 40), (5, 40)],
                                              fill=1)]
    m3 = struct.pack("<9f", 0.9, 0.2, 12.0, -0.1, 1.1, 4.0, 0, 0, 1)
    m44 = struct.pack("<16f", 1.2, 0.1, 0, 0, -0.15, 0.8, 0, 0, 0, 0, 1, 0, 20, 30, 0, 1)
    ops = [K.op_bytes("SAVE"),
           K.op_bytes("TRANSLATE", ("f", 20.0), ("f", 10.0)),
           K.op_bytes("SCALE", ("f", 1.5), ("f", 1.25)),
           K.op_bytes("DRAW_PATH", ("i", 1), ("i", 1)),
           K.op_bytes("ROTATE", ("f", 30.0)),
           K.op_bytes("DRAW_PATH", ("i", 2), ("i", 2)),
           K.op_bytes("RESTORE"),
           K.op_bytes("SKEW", ("f", 0.2), ("f", -0.1)),
           K.op_bytes("DRAW_RECT", ("i", 3), struct.pack("<4f", 100, 100, 180, 150)),
           K.op_bytes("CONCAT", m3),
           K.op_bytes("DRAW_OVAL", ("i", 2), struct.pack("<4f", 30, 120, 110, 170)),
           K.op_bytes("SET_MATRIX", m3),
           K.op_bytes("NOOP", struct.pack("<2I", 7, 9)),
           K.op_bytes("DRAW_PATH", ("i", 3), ("i", 1)),
           K.op_bytes("SET_M44", m44),
           K.op_bytes("DRAW_PATH", ("i", 1), ("i", 2)),
           K.op_bytes("CONCAT44", m44),
           K.op_bytes("DRAW_PAINT", ("i", 3))]
    rec = K.Recording(K.skp_bytes(ops, paints, paths), 256, 256, (1, 0, 3, 0, 1, -2))
    assert rec.rc == 0, rec.error
    got = _oracle(rec)
    rec.close()

    cv = Canvas(256, 256)
    cv.concat((1, 0, 3, 0, 1, -2))
    p_red = Paint(_q(red))
    p_blue = Paint(_q(blue)).set_style(abi.STYLE_STROKE)
    p_blue.set_stroke_width(3.0)
    p_blue.set_stroke_join(abi.JOIN_ROUND)
    p_blue.set_stroke_cap(abi.CAP_ROUND)
    p_green = Paint(_q(green)).set_blend_mode(abi.BLEND_MULTIPLY)
    path1 = Path().move_to(10, 10).line_to(60, 15).line_to(30, 50)
    path1.cubic_to(70, 60, 90, 20, 95, 70).line_to(40, 90)
    path2 = Path().move_to(5, 5).line_to(40, 5).line_to(40, 40).line_to(5, 40).close()
    path2.fill_type = abi.FILL_EVEN_ODD
    cv.save()  # the playback's own Save
    cv.save()
    cv.translate(20.0, 10.0)
    cv.scale(1.5, 1.25)
    cv.draw_path(path1, p_red)
    cv.rotate(30.0)
    cv.draw_path(path2, p_blue)
    cv.restore()
    cv.skew(0.2, -0.1)
    cv.draw_path(Path().add_rect(100, 100, 180, 150), p_green)
    m = (0.9, 0.2, 12.0, -0.1, 1.1, 4.0)
    cv.concat(m)
    cv.draw_path(Path().add_oval(30, 120, 110, 170, start=3), p_blue)
    cv.set_matrix(m)
    cv.draw_path(path1, p_green)
    a = (1.2, -0.15, 20.0, 0.1, 0.8, 30.0)
    cv.set_matrix(a)
    cv.draw_path(path2, p_red)
    cv.concat(a)
    cv.draw_paint(p_green)
    sc = cv.scene()
    want = O.render_ex(sc.scene, 256, 256, abi.ALPHA_PREMUL, abi.COLOR_RGBA)
    assert want.any()
    assert np.array_equal(got, want)


def test_calls_not_drawn_make_the_recording_unsupported():
    paints = [K.paint_bytes((1, 0, 0, 1))]
    paths = [K.path_bytes([0, 1, 1, 5], [(0, 0), (50, 0), (0, 50)])]
    draw = K.op_bytes("DRAW_PATH", ("i", 1), ("i", 1))
    cases = {
        "difference clip rect": K.op_bytes("CLIP_RECT", struct.pack("<4f", 0, 0, 10, 10),
                                           ("I", 0), ("i", 0)),
        "clip path": K.op_bytes("CLIP_PATH", ("i", 1), ("I", 1), ("i", 0)),
        "save layer": K.op_bytes("SAVE_LAYER_SAVELAYERREC", ("I", 0)),
        "clip rrect": K.op_bytes("CLIP_RRECT", struct.pack("<12f", *range(12)), ("I", 1),
                                 ("i", 0)),
        "text blob": K.op_bytes("DRAW_TEXT_BLOB", ("i", 1), ("i", 1), ("f", 0), ("f", 0)),
        "perspective": K.op_bytes("CONCAT", struct.pack("<9f", 1, 0, 0, 0, 1, 0, 0.001, 0, 1)),
    }
    for name, op in cases.items():
        rec = K.Recording(K.skp_bytes([draw, op, draw], paints, paths), 64, 64)
        assert rec.rc_make == 0, name
        assert rec.rc == abi.ERROR_NOT_SUPPORTED, name
        rec.close()
    # a paint with a shader (any non-null effect but the blender) and an
    # inverse fill type are not drawn either
    fx = K.paint_bytes((1, 0, 0, 1), effects=[0, (1, b"\0" * 8), 0, 0, 0, 0])
    rec = K.Recording(K.skp_bytes([draw], [fx], paths, factories=["SkLinearGradient"]), 64, 64)
    assert rec.rc == abi.ERROR_NOT_SUPPORTED
    rec.close()
    blender_only = K.paint_bytes((1, 0, 0, 1), blend=abi.BLEND_SCREEN,
                                 effects=[0, 0, 0, 0, 0, (1, b"\0" * 4)])
    rec = K.Recording(K.skp_bytes([draw], [blender_only], paths,
                                  factories=["SkBlendModeBlender"]), 64, 64)
    assert rec.rc == 0 and rec.scene.contents.n_paths == 1
    assert rec.scene.contents.paints[0].blend == abi.BLEND_SCREEN
    rec.close()
    inv = [K.path_bytes([0, 1, 1, 5], [(0, 0), (50, 0), (0, 50)], fill=2)]
    rec = K.Recording(K.skp_bytes([draw], paints, inv), 64, 64)
    assert rec.rc == abi.ERROR_NOT_SUPPORTED
    rec.close()


def test_clip_rect_narrows_later_draws():
    """CLIP_RECT (intersect) under a scale + translate matrix: the same calls
    on the canvas mirror draw identical pixels; Restore pops the clip."""
    paints = [K.paint_bytes((0.2, 0.6, 0.3, 1.0))]
    paths = [K.path_bytes([0, 1, 1, 1, 5], [(0, 0), (60, 0), (60, 50), (0, 50)])]
    draw = K.op_bytes("DRAW_PATH", ("i", 1), ("i", 1))
    ops = [K.op_bytes("SAVE"), K.op_bytes("SCALE", ("f", 1.5), ("f", 1.25)),
           K.op_bytes("CLIP_RECT", struct.pack("<4f", 5.5, 4.0, 30.25, 22.0), ("I", 1), ("i", 0)),
           draw, K.op_bytes("RESTORE"), K.op_bytes("TRANSLATE", ("f", 50.0), ("f", 40.0)), draw]
    rec = K.Recording(K.skp_bytes(ops, paints, paths), 128, 128)
    assert rec.rc == 0, rec.error
    got = _oracle(rec)
    rec.close()
    cv = Canvas(128, 128)
    p = Paint(_q((0.2, 0.6, 0.3, 1.0)))
    path = Path().move_to(0, 0).line_to(60, 0).line_to(60, 50).line_to(0, 50).close()
    cv.save()
    cv.save()
    cv.scale(1.5, 1.25)
    cv.clip_rect(5.5, 4.0, 30.25, 22.0)
    cv.draw_path(path, p)
    cv.restore()
    cv.translate(50.0, 40.0)
    cv.draw_path(path, p)
    sc = cv.scene()
    want = O.render_ex(sc.scene, 128, 128, abi.ALPHA_PREMUL, abi.COLOR_RGBA)
    assert np.array_equal(got, want)
    assert got[20, 20, 3] == 255 and got[10, 60, 3] == 0  # inside / outside the clip


def test_ops_that_end_playback_end_it():
    """DRAW_PICTURE (and the other ops the reference refuses) end playback:
    the draws after it are not made; DRAW_POINTS is skipped."""
    paints = [K.paint_bytes((0, 0, 1, 1))]
    paths = [K.path_bytes([0, 1, 1, 5], [(0, 0), (50, 0), (0, 50)])]
    draw = K.op_bytes("DRAW_PATH", ("i", 1), ("i", 1))
    pts = K.op_bytes("DRAW_POINTS", ("i", 1), ("I", 0), ("I", 2), struct.pack("<4f", 1, 2, 3, 4))
    rec = K.Recording(K.skp_bytes([draw, pts, draw, K.op_bytes("DRAW_PICTURE", ("i", 1)), draw],
                                  paints, paths), 64, 64)
    assert rec.rc == 0 and rec.scene.contents.n_paths == 2
    rec.close()
    # a bad paint index ends playback too
    rec = K.Recording(K.skp_bytes([draw, K.op_bytes("DRAW_PATH", ("i", 5), ("i", 1)), draw],
                                  paints, paths), 64, 64)
    assert rec.rc == 0 and rec.scene.contents.n_paths == 1
    rec.close()


def test_refused_files():
    good = K.tiger_bytes()
    for bad, why in [(b"skiapicx" + good[8:], "magic"),
                     (good[:8] + struct.pack("<I", 81) + good[12:], "version"),
                     (good[:8] + struct.pack("<I", 110) + good[12:], "version")]:
        rec = K.Recording(bad, 64, 64)
        assert rec.rc_make == abi.ERROR_NOT_SUPPORTED and why in rec.error
    # typefaces: the chunk counts entries, which this reader does not parse
    s = K.skp_bytes([K.op_bytes("SAVE")])
    i = s.index(b"cfpt")
    s = s[:i + 4] + struct.pack("<I", 1) + s[i + 8:]
    rec = K.Recording(s, 64, 64)
    assert rec.rc_make == abi.ERROR_NOT_SUPPORTED and "typeface" in rec.error


def test_truncated_and_corrupted_files_do_not_crash():
    good = K.tiger_bytes()
    rng = np.random.default_rng(5)
    for n in sorted(set(rng.integers(0, len(good), 150).tolist()) | {0, 1, 28, 29, 36, 37}):
        rec = K.Recording(good[:n], 64, 64)
        assert rec.rc_make != 0  # the eof tag is missing
        rec.close()
    off, size = _tiger_ops_offset(good)
    for k in range(60):
        bad = bytearray(good)
        for pos in rng.integers(off, len(good), 8):
            bad[pos] = int(rng.integers(0, 256))
        rec = K.Recording(bytes(bad), 64, 64)
        if rec.rc == 0:
            assert 0 <= rec.scene.contents.n_paths <= 305
        rec.close()


def test_rrect_ops_build_rrect_paths():
    """DRAW_RRECT (SetRectRadii then Path::AddRRect from the legacy start
    index 6: line / conic pairs around the radii points), DRAW_DRRECT (the
    inner then the outer rrect, even-odd) and a rounded-rect path record
    (ReadRRectPath: direction and start index) become the paths the
    reference builds; the oracle draws them inside their bounds."""
    paints = [K.paint_bytes((0.3, 0.4, 0.8, 1.0))]
    rr = struct.pack("<12f", 10, 20, 100, 200, 10, 10, 10, 10, 10, 10, 10, 10)
    inner = struct.pack("<12f", 30, 40, 80, 150, 5, 5, 5, 5, 5, 5, 5, 5)
    # a serialized rrect path: type 1, CCW, start 2, even-odd
    rpath = struct.pack("<I", (1 << 28) | (1 << 26) | (1 << 8) | 5) + rr + struct.pack("<i", 2)
    paths = [rpath]
    ops = [K.op_bytes("DRAW_RRECT", ("i", 1), rr),
           K.op_bytes("DRAW_DRRECT", ("i", 1), rr, inner),
           K.op_bytes("DRAW_PATH", ("i", 1), ("i", 1))]
    rec = K.Recording(K.skp_bytes(ops, paints, paths), 128, 256)
    assert rec.rc == 0, rec.error
    s = rec.scene.contents
    assert s.n_paths == 3
    verbs = [s.verbs[i] for i in range(s.verb_offsets[0], s.verb_offsets[1])]
    pts = [(s.points[2 * i], s.points[2 * i + 1])
           for i in range(s.point_offsets[0], s.point_offsets[1])]
    L, C, M, Z = abi.VERB_LINE, abi.VERB_CONIC, abi.VERB_MOVE, abi.VERB_CLOSE
    assert verbs == [M, L, C, L, C, L, C, L, C, Z]
    assert pts == [(10, 190), (10, 30), (10, 20), (20, 20), (90, 20), (100, 20), (100, 30),
                   (100, 190), (100, 200), (90, 200), (20, 200), (10, 200), (10, 190)]
    # DRRECT: two closed contours, inner first, even-odd
    v1 = [s.verbs[i] for i in range(s.verb_offsets[1], s.verb_offsets[2])]
    assert v1.count(M) == 2 and s.fill_types[1] == abi.FILL_EVEN_ODD
    first = (s.points[2 * s.point_offsets[1]], s.points[2 * s.point_offsets[1] + 1])
    assert first == (30, 145)  # the inner rrect's start point
    # the path record: counter-clockwise from start index 2, even-odd
    v2 = [s.verbs[i] for i in range(s.verb_offsets[2], s.verb_offsets[3])]
    assert v2[0] == M and v2[-1] == Z and s.fill_types[2] == abi.FILL_EVEN_ODD
    assert (s.points[2 * s.point_offsets[2]], s.points[2 * s.point_offsets[2] + 1]) == (100, 30)
    img = _oracle(rec)
    a = img[..., 3]
    assert a[100, 50] > 0 and a[100, 20] > 0 and not a[:19].any() and not a[:, 101:].any()
    rec.close()
