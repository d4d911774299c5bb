"""GPU: SKP playback through display lists (f1).  tiger.skp (the
reference's own SKP) recorded under Translate(-130, 20) and drawn on the
device through skity_hip_display_list_draw equals the oracle's drawing of
the same recording bit for bit, in the exact and learned forms, at the
golden's 1000 x 1000 and scaled to 4096 x 4096 and 8192 x 8192 (the oracle
drawing is pinned against the reference's golden image in
tests/test_picture.py)."""
import numpy as np
import pytest
import torch  # noqa: F401  (the HIP runtime is initialised by torch first)

from oracle import oracle_py as O
from skity_amd import abi, device, presizer

from helpers import diff_report
import skp_kit as K

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    presizer.install(c, *presizer.load())
    yield c
    c.close()


@pytest.mark.parametrize("size,matrix", [(1000, (1, 0, -130, 0, 1, 20)),
                                         (4096, (4.1, 0, -560, 0, 4.1, 60)),
                                         (8192, (8.2, 0, -1120, 0, 8.2, 120))])
@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
def test_tiger_display_list_matches_oracle(ctx, size, matrix, mode):
    rec = K.Recording(K.tiger_bytes(), size, size, matrix)
    assert rec.rc == 0, rec.error
    surf = device.Surface(ctx, size, size)
    surf.clear(0)
    rc = rec.draw(surf, mode)
    assert rc == 0
    got = surf.read()
    surf.close()
    want = O.render_ex(rec.scene, size, size, abi.ALPHA_PREMUL, abi.COLOR_RGBA)
    rec.close()
    mx, cnt = diff_report(got, want)
    print(f"tiger {size}^2 mode {mode}: max diff {mx}, {cnt} pixels differ, "
          f"{int((want[..., 3] > 0).sum())} covered", flush=True)
    assert mx == 0


def test_unsupported_recording_draws_nothing(ctx):
    import struct
    paints = [K.paint_bytes((1, 0, 0, 1))]
    paths = [K.path_bytes([0, 1, 1, 5], [(0, 0), (50, 0), (0, 50)])]
    draw = K.op_bytes("DRAW_PATH", ("i", 1), ("i", 1))
    # a difference clip clips span lists, which the backend does not do
    clip = K.op_bytes("CLIP_RECT", struct.pack("<4f", 0, 0, 10, 10), ("I", 0), ("i", 0))
    rec = K.Recording(K.skp_bytes([draw, clip, draw], paints, paths), 64, 64)
    surf = device.Surface(ctx, 64, 64)
    surf.clear(0)
    rc = rec.draw(surf, abi.MODE_EXACT)
    assert rc == abi.ERROR_NOT_SUPPORTED
    assert not surf.read().any()
    surf.close()
    rec.close()


def test_rrect_ops_match_oracle(ctx):
    """DRAW_RRECT / DRAW_DRRECT / rounded-rect path records (conic outlines
    under a scale + translate and a rotation) drawn by the device equal the
    oracle."""
    import struct
    paints = [K.paint_bytes((0.3, 0.4, 0.8, 0.9)),
              K.paint_bytes((0.8, 0.2, 0.1, 1.0), style=abi.STYLE_STROKE, width=3.0)]
    rr = struct.pack("<12f", 10, 20, 100, 200, 10, 30, 25, 5, 40, 40, 0, 0)
    inner = struct.pack("<12f", 30, 40, 80, 150, 5, 5, 5, 5, 5, 5, 5, 5)
    rpath = struct.pack("<I", (1 << 28) | (1 << 26) | 5) + rr + struct.pack("<i", 5)
    ops = [K.op_bytes("DRAW_RRECT", ("i", 1), rr),
           K.op_bytes("TRANSLATE", ("f", 120.0), ("f", 10.0)),
           K.op_bytes("DRAW_DRRECT", ("i", 2), rr, inner),
           K.op_bytes("ROTATE", ("f", 12.0)),
           K.op_bytes("DRAW_PATH", ("i", 1), ("i", 1))]
    rec = K.Recording(K.skp_bytes(ops, paints, [rpath]), 256, 256, (1.2, 0, 4, 0, 1.1, 3))
    assert rec.rc == 0, rec.error
    surf = device.Surface(ctx, 256, 256)
    surf.clear(0)
    assert rec.draw(surf, abi.MODE_EXACT) == 0
    got = surf.read()
    surf.close()
    want = O.render_ex(rec.scene, 256, 256, abi.ALPHA_PREMUL, abi.COLOR_RGBA)
    rec.close()
    mx, cnt = diff_report(got, want)
    print(f"rrect ops: max diff {mx}, {cnt} pixels differ", flush=True)
    assert mx == 0 and want[..., 3].any()
