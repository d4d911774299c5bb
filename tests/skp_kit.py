"""Test helpers for SKP pictures: a writer of small synthetic SKPs in the
layout skity reads (module/io/src/picture.cc:51-74,249-286,
module/io/src/record/record_playback.cc:417-842, paint_flat.cc:136-190,
path_flat.cc:225-306) and a reader of 8-bit RGBA PNG files (the golden
images)."""
import os
import struct
import zlib

import numpy as np

from skity_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "skp")

# DrawType numbers (module/io/src/record/draw_type.hpp:24-121)
OPS = dict(CLIP_PATH=1, CLIP_RECT=3, CLIP_RRECT=4, CONCAT=5, DRAW_CLEAR=10, DRAW_OVAL=12, DRAW_PAINT=13,
           DRAW_PATH=14, DRAW_PICTURE=15, DRAW_POINTS=16, DRAW_RECT=21, DRAW_RRECT=22,
           RESTORE=28, ROTATE=29, SAVE=30, SCALE=32, SET_MATRIX=33, SKEW=34, TRANSLATE=35,
           NOOP=36, DRAW_DRRECT=40, DRAW_TEXT_BLOB=45, SAVE_LAYER_SAVELAYERREC=52, CONCAT44=68,
           SET_M44=71)


def _tag(s):
    return struct.pack("<I", int.from_bytes(s.encode(), "big"))


def paint_bytes(color=(0, 0, 0, 1), style=abi.STYLE_FILL, width=1.0, miter=4.0,
                cap=abi.CAP_BUTT, join=abi.JOIN_MITER, blend=abi.BLEND_SRC_OVER, effects=()):
    """A flattened paint; effects: flattenable entries after the packed word
    (0 = null, or (factory_index, payload bytes))."""
    flat = 2 if effects else 0
    packed = 1 | (blend << 8) | (cap << 16) | (join << 18) | (style << 20) | (flat << 24)
    b = struct.pack("<6fI", width, miter, *color, packed)
    for e in effects:
        if e == 0:
            b += struct.pack("<i", 0)
        else:
            idx, payload = e
            b += struct.pack("<iI", idx, len(payload)) + payload
    return b


def path_bytes(verbs, points, weights=(), fill=0):
    """A general serialized path (version 5): verbs are Path::Verb values."""
    b = struct.pack("<I3i", (fill << 8) | 5, len(points), len(weights), len(verbs))
    b += b"".join(struct.pack("<2f", *p) for p in points)
    b += b"".join(struct.pack("<f", w) for w in weights)
    vb = bytes(verbs)
    b += vb + b"\0" * ((-len(vb)) % 4)
    return b


def op_bytes(name, *fields):
    """One op: fields are ('i', v), ('f', v), ('I', v) or raw bytes."""
    body = b""
    for f in fields:
        body += f if isinstance(f, bytes) else struct.pack("<" + f[0], f[1])
    return struct.pack("<I", (OPS[name] << 24) | (4 + len(body))) + body


def skp_bytes(ops, paints=(), paths=(), cull=(0, 0, 256, 256), version=109, factories=()):
    out = b"skiapict" + struct.pack("<I4f", version, *cull) + b"\x01"
    stream = b"".join(ops)
    out += _tag("read") + struct.pack("<I", len(stream)) + stream
    fact = struct.pack("<I", len(factories))
    for name in factories:
        fact += bytes([len(name)]) + name.encode()
    out += _tag("fact") + struct.pack("<I", len(fact)) + fact
    out += _tag("tpfc") + struct.pack("<I", 0)
    buf = b""
    if paints:
        buf += _tag("pnt ") + struct.pack("<I", len(paints)) + b"".join(paints)
    if paths:
        buf += _tag("pth ") + struct.pack("<Ii", 1, len(paths)) + b"".join(paths)
    out += _tag("aray") + struct.pack("<I", len(buf)) + buf
    out += _tag("eof ")
    return out


def read_png_rgba(path):
    """An 8-bit RGBA, non-interlaced PNG as an (H, W, 4) uint8 array."""
    d = open(path, "rb").read()
    assert d[:8] == b"\x89PNG\r\n\x1a\n"
    o, idat = 8, b""
    while o < len(d):
        n, = struct.unpack(">I", d[o:o + 4])
        t, body = d[o + 4:o + 8], d[o + 8:o + 8 + n]
        o += 12 + n
        if t == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif t == b"IDAT":
            idat += body
    assert depth == 8 and ctype == 6 and interlace == 0
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    out = np.zeros((h, 4 * w), np.int32)
    prev = np.zeros(4 * w, np.int32)
    for y in range(h):
        f, line = raw[y, 0], raw[y, 1:].astype(np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 255
        else:
            cur = np.zeros(4 * w, np.int32)
            for x in range(4 * w):
                a = cur[x - 4] if x >= 4 else 0
                b = prev[x]
                c = prev[x - 4] if x >= 4 else 0
                if f == 1:
                    p = a
                elif f == 3:
                    p = (a + b) >> 1
                else:  # Paeth
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
                cur[x] = (line[x] + p) & 255
        out[y] = cur
        prev = cur
    return out.reshape(h, w, 4).astype(np.uint8)


from skity_amd.picture import Recording, lib as canvas_lib  # noqa: E402,F401


def tiger_bytes():
    return open(os.path.join(GOLDEN, "tiger.skp"), "rb").read()


def tiger_golden():
    return read_png_rgba(os.path.join(GOLDEN, "tiger_coverage_aa.png"))
