"""Test helper: glyph outlines of a TrueType font as skity_amd.canvas paths.

Reads the 'head', 'maxp', 'loca', 'glyf', 'hmtx', 'hhea' and 'cmap' (format
4) tables of a TrueType file and turns simple glyphs into move / line /
quad / close paths the way a TrueType rasteriser decomposes an outline:
consecutive off-curve points imply an on-curve midpoint, a contour that
starts off-curve starts at its last on-curve point (or the midpoint of its
first and last points), y grows downwards, and font units scale by
size / unitsPerEm.  Composite glyphs are skipped.  These are realistic glyph
inputs for the glyph-run draw; the outlines a FreeType build of the reference
would give (its 26.6 scaling) are not restated here."""
import os
import struct

from skity_amd.canvas import Path

FONT_CANDIDATES = ["/usr/share/fonts/truetype/dejavu/DejaVuSans.ttf",
                   "/usr/share/fonts/truetype/dejavu/DejaVuSerif-Bold.ttf"]


def find_font():
    for f in FONT_CANDIDATES:
        if os.path.exists(f):
            return f
    return None


class TrueTypeFont:
    def __init__(self, path):
        self.data = open(path, "rb").read()
        n, = struct.unpack_from(">H", self.data, 4)
        self.tables = {}
        for i in range(n):
            tag, _, off, ln = struct.unpack_from(">4sIII", self.data, 12 + 16 * i)
            self.tables[tag.decode("latin-1")] = (off, ln)
        head = self.tables["head"][0]
        self.units_per_em, = struct.unpack_from(">H", self.data, head + 18)
        self.loca_long, = struct.unpack_from(">h", self.data, head + 50)
        self.num_glyphs, = struct.unpack_from(">H", self.data, self.tables["maxp"][0] + 4)
        self.num_hmetrics, = struct.unpack_from(">H", self.data, self.tables["hhea"][0] + 34)
        self.cmap = self._cmap4()

    def _cmap4(self):
        base = self.tables["cmap"][0]
        n, = struct.unpack_from(">H", self.data, base + 2)
        for i in range(n):
            plat, enc, off = struct.unpack_from(">HHI", self.data, base + 4 + 8 * i)
            sub = base + off
            fmt, = struct.unpack_from(">H", self.data, sub)
            if fmt == 4 and plat in (0, 3):
                break
        else:
            return {}
        segx2, = struct.unpack_from(">H", self.data, sub + 6)
        seg = segx2 // 2
        ends = struct.unpack_from(">%dH" % seg, self.data, sub + 14)
        starts = struct.unpack_from(">%dH" % seg, self.data, sub + 16 + segx2)
        deltas = struct.unpack_from(">%dh" % seg, self.data, sub + 16 + 2 * segx2)
        ro_base = sub + 16 + 3 * segx2
        ros = struct.unpack_from(">%dH" % seg, self.data, ro_base)
        out = {}
        for s in range(seg):
            for c in range(starts[s], min(ends[s], 0x7F) + 1):
                if ros[s] == 0:
                    g = (c + deltas[s]) & 0xFFFF
                else:
                    a = ro_base + 2 * s + ros[s] + 2 * (c - starts[s])
                    g, = struct.unpack_from(">H", self.data, a)
                    if g:
                        g = (g + deltas[s]) & 0xFFFF
                out[chr(c)] = g
        return out

    def advance(self, gid, size):
        i = min(gid, self.num_hmetrics - 1)
        adv, = struct.unpack_from(">H", self.data, self.tables["hmtx"][0] + 4 * i)
        return adv * size / self.units_per_em

    def _glyph_range(self, gid):
        loca = self.tables["loca"][0]
        if self.loca_long:
            a, b = struct.unpack_from(">II", self.data, loca + 4 * gid)
        else:
            a, b = (2 * v for v in struct.unpack_from(">HH", self.data, loca + 2 * gid))
        return self.tables["glyf"][0] + a, b - a

    def glyph_path(self, gid, size):
        """Path of a simple glyph at `size` px (None for composites)."""
        off, ln = self._glyph_range(gid)
        path = Path()
        if ln == 0:
            return path
        nc, = struct.unpack_from(">h", self.data, off)
        if nc < 0:
            return None
        ends = struct.unpack_from(">%dH" % nc, self.data, off + 10)
        npts = ends[-1] + 1 if nc else 0
        ilen, = struct.unpack_from(">H", self.data, off + 10 + 2 * nc)
        p = off + 12 + 2 * nc + ilen
        flags = []
        while len(flags) < npts:
            f = self.data[p]
            p += 1
            flags.append(f)
            if f & 8:
                r = self.data[p]
                p += 1
                flags += [f] * r
        flags = flags[:npts]

        def coords(short_bit, same_bit):
            nonlocal p
            v, out = 0, []
            for f in flags:
                if f & short_bit:
                    d = self.data[p]
                    p += 1
                    v += d if f & same_bit else -d
                elif not f & same_bit:
                    d, = struct.unpack_from(">h", self.data, p)
                    p += 2
                    v += d
                out.append(v)
            return out

        xs = coords(2, 16)
        ys = coords(4, 32)
        k = size / self.units_per_em
        pts = [(x * k, -y * k, bool(f & 1)) for x, y, f in zip(xs, ys, flags)]
        start = 0
        for e in ends:
            c = pts[start:e + 1]
            start = e + 1
            if not c:
                continue
            # rotate so that the contour starts on-curve (or at an implied point)
            on = [i for i, q in enumerate(c) if q[2]]
            if on:
                i0 = on[0]
                c = c[i0:] + c[:i0]
                first = c[0][:2]
                rest = c[1:] + [c[0]]
            else:
                first = ((c[0][0] + c[-1][0]) / 2, (c[0][1] + c[-1][1]) / 2)
                rest = c + [(first[0], first[1], True)]
            path.move_to(*first)
            ctrl = None
            for x, y, is_on in rest:
                if is_on:
                    if ctrl is None:
                        path.line_to(x, y)
                    else:
                        path.quad_to(ctrl[0], ctrl[1], x, y)
                        ctrl = None
                else:
                    if ctrl is not None:
                        mx, my = (ctrl[0] + x) / 2, (ctrl[1] + y) / 2
                        path.quad_to(ctrl[0], ctrl[1], mx, my)
                    ctrl = (x, y)
            if ctrl is not None:
                path.quad_to(ctrl[0], ctrl[1], *first)
            path.close()
        return path

    def layout(self, text, size, x, y):
        """(glyph paths, xs, ys) of a line of text starting at (x, y)."""
        paths, xs, ys = [], [], []
        for ch in text:
            g = self.cmap.get(ch, 0)
            gp = self.glyph_path(g, size)
            if gp is not None and gp.verbs:
                paths.append(gp)
                xs.append(x)
                ys.append(y)
            x += self.advance(g, size)
        return paths, xs, ys
