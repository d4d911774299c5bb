// hip_picture.cc -- the SKP reader and its playback (see hip_picture.hpp).
//
// Layout of an SKP as skity reads it (module/io/src/picture.cc:51-74,
// 249-286, module/io/src/record/record_playback.cc:417-842):
//   "skiapict", u32 version, 4 f32 cull rect, u8 1,
//   then chunks {u32 tag, u32 size, ...} up to the tag "eof ":
//     "read": the op stream (size bytes),
//     "fact": u32 count, then count packed-length factory names,
//     "tpfc": size typefaces,
//     "pctr": size sub-pictures,
//     "aray": a 4-byte-aligned buffer of sub-chunks {u32 tag, u32 size}:
//             "pnt " size paints, "pth " a count and that many paths, ...
// Ops are {u32 (op << 24 | size), [u32 size if the low 24 bits are all set],
// fields}; paints and paths are referenced by 1-based indices.
#include "hip_picture.hpp"

#include "../../include/skity_hip_picture.h"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace skity_hip {

namespace {

// DrawType (module/io/src/record/draw_type.hpp:24-121): the op numbers of
// the SKP format.
enum Op : uint32_t {
  CLIP_PATH = 1,
  CLIP_REGION = 2,
  CLIP_RECT = 3,
  CLIP_RRECT = 4,
  CONCAT = 5,
  DRAW_CLEAR = 10,
  DRAW_DATA = 11,
  DRAW_OVAL = 12,
  DRAW_PAINT = 13,
  DRAW_PATH = 14,
  DRAW_PICTURE = 15,
  DRAW_POINTS = 16,
  DRAW_RECT = 21,
  DRAW_RRECT = 22,
  RESTORE = 28,
  ROTATE = 29,
  SAVE = 30,
  SCALE = 32,
  SET_MATRIX = 33,
  SKEW = 34,
  TRANSLATE = 35,
  NOOP = 36,
  DRAW_DRRECT = 40,
  PUSH_CULL = 41,
  POP_CULL = 42,
  DRAW_PATCH = 43,
  DRAW_PICTURE_MATRIX_PAINT = 44,
  DRAW_TEXT_BLOB = 45,
  DRAW_IMAGE = 46,
  DRAW_ATLAS = 48,
  DRAW_IMAGE_NINE = 49,
  DRAW_IMAGE_RECT = 50,
  SAVE_LAYER_SAVELAYERREC = 52,
  DRAW_ANNOTATION = 53,
  DRAW_DRAWABLE = 54,
  DRAW_DRAWABLE_MATRIX = 55,
  DRAW_SHADOW_REC = 58,
  DRAW_IMAGE_LATTICE = 59,
  DRAW_ARC = 60,
  DRAW_REGION = 61,
  DRAW_VERTICES_OBJECT = 62,
  FLUSH = 63,
  DRAW_EDGEAA_IMAGE_SET = 64,
  SAVE_BEHIND = 65,
  DRAW_EDGEAA_QUAD = 66,
  DRAW_BEHIND_PAINT = 67,
  CONCAT44 = 68,
  CLIP_SHADER_IN_PAINT = 69,
  SET_M44 = 71,
  DRAW_IMAGE2 = 72,
  DRAW_IMAGE_RECT2 = 73,
  DRAW_IMAGE_LATTICE2 = 74,
  DRAW_EDGEAA_IMAGE_SET2 = 75,
  RESET_CLIP = 76,
  DRAW_SLUG = 77,
  LAST_OP = DRAW_SLUG,
};

constexpr uint32_t Tag(char a, char b, char c, char d) {
  return ((uint32_t)(uint8_t)a << 24) | ((uint32_t)(uint8_t)b << 16) |
         ((uint32_t)(uint8_t)c << 8) | (uint32_t)(uint8_t)d;
}
constexpr uint32_t kReadTag = Tag('r', 'e', 'a', 'd');
constexpr uint32_t kFactoryTag = Tag('f', 'a', 'c', 't');
constexpr uint32_t kTypefaceTag = Tag('t', 'p', 'f', 'c');
constexpr uint32_t kPictureTag = Tag('p', 'c', 't', 'r');
constexpr uint32_t kDrawableTag = Tag('d', 'r', 'a', 'w');
constexpr uint32_t kBufferTag = Tag('a', 'r', 'a', 'y');
constexpr uint32_t kPaintTag = Tag('p', 'n', 't', ' ');
constexpr uint32_t kPathTag = Tag('p', 't', 'h', ' ');
constexpr uint32_t kTextBlobTag = Tag('b', 'l', 'o', 'b');
constexpr uint32_t kSlugTag = Tag('s', 'l', 'u', 'g');
constexpr uint32_t kVerticesTag = Tag('v', 'e', 'r', 't');
constexpr uint32_t kImageTag = Tag('i', 'm', 'a', 'g');
constexpr uint32_t kEofTag = Tag('e', 'o', 'f', ' ');

// Picture versions read (module/io/src/picture_priv.hpp:90-151).
constexpr uint32_t kMinVersion = 82;
constexpr uint32_t kCurrentVersion = 109;
constexpr uint32_t kBlenderInPaintVersion = 87;

std::string g_last_error;

// A bounded little-endian cursor.  `aligned` reads (the flattened buffer and
// the op stream, ReadBuffer in module/io/src/io/memory_read.cc) advance by
// whole 4-byte words; the container stream reads are byte-exact.
struct Cursor {
  const uint8_t* base;
  const uint8_t* p;
  const uint8_t* end;
  bool ok = true;

  Cursor(const uint8_t* b, size_t n) : base(b), p(b), end(b + n) {}
  size_t Offset() const { return (size_t)(p - base); }
  size_t Left() const { return (size_t)(end - p); }
  bool Eof() const { return p >= end; }
  bool Validate(bool c) {
    if (!c && ok) {
      ok = false;
      p = end;
    }
    return ok;
  }
  // ReadBuffer::Skip: size rounded up to a word
  const uint8_t* Skip(size_t n) {
    size_t inc = (n + 3) & ~(size_t)3;
    if (!Validate(inc >= n && ok && Left() >= inc)) return nullptr;
    const uint8_t* a = p;
    p += inc;
    return a;
  }
  // byte-exact read for the container stream
  bool Read(void* dst, size_t n) {
    if (!Validate(ok && Left() >= n)) return false;
    if (n) std::memcpy(dst, p, n);
    p += n;
    return true;
  }
  uint32_t U32() {
    uint32_t v = 0;
    if (const uint8_t* a = Skip(4)) std::memcpy(&v, a, 4);
    return v;
  }
  int32_t I32() { return (int32_t)U32(); }
  float F32() {
    float v = 0;
    if (const uint8_t* a = Skip(4)) std::memcpy(&v, a, 4);
    return v;
  }
  bool Bool() {
    uint32_t v = U32();
    Validate(!(v & ~1u));
    return v != 0;
  }
  bool ReadPad32(void* dst, size_t n) {
    const uint8_t* a = Skip(n);
    if (!a) return false;
    if (n) std::memcpy(dst, a, n);  // dst may be an empty vector's null data()
    return true;
  }
};

// Color4fToColor (src/graphic/color.cc:53-59): each channel clamped to
// [0, 255] after the product and truncated to 8 bits.
Color ToColor(const float c[4]) {
  auto ch = [](float v) {
    float x = v * 255.f;
    x = x < 0.f ? 0.f : (x > 255.f ? 255.f : x);
    return (uint32_t)(uint8_t)x;
  };
  return ColorSetARGB(ch(c[3]), ch(c[0]), ch(c[1]), ch(c[2]));
}

bool AffineFrom3x3(const float m[9], Matrix* out) {
  if (m[6] != 0.f || m[7] != 0.f || m[8] != 1.f) return false;
  out->sx = m[0];
  out->kx = m[1];
  out->tx = m[2];
  out->ky = m[3];
  out->sy = m[4];
  out->ty = m[5];
  return true;
}

// A column-major 4x4 (SkM44) that maps the plane affinely: no perspective
// row and no z coupling, so its products with the canvas matrix equal the
// 2-D products term for term.
bool AffineFrom4x4(const float c[16], Matrix* out) {
  if (c[3] != 0.f || c[7] != 0.f || c[15] != 1.f || c[2] != 0.f || c[6] != 0.f ||
      c[8] != 0.f || c[9] != 0.f || c[11] != 0.f || c[14] != 0.f)
    return false;
  out->sx = c[0];
  out->ky = c[1];
  out->kx = c[4];
  out->sy = c[5];
  out->tx = c[12];
  out->ty = c[13];
  return true;
}

// ReadFromMemory<RRect> (module/io/src/io/flat/rrect_flat.cc:23-39): a
// rectangle and four radii, normalised by SetRectRadii
bool ReadRRect(Cursor& c, RRect* out) {
  float v[12];
  if (!c.ReadPad32(v, sizeof(v))) return false;
  const Point radii[4] = {{v[4], v[5]}, {v[6], v[7]}, {v[8], v[9]}, {v[10], v[11]}};
  out->SetRectRadii(Rect::MakeLTRB(v[0], v[1], v[2], v[3]), radii);
  return true;
}

}  // namespace

struct PictureParser {
  Picture* pic;

  // ReadBuffer::ReadRawFlattenable (memory_read.cc:252-310): a null entry
  // consumes one word; otherwise the factory reference and the recorded
  // size, then the object's own bytes, skipped here by that size.
  // *present is set when the entry is not null.
  bool SkipFlattenable(Cursor& c, bool* present) {
    *present = false;
    if (!pic->factories_.empty()) {
      int32_t index = c.I32();
      if (index == 0 || !c.ok) return c.ok;
      if (!c.Validate(index > 0 && (size_t)index <= pic->factories_.size())) return false;
    } else {
      if (!c.Validate(c.Left() >= 4)) return false;
      if (c.p[0] != 0) {
        // a factory named inline; this reader does not register names
        return c.Validate(false);
      }
      uint32_t index = c.U32() >> 8;
      if (index == 0) return c.ok;
      return c.Validate(false);
    }
    uint32_t size = c.U32();
    *present = true;
    c.Skip(size);
    return c.ok;
  }

  // ReadFromMemory<Paint> (module/io/src/io/flat/paint_flat.cc:136-190)
  bool ReadPaint(Cursor& c) {
    Picture::PaintEntry e;
    Paint& paint = e.paint;
    paint.SetStrokeWidth(c.F32());
    paint.SetStrokeMiter(c.F32());
    float color[4];
    if (!c.ReadPad32(color, sizeof(color))) return false;
    paint.SetColor(ToColor(color));
    uint32_t packed = c.U32();
    // unpack_v68: anti-alias (ignored by the software path), blend mode
    // (0xFF = a custom blender: the mode is left as it is), cap, join,
    // style, filter quality (skipped), flat flags
    uint32_t mode = (packed >> 8) & 0xFF;
    if (mode != 0xFF) {
      uint32_t last = static_cast<uint32_t>(BlendMode::kLuminosity);
      paint.SetBlendMode(static_cast<BlendMode>(mode < last ? mode : last));
    }
    uint32_t cap = (packed >> 16) & 3, join = (packed >> 18) & 3, style = (packed >> 20) & 3;
    if (cap > 2 || join > 2) e.supported = false;
    paint.SetStrokeCap(static_cast<Paint::Cap>(cap > 2 ? 0 : cap));
    paint.SetStrokeJoin(static_cast<Paint::Join>(join > 2 ? 0 : join));
    paint.SetStyle(static_cast<Paint::Style>(style));
    const uint32_t flat = packed >> 24;
    if (flat & 2) {
      // path effect, shader, mask filter, colour filter, [a deprecated word
      // before version 87,] image filter, [the blender from version 87]
      const bool old = pic->version_ > 0 && pic->version_ < kBlenderInPaintVersion;
      bool present = false;
      for (int k = 0; k < 4; k++) {
        if (!SkipFlattenable(c, &present)) return false;
        if (present) e.supported = false;
      }
      if (old) (void)c.U32();
      if (!SkipFlattenable(c, &present)) return false;
      if (present) e.supported = false;
      if (!old) {
        // the blend mode is in the packed word; the blender is skipped
        if (!SkipFlattenable(c, &present)) return false;
      }
    }
    if (!c.ok) return false;
    pic->paints_.push_back(std::move(e));
    return true;
  }

  // ReadFromMemory<Path> (module/io/src/io/flat/path_flat.cc:225-306)
  bool ReadPath(Cursor& c) {
    uint32_t packed = c.U32();
    if (!c.ok) return false;
    uint32_t version = packed & 0xFF;
    if (!c.Validate(version == 4 || version == 5)) return false;
    uint32_t type = (packed >> 28) & 0xF;
    uint32_t fill = (packed >> 8) & 0x3;
    Picture::PathEntry e;
    if (type == 1) {
      // a rounded-rect path: direction, 12 floats, start index
      // ReadRRectPath (path_flat.cc:104-145)
      uint32_t dir = (packed >> 26) & 0x3;
      if (!c.Validate(dir <= 1)) return false;
      RRect rr;
      if (!ReadRRect(c, &rr)) return false;
      int32_t start = c.I32();
      if (!c.Validate(start >= 0 && start <= 7)) return false;
      e.path.AddRRect(rr, dir == 0 ? Path::Direction::kCW : Path::Direction::kCCW,
                      (uint32_t)start);
      if (fill == 1) {
        e.path.SetFillType(Path::PathFillType::kEvenOdd);
      } else if (fill >= 2) {
        e.supported = false;  // inverse fill types
      }
      pic->paths_.push_back(std::move(e));
      return true;
    }
    if (!c.Validate(type == 0)) return false;
    int32_t info[3];
    if (!c.ReadPad32(info, sizeof(info))) return false;
    const int32_t pts = info[0], cnx = info[1], vbs = info[2];
    if (!c.Validate(pts >= 0 && cnx >= 0 && vbs >= 0)) return false;
    if (vbs == 0) {
      if (!c.Validate(pts == 0 && cnx == 0)) return false;
      pic->paths_.push_back(std::move(e));
      return true;
    }
    std::vector<float> points(2 * (size_t)pts), weights((size_t)cnx);
    std::vector<uint8_t> verbs((size_t)vbs);
    if (!c.ReadPad32(points.data(), points.size() * sizeof(float))) return false;
    if (!c.ReadPad32(weights.data(), weights.size() * sizeof(float))) return false;
    if (!c.ReadPad32(verbs.data(), verbs.size())) return false;
    // fill_path (path_flat.cc:147-180): the builder calls in verb order;
    // the counts are trusted as written, so they are checked here first
    size_t need_pts = 0, need_w = 0;
    for (uint8_t v : verbs) {
      need_pts += v == 0 || v == 1 ? 1 : (v == 2 || v == 3 ? 2 : (v == 4 ? 3 : 0));
      need_w += v == 3;
    }
    if (!c.Validate(need_pts <= (size_t)pts && need_w <= (size_t)cnx)) return false;
    Path& path = e.path;
    const float* q = points.data();
    const float* w = weights.data();
    for (uint8_t v : verbs) {
      switch (v) {
        case 0:
          path.MoveTo(q[0], q[1]);
          q += 2;
          break;
        case 1:
          path.LineTo(q[0], q[1]);
          q += 2;
          break;
        case 2:
          path.QuadTo(q[0], q[1], q[2], q[3]);
          q += 4;
          break;
        case 3:
          path.ConicTo(q[0], q[1], q[2], q[3], w[0]);
          q += 4;
          w++;
          break;
        case 4:
          path.CubicTo(q[0], q[1], q[2], q[3], q[4], q[5]);
          q += 6;
          break;
        case 5:
          path.Close();
          break;
        default:
          break;
      }
    }
    if (fill == 1) {
      path.SetFillType(Path::PathFillType::kEvenOdd);
    } else if (fill >= 2) {
      e.supported = false;  // inverse fill types
    }
    pic->paths_.push_back(std::move(e));
    return true;
  }

  // RecordPlayback::ParseBufferTag (record_playback.cc:741-826)
  bool ParseBuffer(Cursor& c) {
    while (!c.Eof() && c.ok) {
      uint32_t tag = c.U32();
      uint32_t size = c.U32();
      if (!c.ok) break;
      switch (tag) {
        case kPaintTag:
          for (uint32_t i = 0; i < size; i++)
            if (!ReadPaint(c)) return Fail("a paint could not be read");
          break;
        case kPathTag:
          if (size > 0) {
            int32_t count = c.I32();
            if (!c.Validate(count > 0)) return Fail("bad path count");
            for (int32_t i = 0; i < count; i++)
              if (!ReadPath(c)) return Fail("a path could not be read");
          }
          break;
        case kSlugTag:
          // the reference refuses slugs and reads on; with none there is
          // nothing to refuse
          if (size > 0) return Fail("slugs are not supported");
          break;
        case kTextBlobTag:
          if (size > 0) return Fail("text blobs are not supported");
          break;
        case kVerticesTag:
          if (size > 0) return Fail("vertices are not supported");
          break;
        case kImageTag:
          if (size > 0) return Fail("images are not supported");
          break;
        case kReadTag: {
          const uint8_t* a = c.Skip(size);
          if (!a) return Fail("truncated op stream");
          pic->ops_.assign(a, a + size);
          break;
        }
        default:
          return Fail("unsupported buffer chunk (sub-pictures, drawables)");
      }
    }
    return c.ok ? true : Fail("truncated buffer");
  }

  bool Fail(const char* why) {
    g_last_error = why;
    return false;
  }

  // Picture::MakeFromStream + RecordPlayback::ParseStream
  // (picture.cc:51-74,249-286, record_playback.cc:417-739)
  bool Parse(const uint8_t* data, size_t size) {
    Cursor s(data, size);
    char magic[8];
    uint32_t version = 0;
    float cull[4];
    uint8_t trailing = 0;
    if (!s.Read(magic, 8) || std::memcmp(magic, "skiapict", 8) != 0)
      return Fail("not an SKP (magic)");
    if (!s.Read(&version, 4) || version < kMinVersion || version > kCurrentVersion)
      return Fail("SKP version outside 82..109");
    if (!s.Read(cull, sizeof(cull)) || !s.Read(&trailing, 1) || trailing != 1)
      return Fail("SKP header");
    pic->version_ = version;
    pic->cull_ = Rect::MakeLTRB(cull[0], cull[1], cull[2], cull[3]);
    while (true) {
      uint32_t tag = 0, sz = 0;
      if (!s.Read(&tag, 4)) return Fail("truncated SKP");
      if (tag == kEofTag) break;
      if (!s.Read(&sz, 4)) return Fail("truncated SKP");
      switch (tag) {
        case kReadTag: {
          if (sz == 0 || s.Left() < sz) return Fail("bad op stream chunk");
          pic->ops_.assign(s.p, s.p + sz);
          s.p += sz;
          break;
        }
        case kFactoryTag: {
          uint32_t count = 0;
          if (!s.Read(&count, 4)) return Fail("truncated factory chunk");
          for (uint32_t i = 0; i < count; i++) {
            // ReadStream::ReadPackedUint (stream.cc:106-131)
            uint8_t b = 0;
            size_t len = 0;
            if (!s.Read(&b, 1)) return Fail("truncated factory name");
            if (b == 0xFE) {
              uint16_t v = 0;
              if (!s.Read(&v, 2)) return Fail("truncated factory name");
              len = v;
            } else if (b == 0xFF) {
              uint32_t v = 0;
              if (!s.Read(&v, 4)) return Fail("truncated factory name");
              len = v;
            } else {
              len = b;
            }
            if (s.Left() < len) return Fail("truncated factory name");
            pic->factories_.emplace_back((const char*)s.p, len);
            s.p += len;
          }
          break;
        }
        case kTypefaceTag:
          if (sz > 0) return Fail("typefaces (text) are not supported");
          break;
        case kPictureTag:
          if (sz > 0) return Fail("sub-pictures are not supported");
          break;
        case kDrawableTag:
          return Fail("drawables are not supported");
        case kBufferTag: {
          if (s.Left() < sz) return Fail("truncated buffer chunk");
          Cursor b(s.p, sz);
          s.p += sz;
          if (!ParseBuffer(b)) return false;
          break;
        }
        default:
          return Fail("unknown SKP chunk");
      }
    }
    if (pic->ops_.empty()) return Fail("no op stream");
    return true;
  }
};

const std::string& Picture::LastError() { return g_last_error; }

std::unique_ptr<Picture> Picture::MakeFromData(const uint8_t* data, size_t size) {
  g_last_error.clear();
  if (!data) {
    g_last_error = "no data";
    return nullptr;
  }
  std::unique_ptr<Picture> pic(new Picture());
  PictureParser parser{pic.get()};
  if (!parser.Parse(data, size)) return nullptr;
  return pic;
}

std::unique_ptr<Picture> Picture::MakeFromFile(const char* path) {
  std::FILE* f = std::fopen(path, "rb");
  if (!f) {
    g_last_error = "cannot open file";
    return nullptr;
  }
  std::vector<uint8_t> bytes;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) bytes.insert(bytes.end(), buf, buf + n);
  std::fclose(f);
  return MakeFromData(bytes.data(), bytes.size());
}

// Picture::PlayBack + Picture::HandleOp (picture.cc:215-247,312-1043)
void Picture::PlayBack(HIPCanvas* canvas) const {
  if (!canvas) return;
  Cursor c(ops_.data(), ops_.size());
  const int restore = canvas->Save();
  // 1-based table lookups (record_playback.cc:844-876); a bad index ends
  // playback as the reference's Validate does
  auto paint_at = [&](int32_t index) -> const PaintEntry* {
    if (!c.Validate(index > 0 && (size_t)index <= paints_.size())) return nullptr;
    return &paints_[index - 1];
  };
  auto path_at = [&](int32_t index) -> const PathEntry* {
    if (!c.Validate(index > 0 && (size_t)index <= paths_.size())) return nullptr;
    return &paths_[index - 1];
  };
  while (!c.Eof() && c.ok) {
    const size_t start = c.Offset();
    uint32_t bits = c.U32();
    uint32_t op = bits >> 24;
    size_t size = bits & 0xFFFFFF;
    if (size == 0xFFFFFF) size = c.U32();
    if (!c.Validate(size > 0 && op > 0 && op <= LAST_OP)) break;
    const size_t next = start + size;
    bool skip_to_next = false;  // fields not interpreted here: jump to the next op
    switch (op) {
      case NOOP:
      case DRAW_DATA:
      case DRAW_POINTS:          // the reference skips the draw
      case DRAW_ATLAS:           // read and not drawn by the reference
      case DRAW_ANNOTATION:
      case CLIP_SHADER_IN_PAINT:
        skip_to_next = true;
        break;
      case FLUSH:
      case RESET_CLIP:
      case PUSH_CULL:
      case POP_CULL:
        break;
      // ops on which the reference ends playback (Validate(false))
      case CLIP_REGION:
      case DRAW_DRAWABLE:
      case DRAW_DRAWABLE_MATRIX:
      case DRAW_EDGEAA_QUAD:
      case DRAW_EDGEAA_IMAGE_SET:
      case DRAW_EDGEAA_IMAGE_SET2:
      case DRAW_IMAGE_LATTICE:
      case DRAW_IMAGE_LATTICE2:
      case DRAW_PATCH:
      case DRAW_PICTURE:
      case DRAW_PICTURE_MATRIX_PAINT:
      case DRAW_REGION:
      case DRAW_SLUG:
      case DRAW_VERTICES_OBJECT:
        c.Validate(false);
        break;
      case CLIP_RECT: {
        // rect, packed op (low 4 bits: 0 = difference), restore offset
        float r[4];
        if (!c.ReadPad32(r, sizeof(r))) break;
        uint32_t packed = c.U32();
        int32_t offset = c.I32();
        if (!c.Validate(offset == 0 || ((offset & 3) == 0 && (size_t)offset >= c.Offset())))
          break;
        canvas->ClipRect(Rect::MakeLTRB(r[0], r[1], r[2], r[3]),
                         (packed & 0xF) == 0 ? HIPCanvas::ClipOp::kDifference
                                             : HIPCanvas::ClipOp::kIntersect);
        break;
      }
      case DRAW_RRECT: {
        const PaintEntry* pe = paint_at(c.I32());
        RRect rr;
        if (!pe || !ReadRRect(c, &rr)) break;
        if (pe->supported) {
          canvas->DrawRRect(rr, pe->paint);
        } else {
          canvas->MarkUnsupported();
        }
        break;
      }
      case DRAW_DRRECT: {
        // the inner rrect first, then the outer, even-odd (picture.cc:565-581)
        const PaintEntry* pe = paint_at(c.I32());
        RRect outer, inner;
        if (!pe || !ReadRRect(c, &outer) || !ReadRRect(c, &inner)) break;
        if (!pe->supported) {
          canvas->MarkUnsupported();
          break;
        }
        Path path;
        path.AddRRect(inner);
        path.AddRRect(outer);
        path.SetFillType(Path::PathFillType::kEvenOdd);
        canvas->DrawPath(path, pe->paint);
        break;
      }
      // canvas calls this backend does not draw
      case CLIP_PATH:
      case CLIP_RRECT:
      case DRAW_ARC:
      case DRAW_TEXT_BLOB:
      case DRAW_IMAGE:
      case DRAW_IMAGE2:
      case DRAW_IMAGE_NINE:
      case DRAW_IMAGE_RECT:
      case DRAW_IMAGE_RECT2:
      case DRAW_SHADOW_REC:
        canvas->MarkUnsupported();
        skip_to_next = true;
        break;
      case SAVE_LAYER_SAVELAYERREC:
        canvas->SaveLayer(Rect{0, 0, 0, 0}, Paint());
        skip_to_next = true;
        break;
      case SAVE_BEHIND:
        canvas->Save();
        skip_to_next = true;
        break;
      case CONCAT:
      case SET_MATRIX: {
        float m[9];
        if (!c.ReadPad32(m, sizeof(m))) break;
        Matrix a;
        if (!AffineFrom3x3(m, &a)) {
          canvas->MarkUnsupported();
          break;
        }
        if (op == CONCAT) {
          canvas->Concat(a);
        } else {
          canvas->SetMatrix(a);
        }
        break;
      }
      case CONCAT44:
      case SET_M44: {
        float m[16];
        if (!c.ReadPad32(m, sizeof(m))) break;
        Matrix a;
        if (!AffineFrom4x4(m, &a)) {
          canvas->MarkUnsupported();
          break;
        }
        if (op == CONCAT44) {
          canvas->Concat(a);
        } else {
          canvas->SetMatrix(a);
        }
        break;
      }
      case DRAW_CLEAR: {
        uint32_t color = c.U32();
        if (c.ok) canvas->Clear(color);
        break;
      }
      case DRAW_OVAL:
      case DRAW_RECT: {
        const PaintEntry* pe = paint_at(c.I32());
        float r[4];
        if (!pe || !c.ReadPad32(r, sizeof(r))) break;
        if (!pe->supported) {
          canvas->MarkUnsupported();
        } else if (op == DRAW_RECT) {
          canvas->DrawRect(Rect::MakeLTRB(r[0], r[1], r[2], r[3]), pe->paint);
        } else {
          canvas->DrawOval(Rect::MakeLTRB(r[0], r[1], r[2], r[3]), pe->paint);
        }
        break;
      }
      case DRAW_PAINT:
      case DRAW_BEHIND_PAINT: {
        const PaintEntry* pe = paint_at(c.I32());
        if (!pe) break;
        if (pe->supported) {
          canvas->DrawPaint(pe->paint);
        } else {
          canvas->MarkUnsupported();
        }
        break;
      }
      case DRAW_PATH: {
        const PaintEntry* pe = paint_at(c.I32());
        if (!pe) break;
        const PathEntry* qe = path_at(c.I32());
        if (!qe) break;
        if (pe->supported && qe->supported) {
          canvas->DrawPath(qe->path, pe->paint);
        } else {
          canvas->MarkUnsupported();
        }
        break;
      }
      case RESTORE:
        canvas->Restore();
        break;
      case SAVE:
        canvas->Save();
        break;
      case ROTATE: {
        float deg = c.F32();
        if (c.ok) canvas->Rotate(deg);
        break;
      }
      case SCALE:
      case SKEW:
      case TRANSLATE: {
        float x = c.F32(), y = c.F32();
        if (!c.ok) break;
        if (op == SCALE) {
          canvas->Scale(x, y);
        } else if (op == SKEW) {
          canvas->Skew(x, y);
        } else {
          canvas->Translate(x, y);
        }
        break;
      }
      default:
        // retired op numbers: the reference's default case ends playback
        c.Validate(false);
        break;
    }
    if (skip_to_next && c.ok) {
      if (!c.Validate(next >= c.Offset() && next <= ops_.size())) break;
      c.p = c.base + next;
    }
  }
  canvas->RestoreToCount(restore);
}

}  // namespace skity_hip

// ---- C binding ---------------------------------------------------------------
using skity_hip::DisplayList;
using skity_hip::HIPCanvas;
using skity_hip::PackedScene;
using skity_hip::Picture;
using skity_hip::PictureRecorder;

struct skity_hip_picture_s {
  std::unique_ptr<Picture> pic;
};
struct skity_hip_display_list_s {
  std::unique_ptr<DisplayList> list;
  std::unique_ptr<PackedScene> packed;
  bool unsupported = false;
  int width = 0, height = 0;
};

extern "C" {

int skity_hip_picture_make(const uint8_t* data, size_t size, skity_hip_picture* out) {
  if (!out) return HIPR_ERROR_INVALID_ARGUMENT;
  *out = nullptr;
  auto pic = Picture::MakeFromData(data, size);
  if (!pic) return HIPR_ERROR_NOT_SUPPORTED;
  *out = new skity_hip_picture_s{std::move(pic)};
  return HIPR_SUCCESS;
}

const char* skity_hip_picture_last_error(void) { return Picture::LastError().c_str(); }

int skity_hip_picture_info(skity_hip_picture pic, float cull_rect[4], uint32_t* version,
                           int32_t* paints, int32_t* paths) {
  if (!pic) return HIPR_ERROR_INVALID_ARGUMENT;
  const auto& r = pic->pic->CullRect();
  if (cull_rect) {
    cull_rect[0] = r.left;
    cull_rect[1] = r.top;
    cull_rect[2] = r.right;
    cull_rect[3] = r.bottom;
  }
  if (version) *version = pic->pic->Version();
  if (paints) *paints = (int32_t)pic->pic->PaintCount();
  if (paths) *paths = (int32_t)pic->pic->PathCount();
  return HIPR_SUCCESS;
}

void skity_hip_picture_destroy(skity_hip_picture pic) { delete pic; }

int skity_hip_picture_record(skity_hip_picture pic, int32_t width, int32_t height,
                             const float matrix[6], skity_hip_display_list* out) {
  if (!pic || !out || width <= 0 || height <= 0) return HIPR_ERROR_INVALID_ARGUMENT;
  PictureRecorder rec;
  rec.BeginRecording(skity_hip::Rect::MakeWH((float)width, (float)height));
  HIPCanvas* canvas = rec.GetRecordingCanvas();
  if (matrix) {
    skity_hip::Matrix m;
    m.sx = matrix[0];
    m.kx = matrix[1];
    m.tx = matrix[2];
    m.ky = matrix[3];
    m.sy = matrix[4];
    m.ty = matrix[5];
    canvas->Concat(m);
  }
  pic->pic->PlayBack(canvas);
  auto* dl = new skity_hip_display_list_s();
  dl->list = rec.FinishRecording();
  dl->width = width;
  dl->height = height;
  *out = dl;
  return HIPR_SUCCESS;
}

void skity_hip_display_list_destroy(skity_hip_display_list dl) { delete dl; }

// The display list replayed onto a fresh canvas (identity matrix), packed
// as the scene its flush would draw.  NOT_SUPPORTED when a replayed call is
// one the backend does not draw.
int skity_hip_display_list_scene(skity_hip_display_list dl, const hipr_scene** scene) {
  if (!dl || !scene) return HIPR_ERROR_INVALID_ARGUMENT;
  *scene = nullptr;
  if (!dl->packed) {
    std::vector<HIPCanvas::Draw> draws;
    dl->unsupported = !skity_hip::ReplayDraws(*dl->list, &draws);
    dl->packed.reset(new PackedScene(draws));
  }
  if (dl->unsupported) return HIPR_ERROR_NOT_SUPPORTED;
  *scene = dl->packed->scene();
  return HIPR_SUCCESS;
}

int skity_hip_display_list_draw(skity_hip_display_list dl, hipr_surface surface, int32_t mode) {
  const hipr_scene* scene = nullptr;
  int r = skity_hip_display_list_scene(dl, &scene);
  if (r != HIPR_SUCCESS) return r;
  if (!surface) return HIPR_ERROR_INVALID_ARGUMENT;
  hipr_options opt{};
  opt.mode = mode;
  return hipr_surface_draw_scene(surface, scene, &opt);
}

}  // extern "C"
