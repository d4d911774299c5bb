// hip_canvas.hpp -- C++ host front end of the MI355X DrawPath backend.
//
// The same shapes as skity's public API for the DrawPath path
// (include/skity/graphic/path.hpp, paint.hpp, include/skity/effect/
// shader.hpp, mask_filter.hpp, include/skity/render/canvas.hpp), so code
// written against skity's Canvas::DrawPath reads the same here:
//
//   auto surface = skity_hip::HIPSurface::Make(ctx, 48, 48);
//   skity_hip::HIPCanvas* canvas = surface->LockCanvas();
//   skity_hip::Paint paint;
//   paint.SetStyle(skity_hip::Paint::kStrokeThenFill_Style);
//   canvas->DrawPath(path, paint);
//   surface->Flush();                     // one hipr_surface_draw_scene call
//   skity_hip::Color c = surface->GetPixel(12, 20);
//
// A HIPCanvas records the DrawPath calls in order and Flush() packs them into
// one hipr_scene (include/skity_hip.h) for the device.  Only the calls of the
// DrawPath path exist; the library refuses what it cannot draw exactly
// (hipr_result != 0) and never falls back to a CPU rasteriser.
#ifndef SKITY_HIP_HOST_HIP_CANVAS_HPP_
#define SKITY_HIP_HOST_HIP_CANVAS_HPP_

#include <cstdint>
#include <memory>
#include <vector>

#include "../../include/skity_hip.h"

namespace skity_hip {

// include/skity/graphic/color.hpp: 0xAARRGGBB
using Color = uint32_t;
constexpr Color Color_TRANSPARENT = 0x00000000;
constexpr Color Color_BLACK = 0xFF000000;
constexpr Color Color_WHITE = 0xFFFFFFFF;
constexpr Color Color_RED = 0xFFFF0000;
constexpr Color Color_GREEN = 0xFF00FF00;
constexpr Color Color_BLUE = 0xFF0000FF;
inline constexpr Color ColorSetARGB(uint32_t a, uint32_t r, uint32_t g, uint32_t b) {
  return (a << 24) | (r << 16) | (g << 8) | b;
}
inline constexpr uint32_t ColorGetA(Color c) { return (c >> 24) & 0xFF; }
inline constexpr uint32_t ColorGetR(Color c) { return (c >> 16) & 0xFF; }
inline constexpr uint32_t ColorGetG(Color c) { return (c >> 8) & 0xFF; }
inline constexpr uint32_t ColorGetB(Color c) { return c & 0xFF; }

struct Color4f {
  float r, g, b, a;
};
// Color4fFromColor (src/graphic/color.cc): channel / 255
Color4f Color4fFromColor(Color c);

struct Point {
  float x, y;
};

struct Rect {
  float left, top, right, bottom;
  static Rect MakeLTRB(float l, float t, float r, float b) { return Rect{l, t, r, b}; }
  static Rect MakeXYWH(float x, float y, float w, float h) { return Rect{x, y, x + w, y + h}; }
  static Rect MakeWH(float w, float h) { return Rect{0, 0, w, h}; }
  float CenterX() const { return 0.5f * (right + left); }
  float CenterY() const { return 0.5f * (top + bottom); }
};

// RRect (include/skity/geometry/rrect.hpp, src/geometry/rrect.cc): a sorted
// rectangle and four corner radii {upper-left, upper-right, lower-right,
// lower-left}, normalised as the reference normalises them.
class RRect {
 public:
  enum Type { kEmpty, kRect, kOval, kSimple, kNinePatch, kComplex };
  enum Corner { kUpperLeft, kUpperRight, kLowerRight, kLowerLeft };
  void SetRect(const Rect& rect);
  void SetOval(const Rect& oval);
  void SetRectXY(const Rect& rect, float x_rad, float y_rad);
  void SetRectRadii(const Rect& rect, const Point radii[4]);
  static RRect MakeRectXY(const Rect& rect, float x_rad, float y_rad) {
    RRect r;
    r.SetRectXY(rect, x_rad, y_rad);
    return r;
  }
  static RRect MakeOval(const Rect& oval) {
    RRect r;
    r.SetOval(oval);
    return r;
  }
  Type GetType() const { return type_; }
  bool IsEmpty() const { return type_ == kEmpty; }
  bool IsRect() const { return type_ == kRect; }
  bool IsOval() const { return type_ == kOval; }
  const Rect& GetBounds() const { return rect_; }
  Point Radii(Corner c) const { return radii_[c]; }

 private:
  bool InitializeRect(const Rect& rect);
  bool ScaleRadii();
  void ComputeType();
  bool RadiiValid() const;
  Rect rect_{0, 0, 0, 0};
  Point radii_[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  Type type_ = kEmpty;
};

// Path storage and builder rules of src/graphic/path.cc.
class Path {
 public:
  enum class Verb : uint8_t { kMove, kLine, kQuad, kConic, kCubic, kClose };
  enum class PathFillType { kWinding, kEvenOdd };
  enum class Direction { kCW, kCCW };

  Path& MoveTo(float x, float y);
  Path& LineTo(float x, float y);
  Path& QuadTo(float x1, float y1, float x2, float y2);
  Path& ConicTo(float x1, float y1, float x2, float y2, float weight);
  Path& CubicTo(float x1, float y1, float x2, float y2, float x3, float y3);
  Path& Close();
  Path& AddRect(const Rect& rect, Direction dir = Direction::kCW, uint32_t start = 0);
  Path& AddOval(const Rect& oval, Direction dir = Direction::kCW, uint32_t start = 1);
  Path& AddCircle(float x, float y, float radius, Direction dir = Direction::kCW);
  // Path::AddRRect (path.cc:881-934): a rect or an oval when the radii
  // collapse, otherwise lines and conics around the eight radii points
  Path& AddRRect(const RRect& rrect, Direction dir = Direction::kCW);
  Path& AddRRect(const RRect& rrect, Direction dir, uint32_t start_index);
  Path& AddRoundRect(const Rect& rect, float rx, float ry, Direction dir = Direction::kCW);
  void SetFillType(PathFillType t) { fill_type_ = t; }
  PathFillType GetFillType() const { return fill_type_; }
  bool IsEmpty() const { return verbs_.empty(); }

  const std::vector<uint8_t>& verbs() const { return verbs_; }
  const std::vector<Point>& points() const { return points_; }
  const std::vector<float>& weights() const { return weights_; }

 private:
  void InjectMoveToIfNeeded();
  std::vector<uint8_t> verbs_;
  std::vector<Point> points_;
  std::vector<float> weights_;
  PathFillType fill_type_ = PathFillType::kWinding;
  int64_t last_move_to_index_ = -1;  // ~index after Close
};

// An affine Matrix (include/skity/geometry/matrix.hpp) restricted to the
// plane: x' = (scale_x x + skew_x y) + trans_x, y' = (skew_y x + scale_y y)
// + trans_y.  Products follow src/geometry/matrix.cc:450-458 (identity
// operands short-cut, otherwise glm's column sums, every operation rounded).
struct Matrix {
  float sx = 1, kx = 0, tx = 0, ky = 0, sy = 1, ty = 0;
  static Matrix Translate(float dx, float dy);
  static Matrix Scale(float x, float y);
  static Matrix RotateDeg(float degrees, float px = 0, float py = 0);
  static Matrix Skew(float kx, float ky);
  bool IsIdentity() const {
    return sx == 1 && kx == 0 && tx == 0 && ky == 0 && sy == 1 && ty == 0;
  }
  Matrix operator*(const Matrix& b) const;
};

// BlendMode (include/skity/graphic/blend_mode.hpp:45-80)
enum class BlendMode {
  kClear, kSrc, kDst, kSrcOver, kDstOver, kSrcIn, kDstIn, kSrcOut, kDstOut, kSrcATop,
  kDstATop, kXor, kPlus, kModulate, kScreen, kOverlay, kDarken, kLighten, kColorDodge,
  kColorBurn, kHardLight, kSoftLight, kDifference, kExclusion, kMultiply, kHue,
  kSaturation, kColor, kLuminosity,
  kDefault = kSrcOver
};

enum class TileMode { kClamp, kRepeat, kMirror, kDecal };

// Gradient shaders (include/skity/effect/shader.hpp:88-146), premultiply
// flag 0 (interpolation of unpremultiplied colours).
// An immutable RGBA8 image (Image::MakeImage of a Pixmap): the pixels are
// copied; alpha_type is HIPR_ALPHA_PREMUL, HIPR_ALPHA_OPAQUE or
// HIPR_ALPHA_UNPREMUL.
class Image {
 public:
  static std::shared_ptr<Image> MakeImage(const uint8_t* rgba, int width, int height,
                                          int alpha_type = HIPR_ALPHA_PREMUL);
  int Width() const { return width_; }
  int Height() const { return height_; }
  int AlphaType() const { return alpha_type_; }
  const uint8_t* Pixels() const { return pixels_.data(); }

 private:
  int width_ = 0, height_ = 0, alpha_type_ = HIPR_ALPHA_PREMUL;
  std::vector<uint8_t> pixels_;
};

enum class FilterMode { kNearest = HIPR_FILTER_NEAREST, kLinear = HIPR_FILTER_LINEAR };
struct SamplingOptions {
  FilterMode filter = FilterMode::kNearest;
};

class Shader {
 public:
  // Shader::MakeShader(image, sampling, tile_x, tile_y, local_matrix): a
  // pixmap shader (include/skity/effect/shader.hpp).
  static std::shared_ptr<Shader> MakeShader(std::shared_ptr<Image> image,
                                            const SamplingOptions& sampling, TileMode tile_x,
                                            TileMode tile_y, const Matrix* local = nullptr);
  static std::shared_ptr<Shader> MakeLinear(const Point pts[2], const Color4f colors[],
                                            const float pos[], int count,
                                            TileMode tile_mode = TileMode::kClamp);
  static std::shared_ptr<Shader> MakeRadial(const Point& center, float radius,
                                            const Color4f colors[], const float pos[],
                                            int count, TileMode tile_mode = TileMode::kClamp);
  static std::shared_ptr<Shader> MakeSweep(float cx, float cy, float start_angle,
                                           float end_angle, const Color4f colors[],
                                           const float pos[], int count,
                                           TileMode tile_mode = TileMode::kClamp);
  // Concentric circles with one radius 0 become a radial gradient (colours
  // reversed when the end radius is 0); equal circles give no shader.
  static std::shared_ptr<Shader> MakeTwoPointConical(const Point& start, float start_radius,
                                                     const Point& end, float end_radius,
                                                     const Color4f colors[], const float pos[],
                                                     int count,
                                                     TileMode tile_mode = TileMode::kClamp);
  // {scale_x, skew_x, trans_x, skew_y, scale_y, trans_y}
  void SetLocalMatrix(const float m[6]);
  const hipr_gradient& info() const { return info_; }
  // pixmap shaders: the image and the shader record (its image index is set
  // when a canvas packs it)
  const std::shared_ptr<Image>& image() const { return image_; }
  const hipr_image_shader& image_info() const { return image_info_; }

 private:
  hipr_gradient info_{};
  std::shared_ptr<Image> image_;
  hipr_image_shader image_info_{};
};

enum class BlurStyle {
  kNormal = HIPR_BLUR_NORMAL,
  kSolid = HIPR_BLUR_SOLID,
  kOuter = HIPR_BLUR_OUTER,
  kInner = HIPR_BLUR_INNER
};

class MaskFilter {
 public:
  static std::shared_ptr<MaskFilter> MakeBlur(BlurStyle style, float radius);
  float radius() const { return radius_; }
  BlurStyle style() const { return style_; }

 private:
  BlurStyle style_ = BlurStyle::kNormal;
  float radius_ = 0;
};

// ColorFilters (include/skity/effect/color_filter.hpp:15-37): the factories
// return nullptr where the reference does (src/effect/color_filter.cc:64-106).
class ColorFilter {
 public:
  const hipr_color_filter& info() const { return info_; }

 private:
  friend class ColorFilters;
  hipr_color_filter info_{};
};

class ColorFilters {
 public:
  static std::shared_ptr<ColorFilter> Compose(std::shared_ptr<ColorFilter> outer,
                                              std::shared_ptr<ColorFilter> inner);
  static std::shared_ptr<ColorFilter> Blend(Color color, BlendMode mode);
  static std::shared_ptr<ColorFilter> Matrix(const float row_major[20]);
  static std::shared_ptr<ColorFilter> LinearToSRGBGamma();
  static std::shared_ptr<ColorFilter> SRGBToLinearGamma();
};

class Paint {
 public:
  enum Style { kFill_Style, kStroke_Style, kStrokeAndFill_Style, kStrokeThenFill_Style };
  enum Cap { kButt_Cap, kRound_Cap, kSquare_Cap };
  enum Join { kMiter_Join, kRound_Join, kBevel_Join };

  void SetStyle(Style s) { style_ = s; }
  Style GetStyle() const { return style_; }
  void SetStrokeWidth(float w) { width_ = w; }
  void SetStrokeMiter(float m) { miter_ = m; }
  void SetStrokeCap(Cap c) { cap_ = c; }
  void SetStrokeJoin(Join j) { join_ = j; }
  void SetColor(Color c);  // fill and stroke colour (paint.cc:80-84)
  void SetFillColor(Color c) { fill_ = Color4fFromColor(c); }
  void SetStrokeColor(Color c) { stroke_ = Color4fFromColor(c); }
  void SetFillColor(float r, float g, float b, float a) { fill_ = Color4f{r, g, b, a}; }
  void SetStrokeColor(float r, float g, float b, float a) { stroke_ = Color4f{r, g, b, a}; }
  void SetAlphaF(float a);  // both colours (paint.cc:100-104)
  void SetAntiAlias(bool) {}  // the software path always anti-aliases
  void SetShader(std::shared_ptr<Shader> s) { shader_ = std::move(s); }
  void SetMaskFilter(std::shared_ptr<MaskFilter> m) { mask_ = std::move(m); }
  void SetBlendMode(BlendMode m) { blend_ = m; }
  BlendMode GetBlendMode() const { return blend_; }
  void SetColorFilter(std::shared_ptr<ColorFilter> f) { color_filter_ = std::move(f); }
  const std::shared_ptr<ColorFilter>& GetColorFilter() const { return color_filter_; }

  hipr_paint Pack(int32_t gradient_index) const;
  const std::shared_ptr<Shader>& shader() const { return shader_; }
  const std::shared_ptr<MaskFilter>& GetMaskFilter() const { return mask_; }

 private:
  Style style_ = kFill_Style;
  float width_ = 1.f, miter_ = 4.f;
  Cap cap_ = kButt_Cap;
  Join join_ = kMiter_Join;
  Color4f fill_{0, 0, 0, 1}, stroke_{0, 0, 0, 1};
  std::shared_ptr<Shader> shader_;
  std::shared_ptr<MaskFilter> mask_;
  std::shared_ptr<ColorFilter> color_filter_;
  BlendMode blend_ = BlendMode::kSrcOver;
};

class HIPSurface;
class DisplayList;
class PictureRecorder;

// Records DrawPath / DrawPaint calls (include/skity/render/canvas.hpp) with
// the canvas matrix of each, and the shape helpers that the reference
// dispatches to OnDrawPath.  The matrix stack follows Canvas::Save/Restore
// and the pre-concatenation of src/render/canvas_state.cc:18-45.
//
// A canvas is either a surface's canvas (Flush draws on the device) or a
// PictureRecorder's recording canvas (every call is appended to a
// DisplayList, as RecordingCanvas does, src/recorder/recording_canvas.cc).
class HIPCanvas {
 public:
  int Save();
  void Restore();
  int GetSaveCount() const { return (int)stack_.size(); }
  void RestoreToCount(int count);
  void Translate(float dx, float dy) { Concat(Matrix::Translate(dx, dy)); }
  void Scale(float sx, float sy) { Concat(Matrix::Scale(sx, sy)); }
  void Rotate(float degrees) { Concat(Matrix::RotateDeg(degrees)); }
  void Rotate(float degrees, float px, float py) { Concat(Matrix::RotateDeg(degrees, px, py)); }
  void Skew(float sx, float sy) { Concat(Matrix::Skew(sx, sy)); }
  void Concat(const Matrix& m);
  void SetMatrix(const Matrix& m);
  void ResetMatrix() { SetMatrix(Matrix()); }
  Matrix GetTotalMatrix() const { return stack_.back(); }

  // Canvas::ClipRect (canvas.cc:130-135,463-476, sw_canvas.cc:297-313): an
  // intersect clip narrows the canvas's global clip bounds (kept per save
  // level), the scan clip of later path draws.  Under a matrix with skew or
  // rotation, or as a difference, the software canvas also clips span lists
  // (SWCanvas::State::FindSpan, sw_canvas.cc:158-265), and so do path clips:
  // those, and save layers, are not drawn by this backend: the next Flush
  // returns HIPR_ERROR_NOT_SUPPORTED (and draws nothing) instead of drawing
  // something other than the software canvas would.  The same holds for any
  // other call the backend cannot draw exactly (MarkUnsupported).
  enum class ClipOp { kDifference, kIntersect };
  void ClipRect(const Rect& rect, ClipOp op = ClipOp::kIntersect);
  void ClipPath(const Path&, ClipOp = ClipOp::kIntersect) { MarkUnsupported(); }
  // the global clip bounds ({-1e9, -1e9, 1e9, 1e9} when nothing clips)
  Rect GetGlobalClipBounds() const { return clips_.back(); }
  void SaveLayer(const Rect&, const Paint&) {
    Save();
    MarkUnsupported();
  }
  void MarkUnsupported();

  void DrawPaint(const Paint& paint);
  // Canvas::DrawColor / Clear (canvas.cc:241-251): DrawPaint with a fill
  // colour and a blend mode (kSrc for Clear).
  void DrawColor(Color color, BlendMode mode = BlendMode::kSrcOver);
  void Clear(Color color) { DrawColor(color, BlendMode::kSrc); }
  void DrawPath(const Path& path, const Paint& paint);
  void DrawRect(const Rect& rect, const Paint& paint);
  void DrawOval(const Rect& oval, const Paint& paint);
  void DrawCircle(float cx, float cy, float radius, const Paint& paint);
  // Canvas::DrawRRect / DrawRoundRect / DrawDRRect (canvas.cc:210-235,
  // 424-457): the paths AddRRect builds (the pair even-odd for DRRect)
  void DrawRRect(const RRect& rrect, const Paint& paint);
  void DrawRoundRect(const Rect& rect, float rx, float ry, const Paint& paint);
  void DrawDRRect(const RRect& outer, const RRect& inner, const Paint& paint);
  void DrawLine(float x0, float y0, float x1, float y1, const Paint& paint);
  // Canvas::DrawGlyphs as the software canvas draws a glyph run
  // (SWCanvas::DrawGlyphsInternal / FillGlyphs / StrokeGlyphs,
  // sw_canvas.cc:542-639): glyph_paths[k] is the outline Font::LoadGlyphPath
  // gives for glyph k; each is drawn under the canvas matrix times
  // Translate(position_x[k], position_y[k]), all fills then all strokes (or
  // the reverse for kStrokeThenFill).  Runs whose paint has a mask filter
  // (one layer for the whole run) or a shader (mapped by the canvas matrix
  // alone) are not drawn by this backend: the next flush returns
  // HIPR_ERROR_NOT_SUPPORTED.
  void DrawGlyphs(uint32_t count, const Path* const glyph_paths[], const float position_x[],
                  const float position_y[], const Paint& paint);
  // Draws everything recorded since the last flush; returns the hipr_result.
  int Flush();
  size_t PendingDraws() const { return draws_.size(); }

  // One draw with the canvas matrix it was made under.
  struct Draw {
    Path path;
    Paint paint;
    Matrix matrix;
    uint8_t kind;  // HIPR_DRAW_*
    Rect clip{-1e9f, -1e9f, 1e9f, 1e9f};  // scan clip (kMaxCullRect: none)
    bool unclipped = false;  // glyph rasters ignore the clip bounds
  };

 private:
  friend class HIPSurface;
  friend class PictureRecorder;
  friend class DisplayList;
  friend bool ReplayDraws(const DisplayList& list, std::vector<Draw>* draws);
  explicit HIPCanvas(HIPSurface* s) : surface_(s) {}
  // d.matrix is the canvas matrix times `local`; a recording keeps `local`
  // and replays it under the replaying canvas's matrix.
  void AddDraw(Draw d, const Matrix& local = Matrix());
  HIPSurface* surface_;
  DisplayList* recording_ = nullptr;  // set on a recording canvas
  bool unsupported_ = false;
  std::vector<Draw> draws_;
  std::vector<Matrix> stack_{Matrix()};
  std::vector<Rect> clips_{Rect{-1e9f, -1e9f, 1e9f, 1e9f}};
};

// The flat arrays of a hipr_scene built from a list of draws (what a flush
// hands to hipr_surface_draw_scene); scene() points into them.
class PackedScene {
 public:
  explicit PackedScene(const std::vector<HIPCanvas::Draw>& draws);
  const hipr_scene* scene() const { return &scene_; }

 private:
  std::vector<uint8_t> verbs_, fills_, kinds_;
  std::vector<float> pts_, ws_, matrices_, clips_;
  std::vector<int32_t> vo_, po_, wo_;
  std::vector<hipr_paint> paints_;
  std::vector<hipr_gradient> grads_;
  std::vector<hipr_color_filter> filters_;
  std::vector<hipr_image_shader> shaders_;
  std::vector<hipr_image> images_;
  hipr_scene scene_{};
};

// A recorded list of canvas calls (include/skity/recorder/display_list.hpp:
// 56-77): the matrix and save/restore calls, the draws (shape helpers as the
// paths they build) and the calls this backend cannot draw.  Draw replays
// them onto a canvas in order (display_list.cc:330-343).
class DisplayList {
 public:
  void Draw(HIPCanvas* canvas) const;
  uint32_t OpCount() const { return (uint32_t)ops_.size(); }
  bool Empty() const { return ops_.empty(); }
  const Rect& GetBounds() const { return bounds_; }

 private:
  friend class HIPCanvas;
  friend class PictureRecorder;
  enum class OpKind : uint8_t {
    kSave, kRestore, kConcat, kSetMatrix, kDraw, kClipRect, kUnsupported
  };
  struct Op {
    OpKind kind;
    Matrix matrix;
    uint32_t draw;  // index into draws_ for kDraw; for kClipRect 1 = difference
    Rect rect{0, 0, 0, 0};
  };
  std::vector<Op> ops_;
  std::vector<HIPCanvas::Draw> draws_;
  Rect bounds_{0, 0, 0, 0};
};

// The draws a display list makes when replayed onto a fresh canvas (identity
// matrix), with the matrices they are made under; false when it makes a call
// this backend does not draw.
bool ReplayDraws(const DisplayList& list, std::vector<HIPCanvas::Draw>* draws);

// PictureRecorder (include/skity/recorder/picture_recorder.hpp): a recording
// canvas whose calls become a DisplayList.
class PictureRecorder {
 public:
  PictureRecorder();
  ~PictureRecorder();
  void BeginRecording(const Rect& bounds);
  HIPCanvas* GetRecordingCanvas() { return canvas_.get(); }
  std::unique_ptr<DisplayList> FinishRecording();

 private:
  std::unique_ptr<HIPCanvas> canvas_;
  std::unique_ptr<DisplayList> list_;
};

// A device RGBA8 target (the GPUSurface pattern of
// include/skity/gpu/gpu_surface.hpp:95-113), premultiplied RGBA by default;
// the alpha and colour types of skity's Bitmap are HIPR_ALPHA_* / HIPR_COLOR_*.
class HIPSurface {
 public:
  static std::unique_ptr<HIPSurface> Make(hipr_context ctx, int width, int height,
                                          int alpha_type = HIPR_ALPHA_PREMUL,
                                          int color_type = HIPR_COLOR_RGBA);
  ~HIPSurface();
  HIPCanvas* LockCanvas(bool clear = true);
  int Flush();
  int Width() const { return width_; }
  int Height() const { return height_; }
  // 4 bytes per pixel in the colour type's order, row-major, width*4 bytes
  // per row.
  int ReadPixels(void* dst, size_t bytes);
  // skity Color (ARGB) of one pixel, read back from the device.
  Color GetPixel(int x, int y);
  hipr_surface handle() const { return surface_; }
  int last_result() const { return last_result_; }

 private:
  HIPSurface() = default;
  hipr_surface surface_ = nullptr;
  int width_ = 0, height_ = 0;
  int color_type_ = HIPR_COLOR_RGBA;
  std::unique_ptr<HIPCanvas> canvas_;
  std::vector<uint8_t> readback_;
  bool readback_valid_ = false;
  int last_result_ = 0;
  friend class HIPCanvas;
};

}  // namespace skity_hip

#endif  // SKITY_HIP_HOST_HIP_CANVAS_HPP_
