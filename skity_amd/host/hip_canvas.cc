// hip_canvas.cc -- C++ host front end over the C ABI (include/skity_hip.h).
#include "hip_canvas.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace skity_hip {

namespace {
constexpr float kRoot2Over2 = 0.707106781f;  // FloatRoot2Over2
uint8_t V(Path::Verb v) { return static_cast<uint8_t>(v); }
}  // namespace

Color4f Color4fFromColor(Color c) {
  return Color4f{ColorGetR(c) / 255.f, ColorGetG(c) / 255.f, ColorGetB(c) / 255.f,
                 ColorGetA(c) / 255.f};
}

// ---- Path (src/graphic/path.cc:494-560, 815-859, 1001-1035) ----------------
Path& Path::MoveTo(float x, float y) {
  if (!verbs_.empty() && verbs_.back() == V(Verb::kMove)) {
    points_.back() = Point{x, y};
  } else {
    last_move_to_index_ = (int64_t)points_.size();
    verbs_.push_back(V(Verb::kMove));
    points_.push_back(Point{x, y});
  }
  return *this;
}

void Path::InjectMoveToIfNeeded() {
  if (last_move_to_index_ < 0) {
    Point p{0, 0};
    if (!points_.empty()) p = points_[(size_t)~last_move_to_index_];
    MoveTo(p.x, p.y);
  }
}

Path& Path::LineTo(float x, float y) {
  InjectMoveToIfNeeded();
  verbs_.push_back(V(Verb::kLine));
  points_.push_back(Point{x, y});
  return *this;
}

Path& Path::QuadTo(float x1, float y1, float x2, float y2) {
  InjectMoveToIfNeeded();
  verbs_.push_back(V(Verb::kQuad));
  points_.push_back(Point{x1, y1});
  points_.push_back(Point{x2, y2});
  return *this;
}

Path& Path::ConicTo(float x1, float y1, float x2, float y2, float w) {
  if (!(w > 0)) return LineTo(x2, y2);
  if (std::isinf(w)) {
    LineTo(x1, y1);
    return LineTo(x2, y2);
  }
  if (w == 1.f) return QuadTo(x1, y1, x2, y2);
  InjectMoveToIfNeeded();
  verbs_.push_back(V(Verb::kConic));
  weights_.push_back(w);
  points_.push_back(Point{x1, y1});
  points_.push_back(Point{x2, y2});
  return *this;
}

Path& Path::CubicTo(float x1, float y1, float x2, float y2, float x3, float y3) {
  InjectMoveToIfNeeded();
  verbs_.push_back(V(Verb::kCubic));
  points_.push_back(Point{x1, y1});
  points_.push_back(Point{x2, y2});
  points_.push_back(Point{x3, y3});
  return *this;
}

Path& Path::Close() {
  if (!verbs_.empty() && verbs_.back() != V(Verb::kClose)) verbs_.push_back(V(Verb::kClose));
  if (last_move_to_index_ >= 0) last_move_to_index_ = ~last_move_to_index_;
  return *this;
}

Path& Path::AddRect(const Rect& r, Direction dir, uint32_t start) {
  const Point pts[4] = {{r.left, r.top}, {r.right, r.top}, {r.right, r.bottom},
                        {r.left, r.bottom}};
  const uint32_t adv = dir == Direction::kCW ? 1 : 3;
  uint32_t i = start % 4;
  MoveTo(pts[i].x, pts[i].y);
  for (int k = 0; k < 3; k++) {
    i = (i + adv) % 4;
    LineTo(pts[i].x, pts[i].y);
  }
  return Close();
}

Path& Path::AddOval(const Rect& r, Direction dir, uint32_t start) {
  const float cx = r.CenterX(), cy = r.CenterY();
  const Point oval[4] = {{cx, r.top}, {r.right, cy}, {cx, r.bottom}, {r.left, cy}};
  const Point rect[4] = {{r.left, r.top}, {r.right, r.top}, {r.right, r.bottom},
                         {r.left, r.bottom}};
  const uint32_t adv = dir == Direction::kCW ? 1 : 3;
  uint32_t oi = start % 4, ri = (start + (dir == Direction::kCW ? 0 : 1)) % 4;
  MoveTo(oval[oi].x, oval[oi].y);
  for (int k = 0; k < 4; k++) {
    ri = (ri + adv) % 4;
    oi = (oi + adv) % 4;
    ConicTo(rect[ri].x, rect[ri].y, oval[oi].x, oval[oi].y, kRoot2Over2);
  }
  return Close();
}

Path& Path::AddCircle(float x, float y, float radius, Direction dir) {
  if (radius > 0) AddOval(Rect::MakeLTRB(x - radius, y - radius, x + radius, y + radius), dir);
  return *this;
}

// ---- RRect (src/geometry/rrect.cc) -------------------------------------------
namespace {
bool RectFinite(const Rect& r) {
  return std::isfinite(r.left) && std::isfinite(r.top) && std::isfinite(r.right) &&
         std::isfinite(r.bottom);
}
bool RectEmpty(const Rect& r) { return !(r.left < r.right && r.top < r.bottom); }
// ScaleToSides::AdjustRadii (rrect.cc:17-53)
void AdjustRadii(double limit, double scale, float* a, float* b) {
  *a = (float)((double)*a * scale);
  *b = (float)((double)*b * scale);
  if (*a + *b > limit) {
    float* min_r = a;
    float* max_r = b;
    if (*min_r > *max_r) std::swap(min_r, max_r);
    float new_min = *min_r;
    float new_max = (float)(limit - new_min);
    while (new_max + new_min > limit) new_max = std::nextafter(new_max, 0.0f);
    *max_r = new_max;
  }
}
void FlushToZero(float& a, float& b) {
  if (a + b == a) {
    b = 0;
  } else if (a + b == b) {
    a = 0;
  }
}
bool ClampToZero(Point radii[4]) {
  bool all_square = true;
  for (int i = 0; i < 4; i++) {
    if (radii[i].x <= 0 || radii[i].y <= 0) {
      radii[i] = {0, 0};
    } else {
      all_square = false;
    }
  }
  return all_square;
}
double MinScale(double r1, double r2, double limit, double cur) {
  return r1 + r2 > limit ? std::min(cur, limit / (r1 + r2)) : cur;
}
bool RadiusPredicates(float rad, float mn, float mx) {
  return (mn <= mx) && (rad <= mx - mn) && (mn + rad <= mx) && (mx - rad >= mn) && rad >= 0;
}
}  // namespace

bool RRect::InitializeRect(const Rect& rect) {
  if (!RectFinite(rect)) {
    *this = RRect();
    return false;
  }
  // Rect::MakeSorted
  rect_ = Rect{std::min(rect.left, rect.right), std::min(rect.top, rect.bottom),
               std::max(rect.left, rect.right), std::max(rect.top, rect.bottom)};
  if (RectEmpty(rect_)) {
    type_ = kEmpty;
    return false;
  }
  return true;
}

void RRect::SetRect(const Rect& rect) {
  if (!InitializeRect(rect)) return;
  for (auto& r : radii_) r = {0, 0};
  type_ = kRect;
}

void RRect::SetOval(const Rect& oval) {
  if (!InitializeRect(oval)) return;
  // Rect::HalfWidth / HalfHeight
  float x_rad = (rect_.right - rect_.left) * 0.5f, y_rad = (rect_.bottom - rect_.top) * 0.5f;
  if (x_rad == 0.f || y_rad == 0.f) {
    type_ = kRect;
  } else {
    for (auto& r : radii_) r = {x_rad, y_rad};
    type_ = kOval;
  }
}

void RRect::SetRectXY(const Rect& rect, float x_rad, float y_rad) {
  if (!InitializeRect(rect)) return;
  if (!std::isfinite(x_rad) && !std::isfinite(y_rad)) x_rad = y_rad = 0;
  const float w = rect_.right - rect_.left, h = rect_.bottom - rect_.top;
  if (w < x_rad + x_rad || h < y_rad + y_rad) {
    // SkityIEEEFloatDivided: IEEE division, inf for a zero divisor
    float scale = std::min(w / (x_rad + x_rad), h / (y_rad + y_rad));
    x_rad *= scale;
    y_rad *= scale;
  }
  if (x_rad <= 0 || y_rad <= 0) {
    SetRect(rect);
    return;
  }
  for (auto& r : radii_) r = {x_rad, y_rad};
  type_ = kSimple;
  if (x_rad >= 0.5f * w && y_rad >= 0.5f * h) type_ = kOval;
}

bool RRect::ScaleRadii() {
  float scale = 1.0f;
  const double width = (double)rect_.right - (double)rect_.left;
  const double height = (double)rect_.bottom - (double)rect_.top;
  scale = (float)MinScale(radii_[0].x, radii_[1].x, width, scale);
  scale = (float)MinScale(radii_[1].y, radii_[2].y, height, scale);
  scale = (float)MinScale(radii_[2].x, radii_[3].x, width, scale);
  scale = (float)MinScale(radii_[3].y, radii_[0].y, height, scale);
  FlushToZero(radii_[0].x, radii_[1].x);
  FlushToZero(radii_[1].y, radii_[2].y);
  FlushToZero(radii_[2].x, radii_[3].x);
  FlushToZero(radii_[3].y, radii_[0].y);
  if (scale < 1.0f) {
    AdjustRadii(width, scale, &radii_[0].x, &radii_[1].x);
    AdjustRadii(height, scale, &radii_[1].y, &radii_[2].y);
    AdjustRadii(width, scale, &radii_[2].x, &radii_[3].x);
    AdjustRadii(height, scale, &radii_[3].y, &radii_[0].y);
  }
  ClampToZero(radii_);
  ComputeType();
  return scale < 1.0f;
}

// AreRectAndRadiiValid (rrect.cc:309-325): the checks ComputeType's complex
// case falls back on
bool RRect::RadiiValid() const {
  if (!RectFinite(rect_) || !(rect_.left <= rect_.right && rect_.top <= rect_.bottom)) return false;
  for (int i = 0; i < 4; i++)
    if (!RadiusPredicates(radii_[i].x, rect_.left, rect_.right) ||
        !RadiusPredicates(radii_[i].y, rect_.top, rect_.bottom))
      return false;
  return true;
}

void RRect::ComputeType() {
  if (RectEmpty(rect_)) {
    type_ = kEmpty;
    return;
  }
  bool all_equal = true;
  bool all_square = 0 == radii_[0].x || 0 == radii_[0].y;
  for (int i = 1; i < 4; i++) {
    if (0 != radii_[i].x && 0 != radii_[i].y) all_square = false;
    if (radii_[i].x != radii_[i - 1].x || radii_[i].y != radii_[i - 1].y) all_equal = false;
  }
  if (all_square) {
    type_ = kRect;
    return;
  }
  if (all_equal) {
    type_ = radii_[0].x >= 0.5f * (rect_.right - rect_.left) &&
                    radii_[0].y >= 0.5f * (rect_.bottom - rect_.top)
                ? kOval
                : kSimple;
    return;
  }
  type_ = kComplex;
  if (!RadiiValid()) SetRect(rect_);
}

void RRect::SetRectRadii(const Rect& rect, const Point radii[4]) {
  if (!InitializeRect(rect)) return;
  for (int i = 0; i < 4; i++) radii_[i] = radii[i];
  if (ClampToZero(radii_)) {
    SetRect(rect);
    return;
  }
  ScaleRadii();
  if (type_ == kComplex && !RadiiValid()) SetRect(rect);
}

Path& Path::AddRRect(const RRect& rrect, Direction dir) {
  // legacy start indices: 6 (CW) and 7 (CCW)
  return AddRRect(rrect, dir, dir == Direction::kCW ? 6 : 7);
}

Path& Path::AddRRect(const RRect& rrect, Direction dir, uint32_t start) {
  const Rect& b = rrect.GetBounds();
  if (rrect.IsRect() || rrect.IsEmpty()) return AddRect(b, dir, (start + 1) / 2);
  if (rrect.IsOval()) return AddOval(b, dir, start / 2);
  const bool cw = dir == Direction::kCW;
  const bool starts_with_conic = ((start & 1) == (cw ? 1u : 0u));
  const float weight = 0.707106781f;  // FloatRoot2Over2
  // Path_RRectPointIterator / Path_RectPointIterator (path.cc:131-178)
  const float L = b.left, T = b.top, R = b.right, B = b.bottom;
  using C = RRect::Corner;
  const Point rp[8] = {{L + rrect.Radii(C::kUpperLeft).x, T},
                       {R - rrect.Radii(C::kUpperRight).x, T},
                       {R, T + rrect.Radii(C::kUpperRight).y},
                       {R, B - rrect.Radii(C::kLowerRight).y},
                       {R - rrect.Radii(C::kLowerRight).x, B},
                       {L + rrect.Radii(C::kLowerLeft).x, B},
                       {L, B - rrect.Radii(C::kLowerLeft).y},
                       {L, T + rrect.Radii(C::kUpperLeft).y}};
  const Point cp[4] = {{L, T}, {R, T}, {R, B}, {L, B}};
  uint32_t ri = start % 8, ci = (start / 2 + (cw ? 0 : 1)) % 4;
  auto rnext = [&]() -> Point {
    ri = (ri + (cw ? 1 : 7)) % 8;
    return rp[ri];
  };
  auto cnext = [&]() -> Point {
    ci = (ci + (cw ? 1 : 3)) % 4;
    return cp[ci];
  };
  MoveTo(rp[ri].x, rp[ri].y);
  if (starts_with_conic) {
    for (int i = 0; i < 3; i++) {
      Point c = cnext(), e = rnext();
      ConicTo(c.x, c.y, e.x, e.y, weight);
      Point l = rnext();
      LineTo(l.x, l.y);
    }
    Point c = cnext(), e = rnext();
    ConicTo(c.x, c.y, e.x, e.y, weight);
  } else {
    for (int i = 0; i < 4; i++) {
      Point l = rnext();
      LineTo(l.x, l.y);
      Point c = cnext(), e = rnext();
      ConicTo(c.x, c.y, e.x, e.y, weight);
    }
  }
  return Close();
}

Path& Path::AddRoundRect(const Rect& rect, float rx, float ry, Direction dir) {
  if (rx < 0 || ry < 0) return *this;
  return AddRRect(RRect::MakeRectXY(rect, rx, ry), dir);
}

// ---- Shader (src/effect/shader.cc) --------------------------------------------
namespace {
std::shared_ptr<Shader> MakeGradient(int type, const Color4f colors[], const float pos[],
                                     int count, TileMode mode, hipr_gradient* g) {
  if (count < 2 || count > HIPR_MAX_GRADIENT_STOPS) return nullptr;
  g->type = type;
  g->tile_mode = static_cast<int32_t>(mode);
  g->color_count = count;
  g->has_offsets = pos ? 1 : 0;
  for (int i = 0; i < count; i++) {
    g->colors[i][0] = colors[i].r;
    g->colors[i][1] = colors[i].g;
    g->colors[i][2] = colors[i].b;
    g->colors[i][3] = colors[i].a;
    g->offsets[i] = pos ? pos[i] : 0.f;
  }
  const float id[6] = {1, 0, 0, 0, 1, 0};
  std::memcpy(g->local_matrix, id, sizeof(id));
  return nullptr;
}
}  // namespace

std::shared_ptr<Shader> Shader::MakeLinear(const Point pts[2], const Color4f colors[],
                                           const float pos[], int count, TileMode mode) {
  if (count < 2 || count > HIPR_MAX_GRADIENT_STOPS) return nullptr;
  auto s = std::make_shared<Shader>();
  MakeGradient(HIPR_GRADIENT_LINEAR, colors, pos, count, mode, &s->info_);
  s->info_.points[0] = pts[0].x;
  s->info_.points[1] = pts[0].y;
  s->info_.points[2] = pts[1].x;
  s->info_.points[3] = pts[1].y;
  return s;
}

std::shared_ptr<Shader> Shader::MakeRadial(const Point& c, float radius, const Color4f colors[],
                                           const float pos[], int count, TileMode mode) {
  if (count < 2 || count > HIPR_MAX_GRADIENT_STOPS) return nullptr;
  auto s = std::make_shared<Shader>();
  MakeGradient(HIPR_GRADIENT_RADIAL, colors, pos, count, mode, &s->info_);
  s->info_.points[0] = c.x;
  s->info_.points[1] = c.y;
  s->info_.radius[0] = radius;
  return s;
}

std::shared_ptr<Shader> Shader::MakeSweep(float cx, float cy, float start_angle, float end_angle,
                                          const Color4f colors[], const float pos[], int count,
                                          TileMode mode) {
  if (count < 2 || count > HIPR_MAX_GRADIENT_STOPS) return nullptr;
  if (start_angle >= end_angle) return nullptr;  // shader.cc:95-97
  auto s = std::make_shared<Shader>();
  MakeGradient(HIPR_GRADIENT_SWEEP, colors, pos, count, mode, &s->info_);
  s->info_.points[0] = cx;
  s->info_.points[1] = cy;
  float t0 = start_angle / 360, t1 = end_angle / 360;
  s->info_.radius[0] = -t0;             // bias
  s->info_.radius[1] = 1 / (t1 - t0);   // scale
  return s;
}

std::shared_ptr<Image> Image::MakeImage(const uint8_t* rgba, int width, int height,
                                        int alpha_type) {
  if (!rgba || width <= 0 || height <= 0) return nullptr;
  auto im = std::make_shared<Image>();
  im->width_ = width;
  im->height_ = height;
  im->alpha_type_ = alpha_type;
  im->pixels_.assign(rgba, rgba + (size_t)width * height * 4);
  return im;
}

std::shared_ptr<Shader> Shader::MakeShader(std::shared_ptr<Image> image,
                                           const SamplingOptions& sampling, TileMode tile_x,
                                           TileMode tile_y, const Matrix* local) {
  if (!image) return nullptr;
  auto s = std::make_shared<Shader>();
  s->image_ = std::move(image);
  s->image_info_.image = -1;
  s->image_info_.filter = static_cast<int32_t>(sampling.filter);
  s->image_info_.tile_x = static_cast<int32_t>(tile_x);
  s->image_info_.tile_y = static_cast<int32_t>(tile_y);
  const Matrix m = local ? *local : Matrix{};
  const float lm[6] = {m.sx, m.kx, m.tx, m.ky, m.sy, m.ty};
  std::memcpy(s->image_info_.local_matrix, lm, sizeof(lm));
  return s;
}

// Shader::MakeTwoPointConical (src/effect/shader.cc:50-85).
std::shared_ptr<Shader> Shader::MakeTwoPointConical(const Point& start, float start_radius,
                                                    const Point& end, float end_radius,
                                                    const Color4f colors[], const float pos[],
                                                    int count, TileMode mode) {
  if (count < 2 || count > HIPR_MAX_GRADIENT_STOPS) return nullptr;
  const float kNear = 1.f / 4096;
  auto near0 = [&](float v) { return std::fabs(v) <= kNear; };
  if (near0(start.x - end.x) && near0(start.y - end.y)) {
    if (near0(start_radius - end_radius)) return nullptr;
    if (near0(start_radius) || near0(end_radius)) {
      if (near0(end_radius)) {
        // the start circle's radial gradient, stops reversed and mirrored
        Color4f rc[HIPR_MAX_GRADIENT_STOPS];
        float rp[HIPR_MAX_GRADIENT_STOPS];
        for (int i = 0; i < count; i++) {
          rc[i] = colors[count - 1 - i];
          if (pos) rp[i] = 1.f - pos[count - 1 - i];
        }
        return MakeRadial(start, start_radius, rc, pos ? rp : nullptr, count, mode);
      }
      return MakeRadial(start, end_radius, colors, pos, count, mode);
    }
  }
  auto s = std::make_shared<Shader>();
  MakeGradient(HIPR_GRADIENT_CONICAL, colors, pos, count, mode, &s->info_);
  s->info_.points[0] = start.x;
  s->info_.points[1] = start.y;
  s->info_.points[2] = end.x;
  s->info_.points[3] = end.y;
  s->info_.radius[0] = start_radius;
  s->info_.radius[1] = end_radius;
  return s;
}

void Shader::SetLocalMatrix(const float m[6]) { std::memcpy(info_.local_matrix, m, 6 * sizeof(float)); }

std::shared_ptr<ColorFilter> ColorFilters::Blend(Color color, BlendMode mode) {
  const uint32_t a = color >> 24;
  if (mode == BlendMode::kDst || (mode == BlendMode::kDstIn && a == 255) ||
      (mode == BlendMode::kDstOut && a == 0))
    return nullptr;
  auto f = std::make_shared<ColorFilter>();
  f->info_.type = HIPR_COLOR_FILTER_BLEND;
  f->info_.color = color;
  f->info_.blend = static_cast<int32_t>(mode);
  return f;
}

std::shared_ptr<ColorFilter> ColorFilters::Matrix(const float row_major[20]) {
  static const float kIdentity[20] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 0,
                                      0, 0, 1, 0, 0, 0, 0, 0, 1, 0};
  if (!row_major || std::memcmp(row_major, kIdentity, sizeof(kIdentity)) == 0) return nullptr;
  auto f = std::make_shared<ColorFilter>();
  f->info_.type = HIPR_COLOR_FILTER_MATRIX;
  std::memcpy(f->info_.matrix, row_major, sizeof(f->info_.matrix));
  return f;
}

std::shared_ptr<ColorFilter> ColorFilters::LinearToSRGBGamma() {
  auto f = std::make_shared<ColorFilter>();
  f->info_.type = HIPR_COLOR_FILTER_LINEAR_TO_SRGB;
  return f;
}

std::shared_ptr<ColorFilter> ColorFilters::SRGBToLinearGamma() {
  auto f = std::make_shared<ColorFilter>();
  f->info_.type = HIPR_COLOR_FILTER_SRGB_TO_LINEAR;
  return f;
}

std::shared_ptr<ColorFilter> ColorFilters::Compose(std::shared_ptr<ColorFilter> outer,
                                                   std::shared_ptr<ColorFilter> inner) {
  if (!inner) return outer;
  if (!outer) return inner;
  auto f = std::make_shared<ColorFilter>();
  f->info_.type = HIPR_COLOR_FILTER_COMPOSE;
  return f;
}

std::shared_ptr<MaskFilter> MaskFilter::MakeBlur(BlurStyle style, float radius) {
  auto m = std::make_shared<MaskFilter>();
  m->style_ = style;
  m->radius_ = radius;
  return m;
}

// ---- Paint ------------------------------------------------------------------
void Paint::SetColor(Color c) { fill_ = stroke_ = Color4fFromColor(c); }

void Paint::SetAlphaF(float a) {
  a = std::min(std::max(a, 0.f), 1.f);
  fill_.a = a;
  stroke_.a = a;
}

hipr_paint Paint::Pack(int32_t gradient_index) const {
  hipr_paint p{};
  p.color[0] = fill_.r;
  p.color[1] = fill_.g;
  p.color[2] = fill_.b;
  p.color[3] = fill_.a;
  p.stroke_color[0] = stroke_.r;
  p.stroke_color[1] = stroke_.g;
  p.stroke_color[2] = stroke_.b;
  p.stroke_color[3] = stroke_.a;
  p.style = style_;
  p.stroke_width = width_;
  p.miter_limit = miter_;
  p.cap = cap_;
  p.join = join_;
  p.gradient = gradient_index;
  p.blur_radius = mask_ ? mask_->radius() : 0.f;
  p.blur_style = mask_ ? static_cast<int32_t>(mask_->style()) : 0;
  p.blend = static_cast<int32_t>(blend_);
  return p;
}

// ---- Matrix -------------------------------------------------------------------
Matrix Matrix::Translate(float dx, float dy) {
  Matrix m;
  m.tx = dx;
  m.ty = dy;
  return m;
}
Matrix Matrix::Scale(float x, float y) {
  Matrix m;
  m.sx = x;
  m.sy = y;
  return m;
}
Matrix Matrix::Skew(float kx, float ky) {
  Matrix m;
  m.kx = kx;
  m.ky = ky;
  return m;
}
Matrix Matrix::RotateDeg(float degrees, float px, float py) {
  const float r = degrees * 0.01745329251994329577f;
  const float c = std::cos(r), s = std::sin(r);
  Matrix rot;
  rot.sx = c;
  rot.kx = -s;
  rot.ky = s;
  rot.sy = c;
  if (px == 0 && py == 0) return rot;
  return Translate(px, py) * rot * Translate(-px, -py);
}
Matrix Matrix::operator*(const Matrix& b) const {
  const Matrix& a = *this;
  if (a.IsIdentity()) return b;
  if (b.IsIdentity()) return a;
  Matrix r;
  r.sx = a.sx * b.sx + a.kx * b.ky;
  r.kx = a.sx * b.kx + a.kx * b.sy;
  r.tx = (a.sx * b.tx + a.kx * b.ty) + a.tx;
  r.ky = a.ky * b.sx + a.sy * b.ky;
  r.sy = a.ky * b.kx + a.sy * b.sy;
  r.ty = (a.ky * b.tx + a.sy * b.ty) + a.ty;
  return r;
}

// ---- Canvas -----------------------------------------------------------------
int HIPCanvas::Save() {
  stack_.push_back(stack_.back());
  clips_.push_back(clips_.back());
  if (recording_) recording_->ops_.push_back({DisplayList::OpKind::kSave, Matrix(), 0});
  return GetSaveCount() - 1;
}

void HIPCanvas::Restore() {
  if (stack_.size() > 1) {
    stack_.pop_back();
    clips_.pop_back();
    if (recording_) recording_->ops_.push_back({DisplayList::OpKind::kRestore, Matrix(), 0});
  }
}

void HIPCanvas::RestoreToCount(int count) {
  if (count < 1) count = 1;
  while ((int)stack_.size() > count) Restore();
}

void HIPCanvas::Concat(const Matrix& m) {
  stack_.back() = stack_.back() * m;
  if (recording_) recording_->ops_.push_back({DisplayList::OpKind::kConcat, m, 0});
}

void HIPCanvas::SetMatrix(const Matrix& m) {
  stack_.back() = m;
  if (recording_) recording_->ops_.push_back({DisplayList::OpKind::kSetMatrix, m, 0});
}

void HIPCanvas::ClipRect(const Rect& rect, ClipOp op) {
  if (recording_) {
    DisplayList::Op o{DisplayList::OpKind::kClipRect, Matrix(),
                      op == ClipOp::kDifference ? 1u : 0u, rect};
    recording_->ops_.push_back(o);
  }
  const Matrix& m = stack_.back();
  if (op == ClipOp::kIntersect) {
    // CalculateGlobalClipBounds: Matrix::MapRect of the rectangle (corners
    // as glm's mat4 * vec4 sums them, then their min / max), intersected
    // with the bounds (Rect::Intersect; SetEmpty when they do not overlap)
    const float xs[4] = {rect.left, rect.right, rect.right, rect.left};
    const float ys[4] = {rect.top, rect.top, rect.bottom, rect.bottom};
    float l = 0, t = 0, r = 0, b = 0;
    for (int i = 0; i < 4; i++) {
      float x = (m.sx * xs[i] + m.kx * ys[i]) + (0.f + m.tx);
      float y = (m.ky * xs[i] + m.sy * ys[i]) + (0.f + m.ty);
      if (i == 0) {
        l = r = x;
        t = b = y;
      } else {
        l = std::min(x, l);
        r = std::max(x, r);
        t = std::min(y, t);
        b = std::max(y, b);
      }
    }
    Rect& c = clips_.back();
    float il = std::max(c.left, l), ir = std::min(c.right, r);
    float it = std::max(c.top, t), ib = std::min(c.bottom, b);
    c = il < ir && it < ib ? Rect{il, it, ir, ib} : Rect{0, 0, 0, 0};
  }
  // a difference clip, or one under skew / rotation, also clips span lists
  // (sw_canvas.cc:297-313, OnClipPath)
  if (!recording_ && (op == ClipOp::kDifference || m.kx != 0.f || m.ky != 0.f))
    unsupported_ = true;
}

void HIPCanvas::MarkUnsupported() {
  if (recording_) {
    recording_->ops_.push_back({DisplayList::OpKind::kUnsupported, Matrix(), 0});
  } else {
    unsupported_ = true;
  }
}

void HIPCanvas::AddDraw(Draw d, const Matrix& local) {
  if (recording_) {
    recording_->ops_.push_back(
        {DisplayList::OpKind::kDraw, local, (uint32_t)recording_->draws_.size()});
    recording_->draws_.push_back(std::move(d));
    return;
  }
  draws_.push_back(std::move(d));
  if (surface_) surface_->readback_valid_ = false;
}

void HIPCanvas::DrawPath(const Path& path, const Paint& paint) {
  AddDraw(Draw{path, paint, stack_.back(), HIPR_DRAW_PATH, clips_.back()});
}

void HIPCanvas::DrawPaint(const Paint& paint) {
  AddDraw(Draw{Path(), paint, stack_.back(), HIPR_DRAW_PAINT});
}

void HIPCanvas::DrawColor(Color color, BlendMode mode) {
  Paint paint;
  paint.SetFillColor(color);
  paint.SetBlendMode(mode);
  DrawPaint(paint);
}

void HIPCanvas::DrawRect(const Rect& rect, const Paint& paint) {
  Path p;
  p.AddRect(rect);
  DrawPath(p, paint);
}

void HIPCanvas::DrawOval(const Rect& oval, const Paint& paint) {
  // OnDrawOval -> RRect::SetOval -> Path::AddRRect(kCW, start 6) -> AddOval
  // with start 6 / 2 (path.cc:881-896)
  Path p;
  p.AddOval(oval, Path::Direction::kCW, 3);
  DrawPath(p, paint);
}

void HIPCanvas::DrawCircle(float cx, float cy, float radius, const Paint& paint) {
  if (radius < 0) radius = 0;
  DrawOval(Rect::MakeLTRB(cx - radius, cy - radius, cx + radius, cy + radius), paint);
}

void HIPCanvas::DrawRRect(const RRect& rrect, const Paint& paint) {
  Path p;
  p.AddRRect(rrect);
  DrawPath(p, paint);
}

void HIPCanvas::DrawRoundRect(const Rect& rect, float rx, float ry, const Paint& paint) {
  if (rx > 0 && ry > 0) {
    DrawRRect(RRect::MakeRectXY(rect, rx, ry), paint);
  } else {
    DrawRect(rect, paint);
  }
}

void HIPCanvas::DrawDRRect(const RRect& outer, const RRect& inner, const Paint& paint) {
  const Rect& o = outer.GetBounds();
  const Rect& i = inner.GetBounds();
  if (!(o.left < o.right && o.top < o.bottom)) return;
  if (!(i.left < i.right && i.top < i.bottom)) {
    DrawRRect(outer, paint);
    return;
  }
  // Rect::Contains: the inner bounds inside the outer ones
  if (!(o.left <= i.left && o.top <= i.top && o.right >= i.right && o.bottom >= i.bottom)) return;
  Path p;
  p.AddRRect(outer);
  p.AddRRect(inner);
  p.SetFillType(Path::PathFillType::kEvenOdd);
  DrawPath(p, paint);
}

void HIPCanvas::DrawLine(float x0, float y0, float x1, float y1, const Paint& paint) {
  Path p;
  p.MoveTo(x0, y0);
  p.LineTo(x1, y1);
  DrawPath(p, paint);
}

void HIPCanvas::DrawGlyphs(uint32_t count, const Path* const glyph_paths[],
                           const float position_x[], const float position_y[],
                           const Paint& paint) {
  if (count == 0) return;
  if (paint.GetMaskFilter() || paint.shader()) {
    MarkUnsupported();
    return;
  }
  const bool need_fill = paint.GetStyle() != Paint::kStroke_Style;
  const bool need_stroke = paint.GetStyle() != Paint::kFill_Style;
  auto run = [&](Paint::Style style) {
    Paint p = paint;
    p.SetStyle(style);
    for (uint32_t k = 0; k < count; k++) {
      const Matrix local = Matrix::Translate(position_x[k], position_y[k]);
      // FillGlyphs / StrokeGlyphs raster without the clip bounds
      Draw d{*glyph_paths[k], p, stack_.back() * local, HIPR_DRAW_PATH};
      d.unclipped = true;
      AddDraw(std::move(d), local);
    }
  };
  // DrawFillStrokeInPaintOrder (src/render/paint_order.hpp:12-31)
  if (paint.GetStyle() == Paint::kStrokeThenFill_Style) {
    if (need_stroke) run(Paint::kStroke_Style);
    if (need_fill) run(Paint::kFill_Style);
  } else {
    if (need_fill) run(Paint::kFill_Style);
    if (need_stroke) run(Paint::kStroke_Style);
  }
}

PackedScene::PackedScene(const std::vector<HIPCanvas::Draw>& draws) {
  const int32_t n = (int32_t)draws.size();
  vo_.assign(n + 1, 0);
  po_.assign(n + 1, 0);
  wo_.assign(n + 1, 0);
  fills_.resize(n);
  paints_.resize(n);
  matrices_.resize(6 * (size_t)n);
  kinds_.resize(n);
  filters_.resize(n);
  shaders_.resize(n);
  clips_.resize(4 * (size_t)n);
  std::vector<const Image*> image_of;
  bool any_matrix = false, any_paint = false, any_filter = false, any_clip = false;
  for (int32_t i = 0; i < n; i++) {
    const Matrix& m = draws[i].matrix;
    const float mf[6] = {m.sx, m.kx, m.tx, m.ky, m.sy, m.ty};
    std::copy(mf, mf + 6, matrices_.begin() + 6 * (size_t)i);
    any_matrix |= !m.IsIdentity();
    kinds_[i] = draws[i].kind;
    any_paint |= kinds_[i] != HIPR_DRAW_PATH;
    const Rect& c = draws[i].clip;
    const float cf[4] = {c.left, c.top, c.right, c.bottom};
    std::copy(cf, cf + 4, clips_.begin() + 4 * (size_t)i);
    any_clip |= !(c.left <= -1e9f && c.top <= -1e9f && c.right >= 1e9f && c.bottom >= 1e9f);
    const Path& p = draws[i].path;
    verbs_.insert(verbs_.end(), p.verbs().begin(), p.verbs().end());
    for (const Point& q : p.points()) {
      pts_.push_back(q.x);
      pts_.push_back(q.y);
    }
    ws_.insert(ws_.end(), p.weights().begin(), p.weights().end());
    vo_[i + 1] = (int32_t)verbs_.size();
    po_[i + 1] = (int32_t)(pts_.size() / 2);
    wo_[i + 1] = (int32_t)ws_.size();
    fills_[i] = p.GetFillType() == Path::PathFillType::kEvenOdd ? HIPR_FILL_EVEN_ODD
                                                                : HIPR_FILL_WINDING;
    int32_t gi = -1;
    shaders_[i].image = -1;
    if (const auto& sh = draws[i].paint.shader()) {
      if (sh->image()) {
        // a pixmap shader: one hipr_image per distinct Image
        const Image* im = sh->image().get();
        int32_t k = 0;
        while (k < (int32_t)image_of.size() && image_of[k] != im) k++;
        if (k == (int32_t)image_of.size()) {
          image_of.push_back(im);
          images_.push_back(hipr_image{im->Width(), im->Height(), im->AlphaType(), im->Pixels()});
        }
        shaders_[i] = sh->image_info();
        shaders_[i].image = k;
      } else {
        gi = (int32_t)grads_.size();
        grads_.push_back(sh->info());
      }
    }
    paints_[i] = draws[i].paint.Pack(gi);
    if (const auto& cf = draws[i].paint.GetColorFilter()) {
      filters_[i] = cf->info();
      any_filter = true;
    }
  }
  scene_.n_paths = n;
  scene_.verbs = verbs_.data();
  scene_.verb_offsets = vo_.data();
  scene_.points = pts_.data();
  scene_.point_offsets = po_.data();
  scene_.weights = ws_.data();
  scene_.weight_offsets = wo_.data();
  scene_.fill_types = fills_.data();
  scene_.paints = paints_.data();
  scene_.gradients = grads_.empty() ? nullptr : grads_.data();
  scene_.n_gradients = (int32_t)grads_.size();
  scene_.matrices = any_matrix ? matrices_.data() : nullptr;
  scene_.draw_kinds = any_paint ? kinds_.data() : nullptr;
  scene_.color_filters = any_filter ? filters_.data() : nullptr;
  scene_.clip_rects = any_clip ? clips_.data() : nullptr;
  if (!images_.empty()) {
    scene_.images = images_.data();
    scene_.n_images = (int32_t)images_.size();
    scene_.image_shaders = shaders_.data();
  }
}

int HIPCanvas::Flush() {
  if (recording_) return HIPR_SUCCESS;  // RecordingCanvas::OnFlush records nothing
  int r = HIPR_SUCCESS;
  if (unsupported_) {
    r = HIPR_ERROR_NOT_SUPPORTED;
  } else if (!draws_.empty()) {
    PackedScene packed(draws_);
    hipr_options opt{};
    opt.mode = HIPR_MODE_EXACT;
    r = hipr_surface_draw_scene(surface_->surface_, packed.scene(), &opt);
  } else {
    return HIPR_SUCCESS;
  }
  draws_.clear();
  unsupported_ = false;
  surface_->last_result_ = r;
  return r;
}

// ---- DisplayList / PictureRecorder --------------------------------------------
void DisplayList::Draw(HIPCanvas* canvas) const {
  if (!canvas) return;
  for (const Op& op : ops_) {
    switch (op.kind) {
      case OpKind::kSave:
        canvas->Save();
        break;
      case OpKind::kRestore:
        canvas->Restore();
        break;
      case OpKind::kConcat:
        canvas->Concat(op.matrix);
        break;
      case OpKind::kSetMatrix:
        canvas->SetMatrix(op.matrix);
        break;
      case OpKind::kDraw: {
        HIPCanvas::Draw d = draws_[op.draw];
        d.matrix = canvas->GetTotalMatrix() * op.matrix;
        if (d.kind == HIPR_DRAW_PATH && !d.unclipped) d.clip = canvas->GetGlobalClipBounds();
        canvas->AddDraw(std::move(d), op.matrix);
        break;
      }
      case OpKind::kClipRect:
        canvas->ClipRect(op.rect, op.draw ? HIPCanvas::ClipOp::kDifference
                                          : HIPCanvas::ClipOp::kIntersect);
        break;
      case OpKind::kUnsupported:
        canvas->MarkUnsupported();
        break;
    }
  }
}

bool ReplayDraws(const DisplayList& list, std::vector<HIPCanvas::Draw>* draws) {
  HIPCanvas canvas(nullptr);
  list.Draw(&canvas);
  *draws = std::move(canvas.draws_);
  return !canvas.unsupported_;
}

PictureRecorder::PictureRecorder() = default;
PictureRecorder::~PictureRecorder() = default;

void PictureRecorder::BeginRecording(const Rect& bounds) {
  list_.reset(new DisplayList());
  list_->bounds_ = bounds;
  canvas_.reset(new HIPCanvas(nullptr));
  canvas_->recording_ = list_.get();
}

std::unique_ptr<DisplayList> PictureRecorder::FinishRecording() {
  canvas_.reset();
  return std::move(list_);
}

// ---- Surface ----------------------------------------------------------------
std::unique_ptr<HIPSurface> HIPSurface::Make(hipr_context ctx, int width, int height,
                                             int alpha_type, int color_type) {
  std::unique_ptr<HIPSurface> s(new HIPSurface());
  if (hipr_surface_create_ex(ctx, width, height, alpha_type, color_type, &s->surface_) !=
      HIPR_SUCCESS)
    return nullptr;
  s->color_type_ = color_type;
  s->width_ = width;
  s->height_ = height;
  s->canvas_.reset(new HIPCanvas(s.get()));
  return s;
}

HIPSurface::~HIPSurface() { hipr_surface_destroy(surface_); }

HIPCanvas* HIPSurface::LockCanvas(bool clear) {
  if (clear) {
    canvas_->draws_.clear();
    last_result_ = hipr_surface_clear(surface_, 0);
    readback_valid_ = false;
  }
  return canvas_.get();
}

int HIPSurface::Flush() {
  int r = canvas_->Flush();
  readback_valid_ = false;
  return r;
}

int HIPSurface::ReadPixels(void* dst, size_t bytes) {
  return hipr_surface_read_pixels(surface_, static_cast<uint8_t*>(dst), bytes);
}

Color HIPSurface::GetPixel(int x, int y) {
  if (x < 0 || y < 0 || x >= width_ || y >= height_) return Color_TRANSPARENT;
  if (!readback_valid_) {
    readback_.resize((size_t)width_ * height_ * 4);
    if (ReadPixels(readback_.data(), readback_.size()) != HIPR_SUCCESS) return Color_TRANSPARENT;
    readback_valid_ = true;
  }
  const uint8_t* p = &readback_[((size_t)y * width_ + x) * 4];
  // Bitmap::GetPixel (bitmap.cc:25-35)
  return color_type_ == HIPR_COLOR_BGRA ? ColorSetARGB(p[3], p[2], p[1], p[0])
                                        : ColorSetARGB(p[3], p[0], p[1], p[2]);
}

}  // namespace skity_hip
