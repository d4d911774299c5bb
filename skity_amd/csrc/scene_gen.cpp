// scene_gen.cpp -- seeded synthetic scenes for tests and the benchmark.
#include "scene_gen.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Pcg32 {
  uint64_t state = 0, inc = 0;
  static uint64_t SplitMix(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Pcg32(uint64_t seed) {
    uint64_t s = seed;
    uint64_t init_state = SplitMix(&s);
    inc = (SplitMix(&s) << 1) | 1u;
    state = 0;
    Next();
    state += init_state;
    Next();
  }
  uint32_t Next() {
    uint64_t old = state;
    state = old * 6364136223846793005ull + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
  }
  // uniform float in [0, 1) with 24 random bits
  float Unit() { return (Next() >> 8) * (1.0f / 16777216.0f); }
  float Uniform(float lo, float hi) { return lo + (hi - lo) * Unit(); }
  // uniform integer in [lo, hi]
  int Int(int lo, int hi) {
    uint32_t span = (uint32_t)(hi - lo + 1);
    return lo + (int)(((uint64_t)Next() * span) >> 32);
  }
};

// Path builder with skity's Path semantics for the verbs the generators use
// (src/graphic/path.cc:494-570,837-859): ConicTo with w == 1 becomes a quad.
struct Builder {
  std::vector<uint8_t> verbs;
  std::vector<float> pts;
  std::vector<float> weights;
  void Move(float x, float y) {
    verbs.push_back(HIPR_VERB_MOVE);
    pts.push_back(x);
    pts.push_back(y);
  }
  void Line(float x, float y) {
    verbs.push_back(HIPR_VERB_LINE);
    pts.push_back(x);
    pts.push_back(y);
  }
  void Quad(float x1, float y1, float x2, float y2) {
    verbs.push_back(HIPR_VERB_QUAD);
    pts.insert(pts.end(), {x1, y1, x2, y2});
  }
  void Conic(float x1, float y1, float x2, float y2, float w) {
    if (!(w > 0)) {
      Line(x2, y2);
    } else if (w == 1.f) {
      Quad(x1, y1, x2, y2);
    } else {
      verbs.push_back(HIPR_VERB_CONIC);
      pts.insert(pts.end(), {x1, y1, x2, y2});
      weights.push_back(w);
    }
  }
  void Cubic(float x1, float y1, float x2, float y2, float x3, float y3) {
    verbs.push_back(HIPR_VERB_CUBIC);
    pts.insert(pts.end(), {x1, y1, x2, y2, x3, y3});
  }
  void Close() { verbs.push_back(HIPR_VERB_CLOSE); }
};

}  // namespace

struct sg_scene_s {
  int32_t width = 0, height = 0;
  std::vector<uint8_t> verbs;
  std::vector<int32_t> verb_off{0};
  std::vector<float> points;
  std::vector<int32_t> point_off{0};
  std::vector<float> weights;
  std::vector<int32_t> weight_off{0};
  std::vector<uint8_t> fills;
  std::vector<hipr_paint> paints;
  std::vector<hipr_gradient> gradients;
  hipr_scene view{};

  void Add(const Builder& b, int fill, const hipr_paint& paint) {
    verbs.insert(verbs.end(), b.verbs.begin(), b.verbs.end());
    verb_off.push_back((int32_t)verbs.size());
    points.insert(points.end(), b.pts.begin(), b.pts.end());
    point_off.push_back((int32_t)(points.size() / 2));
    weights.insert(weights.end(), b.weights.begin(), b.weights.end());
    weight_off.push_back((int32_t)weights.size());
    fills.push_back((uint8_t)fill);
    paints.push_back(paint);
  }
  void Finish() {
    view.n_paths = (int32_t)paints.size();
    view.verbs = verbs.data();
    view.verb_offsets = verb_off.data();
    view.points = points.data();
    view.point_offsets = point_off.data();
    view.weights = weights.data();
    view.weight_offsets = weight_off.data();
    view.fill_types = fills.data();
    view.paints = paints.data();
    view.gradients = gradients.data();
    view.n_gradients = (int32_t)gradients.size();
  }
};

namespace {

hipr_paint SolidPaint(Pcg32& rng, int min_alpha) {
  hipr_paint p{};
  // Paint::SetColor(ColorSetARGB(a, r, g, b)) keeps Color4fFromColor: c/255.f
  int r = rng.Int(0, 255), g = rng.Int(0, 255), b = rng.Int(0, 255);
  int a = rng.Int(min_alpha, 255);
  p.color[0] = r / 255.f;
  p.color[1] = g / 255.f;
  p.color[2] = b / 255.f;
  p.color[3] = a / 255.f;
  for (int c = 0; c < 4; c++) p.stroke_color[c] = p.color[c];
  p.style = HIPR_STYLE_FILL;
  p.stroke_width = 0.f;
  p.miter_limit = 4.f;
  p.cap = HIPR_CAP_BUTT;
  p.join = HIPR_JOIN_MITER;
  p.gradient = -1;
  p.blur_radius = 0.f;
  p.blur_style = HIPR_BLUR_NORMAL;
  p.blend = HIPR_BLEND_SRC_OVER;
  return p;
}

void GenPolygons(sg_scene_s* s, Pcg32& rng, int n) {
  s->width = 800;
  s->height = 600;
  for (int i = 0; i < n; i++) {
    Builder b;
    int k = rng.Int(5, 32);
    for (int v = 0; v < k; v++) {
      float x = rng.Uniform(0.f, 800.f), y = rng.Uniform(0.f, 600.f);
      if (v == 0) {
        b.Move(x, y);
      } else {
        b.Line(x, y);
      }
    }
    b.Close();
    s->Add(b, HIPR_FILL_WINDING, SolidPaint(rng, 77));
  }
}

void GenCurves(sg_scene_s* s, Pcg32& rng, int n, float side) {
  s->width = (int32_t)side;
  s->height = (int32_t)side;
  auto P = [&]() { return rng.Uniform(0.f, side); };
  for (int i = 0; i < n; i++) {
    Builder b;
    int segs = rng.Int(8, 200);
    b.Move(P(), P());
    for (int k = 0; k < segs; k++) {
      int verb = rng.Int(0, 2);
      if (verb == 0) {
        float x1 = P(), y1 = P(), x2 = P(), y2 = P();
        b.Quad(x1, y1, x2, y2);
      } else if (verb == 1) {
        float x1 = P(), y1 = P(), x2 = P(), y2 = P(), x3 = P(), y3 = P();
        b.Cubic(x1, y1, x2, y2, x3, y3);
      } else {
        float x1 = P(), y1 = P(), x2 = P(), y2 = P();
        float w = rng.Uniform(0.2f, 4.f);
        b.Conic(x1, y1, x2, y2, w);
      }
    }
    b.Close();
    int fill = rng.Int(0, 1) ? HIPR_FILL_EVEN_ODD : HIPR_FILL_WINDING;
    hipr_paint paint = SolidPaint(rng, 77);
    if (rng.Unit() < 0.2f) {
      paint.style = HIPR_STYLE_STROKE;
      paint.stroke_width = rng.Uniform(1.f, 16.f);
      paint.join = rng.Int(0, 2);
      paint.cap = rng.Int(0, 1);
    }
    s->Add(b, fill, paint);
  }
}

void GenGradientBlur(sg_scene_s* s, Pcg32& rng, int n) {
  s->width = 4096;
  s->height = 4096;
  for (int i = 0; i < n; i++) {
    float side = std::exp(rng.Uniform(std::log(32.f), std::log(1024.f)));
    float bx = rng.Uniform(0.f, 4096.f - side), by = rng.Uniform(0.f, 4096.f - side);
    auto X = [&]() { return bx + rng.Uniform(0.f, side); };
    auto Y = [&]() { return by + rng.Uniform(0.f, side); };
    Builder b;
    b.Move(X(), Y());
    for (int k = 0; k < 4; k++) {
      float x1 = X(), y1 = Y(), x2 = X(), y2 = Y(), x3 = X(), y3 = Y();
      b.Cubic(x1, y1, x2, y2, x3, y3);
    }
    b.Close();
    hipr_gradient g{};
    g.type = HIPR_GRADIENT_LINEAR + rng.Int(0, 2);
    g.tile_mode = rng.Int(0, 2);
    g.color_count = rng.Int(2, 16);
    g.has_offsets = 1;
    std::vector<float> offs(g.color_count);
    for (auto& o : offs) o = rng.Unit();
    std::sort(offs.begin(), offs.end());
    for (int c = 0; c < g.color_count; c++) {
      g.offsets[c] = offs[c];
      for (int ch = 0; ch < 4; ch++) g.colors[c][ch] = rng.Unit();
    }
    float cx = X(), cy = Y();
    if (g.type == HIPR_GRADIENT_LINEAR) {
      g.points[0] = cx;
      g.points[1] = cy;
      g.points[2] = X();
      g.points[3] = Y();
    } else if (g.type == HIPR_GRADIENT_RADIAL) {
      g.points[0] = cx;
      g.points[1] = cy;
      g.radius[0] = rng.Uniform(0.1f, 0.6f) * side;
    } else {
      // Shader::MakeSweep(cx, cy, start, end): bias = -start/360,
      // scale = 1 / (end/360 - start/360) (src/effect/shader.cc:88-105).
      float start = rng.Uniform(0.f, 180.f);
      float end = start + rng.Uniform(90.f, 360.f);
      float t0 = start / 360, t1 = end / 360;
      g.points[0] = cx;
      g.points[1] = cy;
      g.radius[0] = -t0;
      g.radius[1] = 1 / (t1 - t0);
    }
    // local matrix = T(c) * R * Shear * S * T(-c), centred in the box
    float ang = rng.Uniform(0.f, 6.2831853f);
    float sx = rng.Uniform(0.5f, 2.f), sy = rng.Uniform(0.5f, 2.f);
    float sh = rng.Uniform(-0.3f, 0.3f);
    float c = std::cos(ang), sn = std::sin(ang);
    // M = R * [[sx, sh*sy],[0, sy]]
    float a = c * sx, bb = c * sh * sy - sn * sy;
    float cc = sn * sx, d = sn * sh * sy + c * sy;
    float ox = bx + side * 0.5f, oy = by + side * 0.5f;
    g.local_matrix[0] = a;
    g.local_matrix[1] = bb;
    g.local_matrix[2] = ox - (a * ox + bb * oy);
    g.local_matrix[3] = cc;
    g.local_matrix[4] = d;
    g.local_matrix[5] = oy - (cc * ox + d * oy);
    hipr_paint paint = SolidPaint(rng, 77);
    paint.gradient = (int32_t)s->gradients.size();
    float sigma = rng.Uniform(1.f, 32.f);
    paint.blur_radius = (sigma - 0.5f) / 0.57735f;
    paint.blur_style = HIPR_BLUR_NORMAL;
    s->gradients.push_back(g);
    s->Add(b, rng.Int(0, 1) ? HIPR_FILL_EVEN_ODD : HIPR_FILL_WINDING, paint);
  }
}

// A canvas of uneven scribbles: side 2^U[8,12], paths round(10^U[lo_p,
// hi_p]), segments per path round(10^U[0.602, hi_s]) of lines / quads /
// cubics at 40 / 30 / 30 %, control points uniform in the canvas, 50%
// even-odd; paints 70% solid, 30% linear or radial gradient (2-16 stops,
// clamp / repeat / mirror).  Used scaled (10-1,000 paths of 4-316 segments):
// every path spans the canvas, a stress case for wide active lists.
void GenUneven(sg_scene_s* s, Pcg32& rng, int n, float lo_p, float hi_p, float hi_s) {
  int side = 1 << rng.Int(8, 12);
  s->width = side;
  s->height = side;
  int paths = n > 0 ? n : (int)std::lround(std::pow(10.f, rng.Uniform(lo_p, hi_p)));
  auto P = [&]() { return rng.Uniform(0.f, (float)side); };
  for (int i = 0; i < paths; i++) {
    Builder b;
    int segs = (int)std::lround(std::pow(10.f, rng.Uniform(0.602f, hi_s)));
    b.Move(P(), P());
    for (int k = 0; k < segs; k++) {
      float u = rng.Unit();
      if (u < 0.4f) {
        float x = P(), y = P();
        b.Line(x, y);
      } else if (u < 0.7f) {
        float x1 = P(), y1 = P(), x2 = P(), y2 = P();
        b.Quad(x1, y1, x2, y2);
      } else {
        float x1 = P(), y1 = P(), x2 = P(), y2 = P(), x3 = P(), y3 = P();
        b.Cubic(x1, y1, x2, y2, x3, y3);
      }
    }
    b.Close();
    hipr_paint paint = SolidPaint(rng, 77);
    if (rng.Unit() < 0.3f) {
      hipr_gradient g{};
      g.type = rng.Int(0, 1) ? HIPR_GRADIENT_RADIAL : HIPR_GRADIENT_LINEAR;
      g.tile_mode = rng.Int(0, 2);
      g.color_count = rng.Int(2, 16);
      g.has_offsets = 1;
      std::vector<float> offs(g.color_count);
      for (auto& o : offs) o = rng.Unit();
      std::sort(offs.begin(), offs.end());
      for (int c = 0; c < g.color_count; c++) {
        g.offsets[c] = offs[c];
        for (int ch = 0; ch < 4; ch++) g.colors[c][ch] = rng.Unit();
      }
      g.points[0] = P();
      g.points[1] = P();
      if (g.type == HIPR_GRADIENT_LINEAR) {
        g.points[2] = P();
        g.points[3] = P();
      } else {
        g.radius[0] = rng.Uniform(0.05f, 0.5f) * side;
      }
      g.local_matrix[0] = 1.f;
      g.local_matrix[4] = 1.f;
      paint.gradient = (int32_t)s->gradients.size();
      s->gradients.push_back(g);
    }
    s->Add(b, rng.Int(0, 1) ? HIPR_FILL_EVEN_ODD : HIPR_FILL_WINDING, paint);
  }
}

// A gradient paint for a path in box (cx, cy, b): linear or radial, 2-16
// stops, clamp / repeat / mirror.
void AddGradient(sg_scene_s* s, Pcg32& rng, float cx, float cy, float b, hipr_paint* paint) {
  hipr_gradient g{};
  g.type = rng.Int(0, 1) ? HIPR_GRADIENT_RADIAL : HIPR_GRADIENT_LINEAR;
  g.tile_mode = rng.Int(0, 2);
  g.color_count = rng.Int(2, 16);
  g.has_offsets = 1;
  std::vector<float> offs(g.color_count);
  for (auto& o : offs) o = rng.Unit();
  std::sort(offs.begin(), offs.end());
  for (int c = 0; c < g.color_count; c++) {
    g.offsets[c] = offs[c];
    for (int ch = 0; ch < 4; ch++) g.colors[c][ch] = rng.Unit();
  }
  g.points[0] = cx + rng.Uniform(-0.5f, 0.5f) * b;
  g.points[1] = cy + rng.Uniform(-0.5f, 0.5f) * b;
  if (g.type == HIPR_GRADIENT_LINEAR) {
    g.points[2] = cx + rng.Uniform(-0.5f, 0.5f) * b;
    g.points[3] = cy + rng.Uniform(-0.5f, 0.5f) * b;
  } else {
    g.radius[0] = rng.Uniform(0.1f, 0.7f) * b;
  }
  g.local_matrix[0] = 1.f;
  g.local_matrix[4] = 1.f;
  paint->gradient = (int32_t)s->gradients.size();
  s->gradients.push_back(g);
}

// One canvas of the batched uneven family at its stated scale: side
// 2^U{8..12}, round(10^U[2, 4.699]) paths (100-50,000), each of
// round(10^U[0.602, 4]) segments (4-10,000) of lines / quads / cubics at
// 40 / 30 / 30 %; paints 70% solid, 30% linear or radial gradient; half
// even-odd.  Geometry (a choice the family leaves open): each path is a
// closed outline around a centre uniform in the canvas, inside a box of
// side (canvas side) * 2^U[-5, -1], through s anchors at equal angles whose
// radius follows two random harmonics; quads and cubics bend their chord by
// up to 30% of its length.  Such outlines keep a handful of edges per row
// however many segments they have, as drawn shapes do.  Every path has its
// own random stream (seed, path index), so paths are generated in parallel.
struct OutlinePath {
  Builder b;
  int fill;
  hipr_paint paint;
  bool gradient;
  hipr_gradient g;
};

void GenOutlinePath(uint64_t seed, int index, int side, float box_lo, float box_hi,
                    OutlinePath* out) {
  Pcg32 rng(seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(index + 1)));
  const int segs = (int)std::lround(std::pow(10.f, rng.Uniform(0.602f, 4.f)));
  const float b = (float)side * std::exp2(rng.Uniform(box_lo, box_hi));
  const float cx = rng.Uniform(0.f, (float)side), cy = rng.Uniform(0.f, (float)side);
  const int m1 = rng.Int(2, 7), m2 = rng.Int(5, 13);
  const float ph1 = rng.Uniform(0.f, 6.2831853f), ph2 = rng.Uniform(0.f, 6.2831853f);
  const float a1 = rng.Uniform(0.1f, 0.3f), a2 = rng.Uniform(0.02f, 0.12f);
  const float R = 0.5f * b;
  auto anchor = [&](int k, float* x, float* y) {
    float th = 6.2831853f * (float)(k % segs) / (float)segs;
    float r = R * (0.6f + a1 * std::sin(m1 * th + ph1) + a2 * std::sin(m2 * th + ph2));
    *x = cx + r * std::cos(th);
    *y = cy + r * std::sin(th);
  };
  Builder& bd = out->b;
  bd.verbs.reserve(segs + 2);
  bd.pts.reserve(2 * (2 * segs + 1));
  float x0, y0;
  anchor(0, &x0, &y0);
  bd.Move(x0, y0);
  for (int k = 0; k < segs; k++) {
    float x1, y1;
    anchor(k + 1, &x1, &y1);
    float dx = x1 - x0, dy = y1 - y0;
    float nx = -dy, ny = dx;  // chord normal, chord length
    float u = rng.Unit();
    if (u < 0.4f) {
      bd.Line(x1, y1);
    } else if (u < 0.7f) {
      float o = rng.Uniform(-0.3f, 0.3f);
      bd.Quad(x0 + 0.5f * dx + o * nx, y0 + 0.5f * dy + o * ny, x1, y1);
    } else {
      float o1 = rng.Uniform(-0.3f, 0.3f), o2 = rng.Uniform(-0.3f, 0.3f);
      bd.Cubic(x0 + dx / 3.f + o1 * nx, y0 + dy / 3.f + o1 * ny, x0 + 2.f * dx / 3.f + o2 * nx,
               y0 + 2.f * dy / 3.f + o2 * ny, x1, y1);
    }
    x0 = x1;
    y0 = y1;
  }
  bd.Close();
  out->fill = rng.Int(0, 1) ? HIPR_FILL_EVEN_ODD : HIPR_FILL_WINDING;
  out->paint = SolidPaint(rng, 77);
  out->gradient = rng.Unit() < 0.3f;
  if (out->gradient) {
    sg_scene_s tmp;
    AddGradient(&tmp, rng, cx, cy, b, &out->paint);
    out->g = tmp.gradients[0];
  }
}

// side > 0 fixes the canvas side; first_path offsets the path indices (so
// that a large scene can be generated as consecutive slices).
void GenOutlines(sg_scene_s* s, Pcg32& rng, uint64_t seed, int n, int side, float box_lo,
                 float box_hi, int first_path) {
  if (side <= 0) side = 1 << rng.Int(8, 12);
  s->width = side;
  s->height = side;
  const int paths = n > 0 ? n : (int)std::lround(std::pow(10.f, rng.Uniform(2.f, 4.699f)));
  const int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int chunk = 1024;
  std::vector<OutlinePath> buf;
  for (int p0 = 0; p0 < paths; p0 += chunk) {
    const int m = std::min(chunk, paths - p0);
    buf.assign(m, OutlinePath{});
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
      pool.emplace_back([&]() {
        for (int i; (i = next.fetch_add(1)) < m;)
          GenOutlinePath(seed, first_path + p0 + i, side, box_lo, box_hi, &buf[i]);
      });
    for (auto& th : pool) th.join();
    for (int i = 0; i < m; i++) {
      OutlinePath& op = buf[i];
      if (op.gradient) {
        op.paint.gradient = (int32_t)s->gradients.size();
        s->gradients.push_back(op.g);
      }
      s->Add(op.b, op.fill, op.paint);
    }
  }
}

}  // namespace

extern "C" {

sg_scene sg_generate(int32_t workload, uint64_t seed, int32_t n_paths) {
  sg_scene_s* s = new sg_scene_s();
  Pcg32 rng(seed);
  switch (workload) {
    case 1:
      GenPolygons(s, rng, n_paths > 0 ? n_paths : 64);
      break;
    case 2:
      GenCurves(s, rng, n_paths > 0 ? n_paths : 4096, 2048.f);
      break;
    case 3:
      GenGradientBlur(s, rng, n_paths > 0 ? n_paths : 1024);
      break;
    case 4:  // one canvas of the batched uneven family at its stated scale
      GenOutlines(s, rng, seed, n_paths, 0, -5.f, -1.f, 0);
      break;
    case 5:  // the same family scaled: 10-1,000 paths of 4-316 segments
      GenUneven(s, rng, n_paths, 1.f, 3.f, 2.5f);
      break;
    default:
      delete s;
      return nullptr;
  }
  s->Finish();
  return s;
}

void sg_destroy(sg_scene s) { delete s; }

const hipr_scene* sg_view(sg_scene s) { return s ? &s->view : nullptr; }

void sg_canvas_size(sg_scene s, int32_t* w, int32_t* h) {
  *w = s->width;
  *h = s->height;
}

void sg_struct_sizes(int32_t* paint, int32_t* gradient, int32_t* scene) {
  *paint = (int32_t)sizeof(hipr_paint);
  *gradient = (int32_t)sizeof(hipr_gradient);
  *scene = (int32_t)sizeof(hipr_scene);
}

sg_scene sg_slice(sg_scene s, int32_t begin, int32_t end) {
  sg_scene_s* r = new sg_scene_s();
  r->width = s->width;
  r->height = s->height;
  r->gradients = s->gradients;
  for (int p = begin; p < end && p < (int)s->paints.size(); p++) {
    Builder b;
    b.verbs.assign(s->verbs.begin() + s->verb_off[p],
                   s->verbs.begin() + s->verb_off[p + 1]);
    b.pts.assign(s->points.begin() + 2 * s->point_off[p],
                 s->points.begin() + 2 * s->point_off[p + 1]);
    b.weights.assign(s->weights.begin() + s->weight_off[p],
                     s->weights.begin() + s->weight_off[p + 1]);
    r->Add(b, s->fills[p], s->paints[p]);
  }
  r->Finish();
  return r;
}

sg_scene sg_generate_outlines(uint64_t seed, int32_t side, int32_t first_path, int32_t n_paths,
                              float box_lo, float box_hi) {
  if (side <= 0 || n_paths <= 0 || first_path < 0) return nullptr;
  sg_scene_s* s = new sg_scene_s();
  Pcg32 rng(seed);
  GenOutlines(s, rng, seed, n_paths, side, box_lo, box_hi, first_path);
  s->Finish();
  return s;
}

}  // extern "C"
