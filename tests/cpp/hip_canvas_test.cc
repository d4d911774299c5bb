// Known-answer tests of the reference's SWCanvas / Path suites, written
// against the C++ front end (skity_amd/host/hip_canvas.hpp) and drawn by the
// HIP library.  Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../skity_amd/host/hip_canvas.hpp"

using namespace skity_hip;

static int g_failed = 0;
#define EXPECT_EQ(a, b)                                                              \
  do {                                                                               \
    auto va = (a);                                                                   \
    auto vb = (b);                                                                   \
    if (!(va == vb)) {                                                               \
      std::printf("  FAILED %s:%d: %s == %s (0x%08x vs 0x%08x)\n", __FILE__, __LINE__, #a, \
                  #b, (unsigned)va, (unsigned)vb);                                   \
      g_failed++;                                                                    \
    }                                                                                \
  } while (0)
#define EXPECT_TRUE(c)                                                               \
  do {                                                                               \
    if (!(c)) {                                                                      \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);                   \
      g_failed++;                                                                    \
    }                                                                                \
  } while (0)

// test/ut/render/sw_canvas_test.cc:33-48 (the draw inside the save layer)
static void RedRect(hipr_context ctx) {
  auto surface = HIPSurface::Make(ctx, 16, 16);
  HIPCanvas* canvas = surface->LockCanvas();
  Paint paint;
  paint.SetColor(Color_RED);
  canvas->DrawRect(Rect::MakeXYWH(4.f, 4.f, 4.f, 4.f), paint);
  EXPECT_EQ(surface->Flush(), HIPR_SUCCESS);
  EXPECT_EQ(surface->GetPixel(5, 5), Color_RED);
  EXPECT_EQ(surface->GetPixel(0, 0), Color_TRANSPARENT);
}

// test/ut/render/sw_canvas_test.cc:67-85
static void StrokeThenFillDrawsFillAfterStroke(hipr_context ctx) {
  auto surface = HIPSurface::Make(ctx, 48, 48);
  HIPCanvas* canvas = surface->LockCanvas();
  Paint paint;
  paint.SetStyle(Paint::kStrokeThenFill_Style);
  paint.SetStrokeWidth(10.f);
  paint.SetStrokeColor(Color_RED);
  paint.SetFillColor(Color_WHITE);
  paint.SetAntiAlias(false);
  Path path;
  path.AddRect(Rect::MakeXYWH(10.f, 10.f, 24.f, 24.f));
  canvas->DrawPath(path, paint);
  EXPECT_EQ(surface->Flush(), HIPR_SUCCESS);
  EXPECT_EQ(surface->GetPixel(12, 20), Color_WHITE);
  EXPECT_EQ(surface->GetPixel(7, 20), Color_RED);
}

// test/ut/graphic/path_test.cc:1309-1319: the width-2 stroke of the quad
// contains (120, 30); drawn, the pixels on both sides of y = 30 just left of
// x = 120 are covered and the pixel right of the butt end is not.
static void StrokePathContainsEnd(hipr_context ctx) {
  auto surface = HIPSurface::Make(ctx, 130, 40);
  HIPCanvas* canvas = surface->LockCanvas();
  Path src;
  src.MoveTo(30, 30);
  src.QuadTo(40, 30, 120, 30);
  Paint paint;
  paint.SetStyle(Paint::kStroke_Style);
  paint.SetStrokeWidth(2);
  canvas->DrawPath(src, paint);
  EXPECT_EQ(surface->Flush(), HIPR_SUCCESS);
  EXPECT_EQ(surface->GetPixel(119, 29), Color_BLACK);
  EXPECT_EQ(surface->GetPixel(119, 30), Color_BLACK);
  EXPECT_EQ(surface->GetPixel(120, 30), Color_TRANSPARENT);
}

// Canvas::DrawCircle / DrawLine dispatch to DrawPath (canvas.cc:398-437).
static void ShapesDispatchToDrawPath(hipr_context ctx) {
  auto a = HIPSurface::Make(ctx, 64, 64);
  auto b = HIPSurface::Make(ctx, 64, 64);
  Paint paint;
  paint.SetColor(ColorSetARGB(200, 30, 160, 90));
  a->LockCanvas()->DrawCircle(31.3f, 30.6f, 17.2f, paint);
  Path circle;
  circle.AddOval(Rect::MakeLTRB(31.3f - 17.2f, 30.6f - 17.2f, 31.3f + 17.2f, 30.6f + 17.2f),
                 Path::Direction::kCW, 3);
  b->LockCanvas()->DrawPath(circle, paint);
  Paint line;
  line.SetStyle(Paint::kStroke_Style);
  line.SetStrokeWidth(3.5f);
  line.SetStrokeCap(Paint::kRound_Cap);
  line.SetColor(Color_BLUE);
  a->LockCanvas(false)->DrawLine(4, 60, 60, 4, line);
  Path lp;
  lp.MoveTo(4, 60);
  lp.LineTo(60, 4);
  b->LockCanvas(false)->DrawPath(lp, line);
  EXPECT_EQ(a->Flush(), HIPR_SUCCESS);
  EXPECT_EQ(b->Flush(), HIPR_SUCCESS);
  int diff = 0;
  for (int y = 0; y < 64; y++)
    for (int x = 0; x < 64; x++) diff += a->GetPixel(x, y) != b->GetPixel(x, y);
  EXPECT_EQ(diff, 0);
  // inside the circle, away from the line: 200/255 alpha over transparent
  EXPECT_EQ(ColorGetA(a->GetPixel(31, 20)), 200u);
}

// Shader::MakeLinear end colours under clamp tiling.
static void LinearGradientEnds(hipr_context ctx) {
  auto s = HIPSurface::Make(ctx, 100, 8);
  Point pts[2] = {{10, 0}, {90, 0}};
  Color4f colors[2] = {{1, 0, 0, 1}, {0, 0, 1, 1}};
  Paint paint;
  paint.SetShader(Shader::MakeLinear(pts, colors, nullptr, 2));
  s->LockCanvas()->DrawRect(Rect::MakeWH(100, 8), paint);
  EXPECT_EQ(s->Flush(), HIPR_SUCCESS);
  EXPECT_EQ(s->GetPixel(2, 4), Color_RED);
  EXPECT_EQ(s->GetPixel(97, 4), Color_BLUE);
}

// Canvas matrix: the red rect of RedRect moved by Translate and doubled by
// Scale (Canvas::Translate / Scale pre-concatenate, canvas_state.cc:18-24).
static void MatrixMovesAndScalesDraws(hipr_context ctx) {
  auto surface = HIPSurface::Make(ctx, 32, 32);
  HIPCanvas* canvas = surface->LockCanvas();
  Paint paint;
  paint.SetColor(Color_RED);
  canvas->Save();
  canvas->Translate(10.f, 6.f);
  canvas->Scale(2.f, 2.f);
  canvas->DrawRect(Rect::MakeXYWH(1.f, 1.f, 4.f, 4.f), paint);  // -> (12, 8)-(20, 16)
  canvas->Restore();
  canvas->DrawRect(Rect::MakeXYWH(0.f, 24.f, 4.f, 4.f), paint);  // identity again
  EXPECT_EQ(canvas->GetSaveCount(), 1);
  EXPECT_EQ(surface->Flush(), HIPR_SUCCESS);
  EXPECT_EQ(surface->GetPixel(12, 8), Color_RED);
  EXPECT_EQ(surface->GetPixel(19, 15), Color_RED);
  EXPECT_EQ(surface->GetPixel(20, 16), Color_TRANSPARENT);
  EXPECT_EQ(surface->GetPixel(11, 8), Color_TRANSPARENT);
  EXPECT_EQ(surface->GetPixel(1, 25), Color_RED);
}

// Canvas::Clear is DrawColor(c, kSrc) and replaces the pixels; DrawColor with
// kSrcOver blends (canvas.cc:241-251, SWCanvas::OnDrawPaint).
static void ClearAndDrawColor(hipr_context ctx) {
  for (int ct : {HIPR_COLOR_RGBA, HIPR_COLOR_BGRA}) {
    auto surface = HIPSurface::Make(ctx, 8, 8, HIPR_ALPHA_PREMUL, ct);
    HIPCanvas* canvas = surface->LockCanvas();
    canvas->Clear(Color_BLUE);
    EXPECT_EQ(surface->Flush(), HIPR_SUCCESS);
    EXPECT_EQ(surface->GetPixel(3, 3), Color_BLUE);
    canvas->Clear(Color_TRANSPARENT);
    canvas->DrawColor(Color_GREEN, BlendMode::kDstOver);
    Paint red;
    red.SetColor(Color_RED);
    canvas->DrawRect(Rect::MakeXYWH(0.f, 0.f, 4.f, 8.f), red);
    EXPECT_EQ(surface->Flush(), HIPR_SUCCESS);
    EXPECT_EQ(surface->GetPixel(1, 1), Color_RED);
    EXPECT_EQ(surface->GetPixel(6, 1), Color_GREEN);
  }
}

// Width 0 is stroked as SWCanvas strokes it (PathStroker at radius 0): an
// open line's outline has no area, so nothing is drawn.  A negative width is
// refused, not approximated.
static void WidthZeroAndNegativeStrokes(hipr_context ctx) {
  auto s = HIPSurface::Make(ctx, 16, 16);
  Paint paint;
  paint.SetStyle(Paint::kStroke_Style);
  paint.SetStrokeWidth(0);
  s->LockCanvas()->DrawLine(1, 1, 10, 10, paint);
  EXPECT_EQ(s->Flush(), HIPR_SUCCESS);
  for (int y = 0; y < 16; y++)
    for (int x = 0; x < 16; x++) EXPECT_EQ(s->GetPixel(x, y), Color_TRANSPARENT);
  paint.SetStrokeWidth(-1);
  s->LockCanvas()->DrawLine(1, 1, 10, 10, paint);
  EXPECT_EQ(s->Flush(), HIPR_ERROR_NOT_SUPPORTED);
}

// Shader::MakeTwoPointConical's routing (src/effect/shader.cc:50-85).
static void TwoPointConicalFactory(hipr_context) {
  Color4f c[2] = {{1, 0, 0, 1}, {0, 0, 1, 1}};
  float pos[2] = {0.f, 0.25f};
  Point p{10, 20}, q{30, 25};
  EXPECT_EQ(Shader::MakeTwoPointConical(p, 5, p, 5, c, pos, 2) == nullptr, true);
  auto a = Shader::MakeTwoPointConical(p, 0, p, 40, c, pos, 2);
  EXPECT_EQ(a->info().type, HIPR_GRADIENT_RADIAL);
  EXPECT_EQ(a->info().radius[0], 40.f);
  auto b = Shader::MakeTwoPointConical(p, 30, p, 0, c, pos, 2);
  EXPECT_EQ(b->info().type, HIPR_GRADIENT_RADIAL);
  EXPECT_EQ(b->info().radius[0], 30.f);
  EXPECT_EQ(b->info().colors[0][2], 1.f);     // colours reversed
  EXPECT_EQ(b->info().offsets[0], 0.75f);     // 1 - 0.25
  EXPECT_EQ(b->info().offsets[1], 1.f);       // 1 - 0
  auto g = Shader::MakeTwoPointConical(p, 5, q, 20, c, pos, 2);
  EXPECT_EQ(g->info().type, HIPR_GRADIENT_CONICAL);
  EXPECT_EQ(g->info().points[2], 30.f);
  EXPECT_EQ(g->info().radius[1], 20.f);
}

// A pixmap shader through Shader::MakeShader: an image drawn over a rect of
// its own size at an integer offset, nearest sampling, reproduces its texels
// (through the Color4f round trip, exact for 0 and 255).
static void PixmapShaderDrawsTheImage(hipr_context ctx) {
  const int w = 4, h = 3;
  uint8_t px[w * h * 4];
  for (int i = 0; i < w * h; i++) {
    px[4 * i + 0] = (i & 1) ? 255 : 0;
    px[4 * i + 1] = (i & 2) ? 255 : 0;
    px[4 * i + 2] = (i & 4) ? 255 : 0;
    px[4 * i + 3] = 255;
  }
  auto s = HIPSurface::Make(ctx, 16, 16);
  Matrix local = Matrix::Translate(5, 6);
  Paint paint;
  paint.SetShader(Shader::MakeShader(Image::MakeImage(px, w, h), SamplingOptions{},
                                     TileMode::kDecal, TileMode::kDecal, &local));
  s->LockCanvas()->DrawRect(Rect::MakeXYWH(0.f, 0.f, 16.f, 16.f), paint);
  EXPECT_EQ(s->Flush(), HIPR_SUCCESS);
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const uint8_t* q = px + 4 * (y * w + x);
      Color want = (0xFFu << 24) | (q[0] << 16) | (q[1] << 8) | q[2];
      EXPECT_EQ(s->GetPixel(5 + x, 6 + y), want);
    }
  EXPECT_EQ(s->GetPixel(4, 6), Color_TRANSPARENT);
  EXPECT_EQ(s->GetPixel(5 + w, 6), Color_TRANSPARENT);
}

// PictureRecorder + DisplayList::Draw (display_list.cc:330-343): replaying
// the recording onto a canvas under a matrix draws what the same calls made
// directly under that matrix draw; a recorded difference clip makes the
// flush NOT_SUPPORTED and draws nothing.
static void DisplayListReplayEqualsDirectDraws(hipr_context ctx) {
  auto calls = [](HIPCanvas* c) {
    Paint red;
    red.SetColor(Color_RED);
    Paint blue;
    blue.SetColor(0x800000FF);
    blue.SetStyle(Paint::kStroke_Style);
    blue.SetStrokeWidth(3.f);
    c->Save();
    c->Translate(10.f, 5.f);
    c->Rotate(20.f);
    c->DrawRect(Rect::MakeXYWH(4.f, 4.f, 30.f, 20.f), red);
    c->Restore();
    c->Scale(1.5f, 1.25f);
    c->DrawCircle(30.f, 30.f, 12.f, blue);
    c->DrawPaint(blue);
  };
  PictureRecorder rec;
  rec.BeginRecording(Rect::MakeWH(64.f, 64.f));
  calls(rec.GetRecordingCanvas());
  std::unique_ptr<DisplayList> dl = rec.FinishRecording();
  EXPECT_EQ(dl->OpCount(), 8u);

  auto a = HIPSurface::Make(ctx, 64, 64);
  auto b = HIPSurface::Make(ctx, 64, 64);
  HIPCanvas* ca = a->LockCanvas();
  ca->Translate(3.f, -2.f);
  dl->Draw(ca);
  EXPECT_EQ(a->Flush(), HIPR_SUCCESS);
  HIPCanvas* cb = b->LockCanvas();
  cb->Translate(3.f, -2.f);
  calls(cb);
  EXPECT_EQ(b->Flush(), HIPR_SUCCESS);
  std::vector<uint8_t> pa(64 * 64 * 4), pb(64 * 64 * 4);
  EXPECT_EQ(a->ReadPixels(pa.data(), pa.size()), HIPR_SUCCESS);
  EXPECT_EQ(b->ReadPixels(pb.data(), pb.size()), HIPR_SUCCESS);
  EXPECT_TRUE(pa == pb);
  EXPECT_TRUE(pa[4 * (20 * 64 + 20) + 3] != 0);

  PictureRecorder rec2;
  rec2.BeginRecording(Rect::MakeWH(64.f, 64.f));
  calls(rec2.GetRecordingCanvas());
  rec2.GetRecordingCanvas()->ClipRect(Rect::MakeWH(8.f, 8.f), HIPCanvas::ClipOp::kDifference);
  auto dl2 = rec2.FinishRecording();
  HIPCanvas* cc = a->LockCanvas();
  dl2->Draw(cc);
  EXPECT_EQ(a->Flush(), HIPR_ERROR_NOT_SUPPORTED);
  EXPECT_EQ(a->GetPixel(20, 20), Color_TRANSPARENT);
}

// DrawGlyphs (sw_canvas.cc:542-639): a glyph run equals its glyphs drawn as
// paths under Translate(x, y), all fills before all strokes; a run with a
// shader is refused.
static void GlyphRunEqualsTranslatedPaths(hipr_context ctx) {
  Path a, b;
  a.AddRect(Rect::MakeLTRB(0.f, -10.f, 8.f, 0.f));
  b.MoveTo(0.f, 0.f).LineTo(6.f, -12.f).LineTo(12.f, 0.f).Close();
  const Path* glyphs[2] = {&a, &b};
  const float xs[2] = {5.f, 9.f}, ys[2] = {20.f, 21.5f};
  Paint paint;
  paint.SetColor(0xC0204080);
  paint.SetStyle(Paint::kStrokeAndFill_Style);
  paint.SetStrokeWidth(2.f);
  paint.SetStrokeColor(0xFFFFA000);
  auto s1 = HIPSurface::Make(ctx, 40, 40);
  auto s2 = HIPSurface::Make(ctx, 40, 40);
  HIPCanvas* c1 = s1->LockCanvas();
  c1->Rotate(5.f);
  c1->DrawGlyphs(2, glyphs, xs, ys, paint);
  EXPECT_EQ(s1->Flush(), HIPR_SUCCESS);
  HIPCanvas* c2 = s2->LockCanvas();
  c2->Rotate(5.f);
  Paint fill = paint, stroke = paint;
  fill.SetStyle(Paint::kFill_Style);
  stroke.SetStyle(Paint::kStroke_Style);
  for (const Paint* p : {&fill, &stroke})
    for (int k = 0; k < 2; k++) {
      c2->Save();
      c2->Translate(xs[k], ys[k]);
      c2->DrawPath(*glyphs[k], *p);
      c2->Restore();
    }
  EXPECT_EQ(s2->Flush(), HIPR_SUCCESS);
  std::vector<uint8_t> p1(40 * 40 * 4), p2(40 * 40 * 4);
  EXPECT_EQ(s1->ReadPixels(p1.data(), p1.size()), HIPR_SUCCESS);
  EXPECT_EQ(s2->ReadPixels(p2.data(), p2.size()), HIPR_SUCCESS);
  EXPECT_TRUE(p1 == p2);
  EXPECT_TRUE(s1->GetPixel(9, 15) != Color_TRANSPARENT);
  Point pts[2] = {{0.f, 0.f}, {10.f, 0.f}};
  Color4f cols[2] = {{1, 0, 0, 1}, {0, 0, 1, 1}};
  Paint shaded;
  shaded.SetShader(Shader::MakeLinear(pts, cols, nullptr, 2));
  s1->LockCanvas()->DrawGlyphs(2, glyphs, xs, ys, shaded);
  EXPECT_EQ(s1->Flush(), HIPR_ERROR_NOT_SUPPORTED);
}

// Canvas::ClipRect (canvas.cc:130-135, sw_canvas.cc:297-313): an intersect
// clip under scale + translate limits later path draws to the floored /
// ceiled clip bounds, Restore lifts it, DrawPaint ignores it; a difference
// clip makes the flush NOT_SUPPORTED.
static void ClipRectLimitsDraws(hipr_context ctx) {
  auto s = HIPSurface::Make(ctx, 40, 40);
  HIPCanvas* c = s->LockCanvas();
  Paint red;
  red.SetColor(Color_RED);
  c->Save();
  c->Translate(2.f, 1.f);
  c->Scale(2.f, 2.f);
  c->ClipRect(Rect::MakeLTRB(2.2f, 3.f, 8.f, 9.6f));  // device [6.4, 7) - (18, 20.2)
  c->DrawRect(Rect::MakeWH(40.f, 40.f), red);
  c->Restore();
  EXPECT_EQ(s->Flush(), HIPR_SUCCESS);
  EXPECT_EQ(s->GetPixel(6, 7), Color_RED);
  EXPECT_EQ(s->GetPixel(17, 20), Color_RED);
  EXPECT_EQ(s->GetPixel(5, 10), Color_TRANSPARENT);
  EXPECT_EQ(s->GetPixel(18, 10), Color_TRANSPARENT);
  EXPECT_EQ(s->GetPixel(10, 6), Color_TRANSPARENT);
  EXPECT_EQ(s->GetPixel(10, 21), Color_TRANSPARENT);
  c = s->LockCanvas();
  c->ClipRect(Rect::MakeLTRB(0.f, 0.f, 4.f, 4.f));
  c->DrawColor(Color_BLUE);  // DrawPaint: not clipped
  EXPECT_EQ(s->Flush(), HIPR_SUCCESS);
  EXPECT_EQ(s->GetPixel(30, 30), Color_BLUE);
  c = s->LockCanvas();
  c->ClipRect(Rect::MakeLTRB(0.f, 0.f, 4.f, 4.f), HIPCanvas::ClipOp::kDifference);
  c->DrawRect(Rect::MakeWH(40.f, 40.f), red);
  EXPECT_EQ(s->Flush(), HIPR_ERROR_NOT_SUPPORTED);
}

// RRect normalisation, the reference's own known answers
// (test/ut/geometry/rrect_test.cc:101-241,450-471), and the outline
// Path::AddRRect builds for a simple rrect (path.cc:881-934: from the legacy
// start index 6, line / conic pairs around the eight radii points).
static void RRectKnownAnswers(hipr_context) {
  const Rect rect = Rect::MakeWH(100.f, 100.f);
  RRect r1;
  r1.SetRectXY(rect, 0.f, 0.f);
  EXPECT_EQ(r1.GetType(), RRect::kRect);
  const Point zero[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  RRect r2;
  r2.SetRectRadii(rect, zero);
  EXPECT_EQ(r2.GetType(), RRect::kRect);
  const Point mixed[4] = {{0, 0}, {20, 20}, {50, 50}, {20, 50}};
  RRect r3;
  r3.SetRectRadii(rect, mixed);
  EXPECT_EQ(r3.GetType(), RRect::kComplex);
  RRect oval;
  oval.SetRectXY(rect, 50.f, 50.f);
  EXPECT_EQ(oval.GetType(), RRect::kOval);
  RRect simple;
  simple.SetRectXY(rect, 20.f, 20.f);
  EXPECT_EQ(simple.GetType(), RRect::kSimple);
  RRect inf;
  inf.SetRectXY(rect, INFINITY, INFINITY);
  EXPECT_EQ(inf.GetType(), RRect::kRect);
  RRect big;
  big.SetRectXY(rect, 100.f, 100.f);
  EXPECT_EQ(big.GetType(), RRect::kOval);
  const Point wide[4] = {{50, 100}, {100, 50}, {50, 100}, {100, 50}};
  RRect scaled;
  scaled.SetRectRadii(rect, wide);
  EXPECT_EQ(scaled.GetType(), RRect::kComplex);
  Point ul = scaled.Radii(RRect::kUpperLeft);
  EXPECT_TRUE(std::fabs(ul.x - 33.333333f) <= 1.f / 4096 && std::fabs(ul.y - 66.666666f) <= 1.f / 4096);
  RRect neg;
  neg.SetRectXY(rect, -10.f, -20.f);
  EXPECT_EQ(neg.GetType(), RRect::kRect);
  const Point one[4] = {{0, 0}, {48.8800011f, 48.8800011f}, {0, 0}, {0, 0}};
  RRect sr;
  sr.SetRectRadii(Rect::MakeXYWH(1.f, 1.12f, 47.7599983f, 47.7599983f), one);
  EXPECT_EQ(sr.GetType(), RRect::kComplex);
  // EXPECT_FLOAT_EQ: within 4 ULPs
  auto float_eq = [](float a, float b) {
    int32_t ia, ib;
    std::memcpy(&ia, &a, 4);
    std::memcpy(&ib, &b, 4);
    return std::abs(ia - ib) <= 4;
  };
  EXPECT_TRUE(float_eq(sr.Radii(RRect::kUpperRight).x, 47.7599983f));
  EXPECT_TRUE(float_eq(sr.Radii(RRect::kUpperRight).y, 47.7599983f));
  EXPECT_TRUE(sr.Radii(RRect::kUpperLeft).x == 0.f && sr.Radii(RRect::kLowerLeft).y == 0.f);
  // AddRoundRect(10, 20, 100, 200, 10, 10): M L C L C L C L C Z
  Path p;
  p.AddRoundRect(Rect::MakeLTRB(10, 20, 100, 200), 10, 10);
  const uint8_t want_v[10] = {0, 1, 3, 1, 3, 1, 3, 1, 3, 5};
  EXPECT_EQ(p.verbs().size(), 10u);
  bool verbs_ok = p.verbs().size() == 10;
  for (size_t i = 0; verbs_ok && i < 10; i++) verbs_ok = p.verbs()[i] == want_v[i];
  EXPECT_TRUE(verbs_ok);
  const float want_p[17][2] = {{10, 190}, {10, 30},  {10, 20},  {20, 20},  {90, 20},  {100, 20},
                               {100, 30}, {100, 190}, {100, 200}, {90, 200}, {20, 200}, {10, 200},
                               {10, 190}};
  bool pts_ok = p.points().size() == 13;
  for (size_t i = 0; pts_ok && i < 13; i++)
    pts_ok = p.points()[i].x == want_p[i][0] && p.points()[i].y == want_p[i][1];
  EXPECT_TRUE(pts_ok);
}

int main() {
  hipr_context ctx = nullptr;
  if (hipr_context_create(0, &ctx) != HIPR_SUCCESS) {
    std::printf("no device\n");
    return 2;
  }
  struct {
    const char* name;
    void (*fn)(hipr_context);
  } tests[] = {{"SWCanvas.RedRect", RedRect},
               {"SWCanvas.StrokeThenFillDrawsFillAfterStroke", StrokeThenFillDrawsFillAfterStroke},
               {"Path.StrokePathContainsEnd", StrokePathContainsEnd},
               {"Canvas.ShapesDispatchToDrawPath", ShapesDispatchToDrawPath},
               {"Shader.LinearGradientEnds", LinearGradientEnds},
               {"Canvas.MatrixMovesAndScalesDraws", MatrixMovesAndScalesDraws},
               {"Canvas.ClearAndDrawColor", ClearAndDrawColor},
               {"HIPCanvas.WidthZeroAndNegativeStrokes", WidthZeroAndNegativeStrokes},
               {"Shader.TwoPointConicalFactory", TwoPointConicalFactory},
               {"Shader.PixmapShaderDrawsTheImage", PixmapShaderDrawsTheImage},
               {"DisplayList.ReplayEqualsDirectDraws", DisplayListReplayEqualsDirectDraws},
               {"Canvas.GlyphRunEqualsTranslatedPaths", GlyphRunEqualsTranslatedPaths},
               {"Canvas.ClipRectLimitsDraws", ClipRectLimitsDraws},
               {"RRect.KnownAnswers", RRectKnownAnswers}};
  for (auto& t : tests) {
    int before = g_failed;
    t.fn(ctx);
    std::printf("[%s] %s\n", g_failed == before ? "  OK  " : "FAILED", t.name);
  }
  hipr_context_destroy(ctx);
  std::printf("%d check(s) failed\n", g_failed);
  return g_failed ? 1 : 0;
}
