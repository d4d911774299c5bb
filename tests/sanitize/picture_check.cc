// Sanitizer run of the SKP reader, its playback into display lists, the
// replay and the scene packing (skity_amd/host/hip_picture.cc,
// hip_canvas.cc), over a real SKP and seeded corruptions and truncations of
// it.  Only recording canvases are used: nothing reaches the device.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../skity_amd/host/hip_picture.hpp"

using namespace skity_hip;

static uint64_t g_state = 0x5EED5EEDull;
static uint32_t Next() {
  g_state = g_state * 6364136223846793005ull + 1442695040888963407ull;
  return (uint32_t)(g_state >> 33);
}

static size_t Exercise(const std::vector<uint8_t>& bytes) {
  auto pic = Picture::MakeFromData(bytes.data(), bytes.size());
  if (!pic) return 0;
  PictureRecorder rec;
  rec.BeginRecording(Rect::MakeWH(1000.f, 1000.f));
  HIPCanvas* c = rec.GetRecordingCanvas();
  c->Translate(-130.f, 20.f);
  c->ClipRect(Rect::MakeLTRB(10.5f, 20.f, 900.f, 950.25f));
  pic->PlayBack(c);
  Path glyph;
  glyph.MoveTo(0, 0).LineTo(5, -9).QuadTo(8, -12, 10, 0).Close();
  const Path* glyphs[2] = {&glyph, &glyph};
  const float xs[2] = {1.f, 12.f}, ys[2] = {30.f, 30.f};
  Paint p;
  p.SetStyle(Paint::kStrokeAndFill_Style);
  c->DrawGlyphs(2, glyphs, xs, ys, p);
  std::unique_ptr<DisplayList> dl = rec.FinishRecording();
  std::vector<HIPCanvas::Draw> draws;
  ReplayDraws(*dl, &draws);
  PackedScene packed(draws);
  const hipr_scene* s = packed.scene();
  size_t sum = 0;
  for (int i = 0; i < s->n_paths; i++) {
    for (int v = s->verb_offsets[i]; v < s->verb_offsets[i + 1]; v++) sum += s->verbs[v];
    if (s->clip_rects) sum += (size_t)(s->clip_rects[4 * i] > 0);
  }
  return sum + (size_t)s->n_paths;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: picture_check file.skp\n");
    return 2;
  }
  std::FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<uint8_t> good;
  int ch;
  while ((ch = std::fgetc(f)) != EOF) good.push_back((uint8_t)ch);
  std::fclose(f);
  size_t base = Exercise(good);
  if (base == 0) {
    std::printf("the SKP did not parse: %s\n", Picture::LastError().c_str());
    return 1;
  }
  int parsed = 0;
  for (int k = 0; k < 400; k++) {
    std::vector<uint8_t> b = good;
    if (k % 4 == 0) {
      b.resize(Next() % good.size());
    } else {
      int flips = 1 + (int)(Next() % 16);
      for (int i = 0; i < flips; i++) b[29 + Next() % (good.size() - 29)] = (uint8_t)Next();
    }
    parsed += Exercise(b) != 0;
  }
  std::printf("tiger draws packed: %zu; %d of 400 corrupted files parsed\n", base, parsed);
  std::printf("picture run clean\n");
  return 0;
}
