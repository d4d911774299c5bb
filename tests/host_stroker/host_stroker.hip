// Host build of the device stroker (StrokeOne) for CPU comparison with the oracle.
#include <cstdint>
#include <vector>
#include "../../include/skity_hip.h"
#include "../../skity_amd/csrc/hipr_stroke.hpp"
using namespace hipr;
extern "C" __attribute__((visibility("default")))
int64_t hs_stroke_outline(const hipr_scene* s, int32_t p, uint8_t* verbs, int64_t vcap,
                          float* pts, int64_t pcap, float* ws, int64_t wcap, int64_t* counts) {
  Job job{};
  job.path = p;
  job.kind = kJobStroke;
  job.verb_begin = s->verb_offsets[p];
  job.verb_end = s->verb_offsets[p + 1];
  job.pt_begin = s->point_offsets[p];
  job.w_begin = s->weight_offsets[p];
  GeomSrc g{s->verbs, s->points, s->weights};
  StrokeOut so{};
  stroke::StrokeStack stack;
  StrokeOne(job, s->paints[p], g, &so, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
            &stack);
  int scratch = so.inner_nv > so.inner_np ? so.inner_nv : so.inner_np;
  if (so.inner_nw > scratch) scratch = so.inner_nw;
  std::vector<uint8_t> ov(so.nv + 1), sv(scratch + 1);
  std::vector<float> op(2 * so.np + 2), ow(so.nw + 1), sp(2 * scratch + 2), sw(scratch + 1);
  StrokeOut w = so;
  w.verb_begin = w.pt_begin = w.w_begin = w.scratch_begin = 0;
  StrokeOne(job, s->paints[p], g, &w, 1, ov.data(), op.data(), ow.data(), sv.data(), sp.data(),
            sw.data(), &stack);
  for (int i = 0; i < so.nv && i < vcap; i++) verbs[i] = ov[i];
  for (int i = 0; i < 2 * so.np && i < 2 * pcap; i++) pts[i] = op[i];
  for (int i = 0; i < so.nw && i < wcap; i++) ws[i] = ow[i];
  counts[0] = so.nv;
  counts[1] = so.np;
  counts[2] = so.nw;
  return so.nv;
}

// The same outline through the stroke kernel's chunked feed (segments
// stroked ahead, then taken by the stroker), its lanes run one after another.
extern "C" __attribute__((visibility("default")))
int64_t hs_stroke_outline_chunked(const hipr_scene* s, int32_t p, uint8_t* verbs, int64_t vcap,
                                  float* pts, int64_t pcap, float* ws, int64_t wcap,
                                  int64_t* counts) {
  Job job{};
  job.path = p;
  job.kind = kJobStroke;
  job.verb_begin = s->verb_offsets[p];
  job.verb_end = s->verb_offsets[p + 1];
  job.pt_begin = s->point_offsets[p];
  job.w_begin = s->weight_offsets[p];
  GeomSrc g{s->verbs, s->points, s->weights};
  std::vector<stroke::WaveLds> lds(1);
  stroke::WaveLds* L = lds.data();
  auto run = [&](StrokeOut* so, int mode, uint8_t* ov, float* op, float* ow, uint8_t* sv,
                 float* sp, float* sw) {
    StrokeOutline(job, s->paints[p], g, so, mode, ov, op, ow, sv, sp, sw, &L->stack.serial,
                  [&](stroke::Stroker* st, stroke::OpIter* it) {
                    stroke::FeedChunks(st, it, L, true, []() {}, [&](int nq) {
                      for (int l = 0; l < nq; l++) stroke::StrokeAhead(*st, L, l);
                    });
                  });
  };
  StrokeOut so{};
  run(&so, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  int scratch = so.inner_nv > so.inner_np ? so.inner_nv : so.inner_np;
  if (so.inner_nw > scratch) scratch = so.inner_nw;
  std::vector<uint8_t> ov(so.nv + 1), sv(scratch + 1);
  std::vector<float> op(2 * so.np + 2), ow(so.nw + 1), sp(2 * scratch + 2), sw(scratch + 1);
  StrokeOut w = so;
  w.verb_begin = w.pt_begin = w.w_begin = w.scratch_begin = 0;
  run(&w, 1, ov.data(), op.data(), ow.data(), sv.data(), sp.data(), sw.data());
  for (int i = 0; i < so.nv && i < vcap; i++) verbs[i] = ov[i];
  for (int i = 0; i < 2 * so.np && i < 2 * pcap; i++) pts[i] = op[i];
  for (int i = 0; i < so.nw && i < wcap; i++) ws[i] = ow[i];
  counts[0] = so.nv;
  counts[1] = so.np;
  counts[2] = so.nw;
  return so.nv;
}
