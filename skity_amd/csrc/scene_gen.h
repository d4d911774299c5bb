/* scene_gen.h -- seeded synthetic DrawPath scenes (inputs for tests and the
 * benchmark).  Not part of the drop-in boundary: it only produces hipr_scene
 * arrays the way a skity application would build Paths and Paints.
 *
 * Random numbers: SplitMix64 seeds a PCG32 (XSH-RR) stream; the default seed
 * of workload k is 0x5EED0000 + k. */
#ifndef SKITY_HIP_SCENE_GEN_H_
#define SKITY_HIP_SCENE_GEN_H_

#include <stdint.h>

#include "../../include/skity_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sg_scene_s* sg_scene;

/* Workloads:
 *   1  plain fill: 64 closed random polygons (5-32 vertices) in 800x600,
 *      solid colours with alpha 77..255, nonzero winding;
 *   2  curves + strokes: quad/cubic/conic paths (8-200 segments, control
 *      points in 2048^2, conic weights 0.2-4), half even-odd, 20% stroked
 *      (width 1-16, miter/round/bevel joins, butt/round caps);
 *   3  gradients + blur: 4-cubic closed paths in boxes of side 32-1024 in
 *      4096^2 with linear/radial/sweep gradients (2-16 stops, clamp/repeat/
 *      mirror, random affine local matrix) and a normal blur, sigma 1-32;
 *   4  one canvas of the batched uneven family at its stated scale: side
 *      256-4096, 100-50,000 closed outlines of 4-10,000 line/quad/cubic
 *      segments (log-uniform), 30% gradients;
 *   5  uneven scribbles, scaled: side 256-4096, 10-1,000 paths of 4-316
 *      segments with control points anywhere in the canvas.
 * n_paths <= 0 selects the workload's default count. */
__attribute__((visibility("default"))) sg_scene sg_generate(int32_t workload,
                                                            uint64_t seed,
                                                            int32_t n_paths);
__attribute__((visibility("default"))) void sg_destroy(sg_scene s);
/* The scene view (valid until sg_destroy). */
__attribute__((visibility("default"))) const hipr_scene* sg_view(sg_scene s);
__attribute__((visibility("default"))) void sg_canvas_size(sg_scene s,
                                                           int32_t* w,
                                                           int32_t* h);
/* sizeof(hipr_paint), sizeof(hipr_gradient) and sizeof(hipr_scene) this
 * library was built with, so that a caller can refuse a library built from
 * another header. */
__attribute__((visibility("default"))) void sg_struct_sizes(int32_t* paint,
                                                            int32_t* gradient,
                                                            int32_t* scene);
/* Paths [first_path, first_path + n_paths) of the outline family of
 * workload 4 on a side x side canvas, with path boxes of side
 * side * 2^U[box_lo, box_hi]: a large scene (the learned pre-sizer's 1M-path
 * evaluation on 8192^2) is generated as consecutive slices, each its own
 * scene, which together are the same draws in the same order. */
__attribute__((visibility("default"))) sg_scene sg_generate_outlines(uint64_t seed,
                                                                     int32_t side,
                                                                     int32_t first_path,
                                                                     int32_t n_paths,
                                                                     float box_lo,
                                                                     float box_hi);
/* Keep only paths [begin, end) (for bounded samples). */
__attribute__((visibility("default"))) sg_scene sg_slice(sg_scene s,
                                                         int32_t begin,
                                                         int32_t end);

#ifdef __cplusplus
}
#endif

#endif /* SKITY_HIP_SCENE_GEN_H_ */
