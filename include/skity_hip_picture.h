/* skity_hip_picture.h -- C binding of the SKP reader and display lists of
 * the C++ front end (libskity_hip_canvas.so), for callers that drive the
 * HIP canvas from recorded scenes.
 *
 * skity_hip_picture_make       Picture::MakeFromStream
 *                              (module/io/include/skity/io/picture.hpp:70)
 * skity_hip_picture_record     PictureRecorder::BeginRecording(w, h) +
 *                              Canvas::Concat(matrix) + Picture::PlayBack +
 *                              FinishRecording (picture.hpp:76,
 *                              include/skity/recorder/picture_recorder.hpp)
 * skity_hip_display_list_draw  DisplayList::Draw onto a surface's canvas +
 *                              Flush (include/skity/recorder/display_list.hpp:76)
 * skity_hip_display_list_scene the hipr_scene that draw hands to
 *                              hipr_surface_draw_scene
 *
 * Results are hipr_result codes.  HIPR_ERROR_NOT_SUPPORTED from
 * skity_hip_picture_make names SKP contents the reader does not parse
 * (skity_hip_picture_last_error says which); from the display-list calls it
 * means the recording holds a call the backend does not draw (span-list clips, save
 * layers, text, images, shadows, arcs, paint effects other
 * than a blend mode), and nothing is drawn. */
#ifndef SKITY_HIP_PICTURE_H_
#define SKITY_HIP_PICTURE_H_

#include <stddef.h>
#include <stdint.h>

#include "skity_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skity_hip_picture_s* skity_hip_picture;
typedef struct skity_hip_display_list_s* skity_hip_display_list;

int skity_hip_picture_make(const uint8_t* data, size_t size, skity_hip_picture* out);
const char* skity_hip_picture_last_error(void);
/* cull rect {l, t, r, b}, SKP version, paint and path table sizes; any
 * output may be NULL */
int skity_hip_picture_info(skity_hip_picture pic, float cull_rect[4], uint32_t* version,
                           int32_t* paints, int32_t* paths);
void skity_hip_picture_destroy(skity_hip_picture pic);

/* matrix: {scale_x, skew_x, trans_x, skew_y, scale_y, trans_y} concatenated
 * before playback (NULL = none) */
int skity_hip_picture_record(skity_hip_picture pic, int32_t width, int32_t height,
                             const float matrix[6], skity_hip_display_list* out);
void skity_hip_display_list_destroy(skity_hip_display_list dl);
/* The scene stays valid until the display list is destroyed. */
int skity_hip_display_list_scene(skity_hip_display_list dl, const hipr_scene** scene);
/* mode: HIPR_MODE_* */
int skity_hip_display_list_draw(skity_hip_display_list dl, hipr_surface surface, int32_t mode);

#ifdef __cplusplus
}
#endif

#endif /* SKITY_HIP_PICTURE_H_ */
