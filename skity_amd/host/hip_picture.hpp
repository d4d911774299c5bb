// hip_picture.hpp -- SKP pictures played back onto a HIPCanvas.
//
// Picture::MakeFromStream / Picture::PlayBack of skity's IO module
// (module/io/include/skity/io/picture.hpp:57-80, module/io/src/picture.cc,
// module/io/src/record/record_playback.cc): the SKP container (header,
// chunks, the paint and path tables of the flattened buffer, the op stream)
// is read here, and playback issues the same canvas calls in the same order
// as the reference's HandleOp.  Calls this backend cannot draw exactly
// (span-list clips, save layers, text, images, shadows, arcs, paints
// with effects other than a blend mode, inverse fill types) mark the canvas
// unsupported, so that its next Flush returns HIPR_ERROR_NOT_SUPPORTED; ops
// on which the reference ends playback end it here too.
#ifndef SKITY_HIP_HOST_HIP_PICTURE_HPP_
#define SKITY_HIP_HOST_HIP_PICTURE_HPP_

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "hip_canvas.hpp"

namespace skity_hip {

class Picture {
 public:
  // nullptr when the data is not an SKP this reader handles: bad magic,
  // a version outside 82..109, a truncated chunk, or contents it does not
  // parse (typefaces, text blobs, images, vertices, sub-pictures,
  // drawables).  Why is in LastError().
  static std::unique_ptr<Picture> MakeFromData(const uint8_t* data, size_t size);
  static std::unique_ptr<Picture> MakeFromFile(const char* path);
  static const std::string& LastError();

  // Canvas calls of the op stream, between a Save and a RestoreToCount
  // (picture.cc:215-247).
  void PlayBack(HIPCanvas* canvas) const;

  const Rect& CullRect() const { return cull_; }
  uint32_t Version() const { return version_; }
  size_t PaintCount() const { return paints_.size(); }
  size_t PathCount() const { return paths_.size(); }

 private:
  struct PaintEntry {
    Paint paint;
    bool supported = true;
  };
  struct PathEntry {
    Path path;
    bool supported = true;
  };
  friend struct PictureParser;
  std::vector<uint8_t> ops_;
  std::vector<PaintEntry> paints_;
  std::vector<PathEntry> paths_;
  std::vector<std::string> factories_;
  Rect cull_{0, 0, 0, 0};
  uint32_t version_ = 0;
};

}  // namespace skity_hip

#endif  // SKITY_HIP_HOST_HIP_PICTURE_HPP_
