// Subset of ugu/image_io.h (include/ugu/image_io.h:20-72 in UGU) without
// OpenCV: PNG read/write (8-bit 1/3/4 channels, 16-bit 1 channel) and raw .bin
// write, on the typed images of ugu/image.h.
#pragma once

#include <string>
#include <vector>

#include "ugu/image.h"

namespace ugu {

enum ImreadModes {
  IMREAD_UNCHANGED = -1,
  IMREAD_GRAYSCALE = 0,
  IMREAD_COLOR = 1,
  IMREAD_ANYDEPTH = 2,
};

// imwrite (src/image_io.cc:182-206): the extension picks the format; .png
// (8-bit 1/3/4 channels, or 16-bit 1 channel as lodepng writes it), .bin (raw
// bytes). JPEG is not available in this build (LOGE, false).
bool imwrite(const std::string& filename, const Image1b& img);
bool imwrite(const std::string& filename, const Image3b& img);
bool imwrite(const std::string& filename, const Image4b& img);
bool imwrite(const std::string& filename, const Image1w& img);

// imread + Imread<T> (src/image_io.cc:208-222, include/ugu/image_io.h:58-70):
// decodes a PNG with its own channel count and bit depth (as stbi_load /
// stbi_load_16 with no requested channels, palette expanded); IMREAD_COLOR
// drops the alpha of a 4-channel image. An image whose channel count or
// sample size differs from T gives LOGE and an empty image. Interlaced PNGs
// are not supported (LOGE, empty).
template <typename T>
T Imread(const std::string& filename, int flags = IMREAD_COLOR);

}  // namespace ugu
