// Subset of ugu/image.h: dense interleaved HWC images, row-major and
// contiguous (include/ugu/image.h:347-427,934-985 in UGU).
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

namespace ugu {

template <typename TT, int N>
struct Vec_ {
  using value_type = TT;
  static constexpr int channels = N;
  TT val[N];
  TT& operator[](int i) { return val[i]; }
  const TT& operator[](int i) const { return val[i]; }
};

using Vec3b = Vec_<uint8_t, 3>;
using Vec4b = Vec_<uint8_t, 4>;
using Vec3f = Vec_<float, 3>;

template <typename T>
struct PixelTraits {
  using value_type = T;
  static constexpr int channels = 1;
};
template <typename TT, int N>
struct PixelTraits<Vec_<TT, N>> {
  using value_type = TT;
  static constexpr int channels = N;
};

template <typename T>
class Image {
 public:
  using value_type = typename PixelTraits<T>::value_type;
  int rows{0};
  int cols{0};
  uint8_t* data{nullptr};

  Image() {}
  Image(int rows_, int cols_) { Allocate(rows_, cols_); }
  Image(const Image& o) = default;  // shares the buffer, like ugu::Image
  Image& operator=(const Image& o) = default;

  static Image zeros(int rows_, int cols_) { return Image(rows_, cols_); }

  int channels() const { return PixelTraits<T>::channels; }
  bool empty() const { return rows < 1 || cols < 1; }
  size_t total() const { return size_t(rows) * size_t(cols); }
  size_t elemSize() const { return sizeof(T); }

  template <typename TT>
  TT& at(int y, int x) {
    return reinterpret_cast<TT*>(data + size_t(cols) * sizeof(T) * size_t(y))[x];
  }
  template <typename TT>
  const TT& at(int y, int x) const {
    return reinterpret_cast<const TT*>(data + size_t(cols) * sizeof(T) * size_t(y))[x];
  }

  void setTo(value_type v) {
    value_type* p = reinterpret_cast<value_type*>(data);
    const size_t n = total() * size_t(channels());
    for (size_t i = 0; i < n; i++) p[i] = v;
  }

 private:
  std::shared_ptr<std::vector<uint8_t>> buf_;
  void Allocate(int rows_, int cols_) {
    if (rows_ < 0 || cols_ < 0) throw std::runtime_error("Type error");
    rows = rows_;
    cols = cols_;
    buf_ = std::make_shared<std::vector<uint8_t>>(size_t(rows) * size_t(cols) * sizeof(T), 0);
    data = buf_->data();
  }
};

using Image1b = Image<uint8_t>;   // gray / mask
using Image1w = Image<uint16_t>;  // 16-bit depth
using Image1i = Image<int32_t>;   // face id
using Image1f = Image<float>;     // depth
using Image3f = Image<Vec3f>;     // normal, XYZ
using Image3b = Image<Vec3b>;     // colour, RGB
using Image4b = Image<Vec4b>;     // colour with alpha, RGBA

// ugu::Init (image.h:964-985): reallocate when the size differs, then fill.
template <typename T>
inline void Init(Image<T>* image, int width, int height,
                 typename Image<T>::value_type val) {
  if (image->cols != width || image->rows != height)
    *image = Image<T>::zeros(height, width);
  image->setTo(val);
}

// ConvertTo (image_proc.h:453-520) for the float depth -> uint16 case of
// RenderW: static_cast<uint16_t>(scale * (double)v).
bool ConvertTo(const Image1f& src, Image1w* dst, float scale = 1.0f);

}  // namespace ugu
