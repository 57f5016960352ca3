// PNG and raw image I/O for the C++ API (src/image_io.cc in UGU, which uses
// stb_image / stb_image_write / lodepng; this build has none of them and uses
// zlib for the deflate streams).
//
// Reading follows LoadByStb (src/image_io.cc:14-59): the bit depth byte of
// the header picks stbi_load (8) or stbi_load_16 (anything else), with no
// requested channel count: the image's own channels (gray, gray+alpha, RGB,
// RGBA; palette expanded to RGB, or RGBA with a tRNS chunk; a tRNS colour key
// adds an alpha channel). 1/2/4-bit gray is scaled to 8 bits and, as
// stbi_load_16 does for such files, widened to 16 bits by v * 257 (so are
// palette images of depth < 8). Writing stores filter type 0 rows in one zlib
// stream: the bytes differ from stb's and lodepng's encoders, the decoded
// pixels do not.

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "ugu/image_io.h"
#include "ugu/log.h"
#include "ugu/util/io_util.h"

namespace ugu {

namespace {

struct Decoded {
  int w = 0, h = 0, ch = 0, bytes = 0;  // bytes per sample: 1 or 2
  std::vector<uint8_t> px;              // row-major HWC, 16-bit in host order
};

uint32_t be32(const uint8_t* p) {
  return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3];
}

bool read_file(const std::string& path, std::vector<uint8_t>* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  try {
    out->assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  } catch (const std::exception&) {  // e.g. a directory: opens, then fails to read
    return false;
  }
  return true;
}

uint8_t paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
  if (pa <= pb && pa <= pc) return uint8_t(a);
  if (pb <= pc) return uint8_t(b);
  return uint8_t(c);
}

// out8: the file is decoded as by stbi_load (16-bit samples reduced to their
// high byte), else by the bit-depth rule above.
bool decode_png(const std::vector<uint8_t>& f, bool out8, Decoded* d, std::string* err) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) {
    *err = "not a PNG file";
    return false;
  }
  int bd = 0, ct = -1, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  size_t pos = 8;
  bool have_ihdr = false;
  while (pos + 12 <= f.size()) {
    const uint32_t len = be32(&f[pos]);
    if (len > f.size() - pos - 12) {
      *err = "truncated chunk";
      return false;
    }
    const uint8_t* type = &f[pos + 4];
    const uint8_t* data = &f[pos + 8];
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len < 13) {
        *err = "bad IHDR";
        return false;
      }
      d->w = int(be32(data));
      d->h = int(be32(data + 4));
      bd = data[8];
      ct = data[9];
      interlace = data[12];
      have_ihdr = true;
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(data, data + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(data, data + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), data, data + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    pos += 12 + len;
  }
  if (!have_ihdr || d->w < 1 || d->h < 1 || d->w > (1 << 24) || d->h > (1 << 24)) {
    *err = "missing or invalid IHDR";
    return false;
  }
  if (interlace != 0) {
    *err = "interlaced PNG is not supported";
    return false;
  }
  int n = 0;  // samples per pixel in the file
  switch (ct) {
    case 0: n = 1; break;
    case 2: n = 3; break;
    case 3: n = 1; break;
    case 4: n = 2; break;
    case 6: n = 4; break;
    default: *err = "unknown colour type"; return false;
  }
  const bool ok_bd = (ct == 0 && (bd == 1 || bd == 2 || bd == 4 || bd == 8 || bd == 16)) ||
                     (ct == 3 && (bd == 1 || bd == 2 || bd == 4 || bd == 8)) ||
                     ((ct == 2 || ct == 4 || ct == 6) && (bd == 8 || bd == 16));
  if (!ok_bd) {
    *err = "invalid bit depth";
    return false;
  }
  const size_t W = size_t(d->w), H = size_t(d->h);
  const size_t row = (W * size_t(n) * size_t(bd) + 7) / 8;
  const size_t bpp = std::max<size_t>(1, size_t(n) * size_t(bd) / 8);
  // Deflate expands at most 1032:1 (258-byte matches coded in two bits), so
  // a header whose rows cannot come out of the compressed data is rejected
  // before anything that large is allocated.
  if ((row + 1) * H > idat.size() * 1032 + 1024) {
    *err = "image data too short for the header's size";
    return false;
  }
  std::vector<uint8_t> raw((row + 1) * H);
  uLongf raw_len = uLongf(raw.size());
  if (uncompress(raw.data(), &raw_len, idat.data(), uLong(idat.size())) != Z_OK ||
      raw_len != raw.size()) {
    *err = "corrupt image data";
    return false;
  }
  // undo the row filters in place
  std::vector<uint8_t> cur(row), prev(row, 0);
  std::vector<uint8_t> img(row * H);
  for (size_t y = 0; y < H; y++) {
    const uint8_t ft = raw[y * (row + 1)];
    const uint8_t* src = &raw[y * (row + 1) + 1];
    for (size_t i = 0; i < row; i++) {
      const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
      int v = src[i];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, c); break;
        default: *err = "unknown row filter"; return false;
      }
      cur[i] = uint8_t(v);
    }
    std::memcpy(&img[y * row], cur.data(), row);
    std::swap(cur, prev);
  }
  // expand to the output layout
  const bool key = !trns.empty() && (ct == 0 || ct == 2);
  d->ch = ct == 3 ? (trns.empty() ? 3 : 4) : n + (key ? 1 : 0);
  d->bytes = (bd == 8 || out8) ? 1 : 2;
  const uint32_t widen = (bd < 8 && d->bytes == 2) ? 257u : 1u;  // 8-bit expansion -> 16 bits
  const int shift = (bd == 16 && d->bytes == 1) ? 8 : 0;         // 16 -> 8 bits
  d->px.assign(W * H * size_t(d->ch) * size_t(d->bytes), 0);
  for (size_t y = 0; y < H; y++) {
    const uint8_t* r = &img[y * row];
    for (size_t x = 0; x < W; x++) {
      uint32_t s[4] = {0, 0, 0, 0};
      if (bd < 8) {
        const size_t bit = x * size_t(bd);
        const uint32_t v = (r[bit >> 3] >> (8 - bd - (bit & 7))) & ((1u << bd) - 1);
        s[0] = ct == 3 ? v : v * (255u / ((1u << bd) - 1));
      } else {
        for (int k = 0; k < n; k++)
          s[k] = bd == 16 ? (uint32_t(r[(x * n + k) * 2]) << 8) | r[(x * n + k) * 2 + 1]
                          : r[x * n + k];
      }
      const size_t o = (y * W + x) * size_t(d->ch);
      if (ct == 3) {
        const uint32_t i = s[0];
        if (size_t(i) * 3 + 2 >= plte.size()) {
          *err = "palette index out of range";
          return false;
        }
        uint32_t c[4] = {plte[i * 3], plte[i * 3 + 1], plte[i * 3 + 2],
                         i < trns.size() ? trns[i] : 255u};
        for (int k = 0; k < d->ch; k++) {
          if (d->bytes == 2) {
            const uint16_t v16 = uint16_t(c[k] * widen);
            std::memcpy(&d->px[(o + size_t(k)) * 2], &v16, 2);
          } else {
            d->px[o + size_t(k)] = uint8_t(c[k]);
          }
        }
        continue;
      }
      const uint32_t full = bd == 16 ? 65535u : 255u;
      uint32_t alpha = full;
      if (key) {
        // colour key: 16-bit samples in tRNS, compared at the image's depth
        bool match = true;
        for (int k = 0; k < n; k++) {
          uint32_t kv = size_t(k) * 2 + 1 < trns.size()
                            ? (uint32_t(trns[k * 2]) << 8) | trns[k * 2 + 1]
                            : 0;
          if (bd < 8) kv *= 255u / ((1u << bd) - 1);
          match = match && (s[k] == kv);
        }
        alpha = match ? 0 : full;
      }
      for (int k = 0; k < d->ch; k++) {
        const uint32_t v = k < n ? s[k] : alpha;
        if (d->bytes == 2) {
          const uint16_t v16 = uint16_t(v * widen);
          std::memcpy(&d->px[(o + size_t(k)) * 2], &v16, 2);
        } else {
          d->px[o + size_t(k)] = uint8_t(v >> shift);
        }
      }
    }
  }
  return true;
}

void put32(std::vector<uint8_t>* o, uint32_t v) {
  o->push_back(uint8_t(v >> 24));
  o->push_back(uint8_t(v >> 16));
  o->push_back(uint8_t(v >> 8));
  o->push_back(uint8_t(v));
}

void chunk(std::vector<uint8_t>* o, const char* type, const std::vector<uint8_t>& data) {
  put32(o, uint32_t(data.size()));
  const size_t start = o->size();
  o->insert(o->end(), type, type + 4);
  o->insert(o->end(), data.begin(), data.end());
  const uLong crc = crc32(0L, o->data() + start, uInt(o->size() - start));
  put32(o, uint32_t(crc));
}

// ch: 1, 3 or 4 samples per pixel; bytes: 1 or 2 (16-bit samples in host
// order, written big-endian)
bool write_png(const std::string& path, int w, int h, int ch, int bytes, const uint8_t* data) {
  if (w < 1 || h < 1) {
    LOGE("image is empty\n");
    return false;
  }
  const size_t row = size_t(w) * size_t(ch) * size_t(bytes);
  std::vector<uint8_t> raw((row + 1) * size_t(h));
  for (size_t y = 0; y < size_t(h); y++) {
    uint8_t* dst = &raw[y * (row + 1)];
    dst[0] = 0;
    const uint8_t* src = data + y * row;
    if (bytes == 1) {
      std::memcpy(dst + 1, src, row);
    } else {
      for (size_t i = 0; i < row; i += 2) {
        uint16_t v;
        std::memcpy(&v, src + i, 2);
        dst[1 + i] = uint8_t(v >> 8);
        dst[2 + i] = uint8_t(v);
      }
    }
  }
  std::vector<uint8_t> z(compressBound(uLong(raw.size())));
  uLongf zlen = uLongf(z.size());
  if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK) {
    LOGE("deflate failed\n");
    return false;
  }
  z.resize(zlen);
  std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
  std::vector<uint8_t> ihdr;
  put32(&ihdr, uint32_t(w));
  put32(&ihdr, uint32_t(h));
  ihdr.push_back(uint8_t(bytes * 8));
  ihdr.push_back(uint8_t(ch == 1 ? 0 : (ch == 3 ? 2 : 6)));
  ihdr.push_back(0);
  ihdr.push_back(0);
  ihdr.push_back(0);
  chunk(&out, "IHDR", ihdr);
  chunk(&out, "IDAT", z);
  chunk(&out, "IEND", {});
  std::ofstream f(path, std::ios::binary);
  if (!f) {
    LOGE("can't open %s\n", path.c_str());
    return false;
  }
  f.write(reinterpret_cast<const char*>(out.data()), std::streamsize(out.size()));
  return bool(f);
}

std::string extension(const std::string& filename) {
  const size_t i = filename.find_last_of('.');
  return i == std::string::npos ? "" : filename.substr(i);
}

template <typename T>
bool imwrite_any(const std::string& filename, const Image<T>& img) {
  if (filename.size() < 4) return false;
  const std::string ext = extension(filename);
  const int ch = PixelTraits<T>::channels;
  const int bytes = int(sizeof(typename PixelTraits<T>::value_type));
  if (ext == ".png" || ext == ".PNG") {
    if (bytes == 2 && ch != 1) {
      LOGE("1 byte per channel is required to save by stb_image: actual %d\n", bytes);
      return false;
    }
    return write_png(filename, img.cols, img.rows, ch, bytes, img.data);
  }
  if (ext == ".bin" || ext == ".BIN") {
    std::ofstream f(filename, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char*>(img.data), std::streamsize(img.total() * sizeof(T)));
    return bool(f);
  }
  if (ext == ".jpg" || ext == ".jpeg" || ext == ".JPG" || ext == ".JPEG") {
    LOGE("can't write image with this configuration\n");
    return false;
  }
  LOGE("acceptable extention is .png, .jpg or .jpeg. this extention is not supported: %s\n",
       filename.c_str());
  return false;
}

}  // namespace

bool imwrite(const std::string& filename, const Image1b& img) { return imwrite_any(filename, img); }
bool imwrite(const std::string& filename, const Image3b& img) { return imwrite_any(filename, img); }
bool imwrite(const std::string& filename, const Image4b& img) { return imwrite_any(filename, img); }
bool imwrite(const std::string& filename, const Image1w& img) { return imwrite_any(filename, img); }

template <typename T>
T Imread(const std::string& filename, int flags) {
  std::vector<uint8_t> file;
  Decoded d;
  std::string err;
  const std::string ext = extension(filename);
  const bool png_name = ext == ".png" || ext == ".PNG";
  if (!read_file(filename, &file) || !decode_png(file, !png_name, &d, &err)) {
    LOGE("failed to load %s: %s\n", filename.c_str(), err.empty() ? "can't read" : err.c_str());
    return T();
  }
  if (flags == IMREAD_COLOR && d.ch == 4 && d.bytes == 1) {
    // imread's IMREAD_COLOR case: 4 channels -> 3 (src/image_io.cc:213-221)
    std::vector<uint8_t> rgb(size_t(d.w) * size_t(d.h) * 3);
    for (size_t i = 0; i < size_t(d.w) * size_t(d.h); i++)
      for (int k = 0; k < 3; k++) rgb[i * 3 + k] = d.px[i * 4 + k];
    d.px.swap(rgb);
    d.ch = 3;
  }
  const int want_ch = T().channels();
  const int want_bytes = int(sizeof(typename T::value_type));
  if (d.ch != want_ch || d.bytes != want_bytes) {
    LOGE("desired channel %d/%d, actual %d/%d\n", want_ch, want_bytes, d.ch, d.bytes);
    return T();
  }
  T img = T::zeros(d.h, d.w);
  std::memcpy(img.data, d.px.data(), d.px.size());
  return img;
}

template Image1b Imread<Image1b>(const std::string&, int);
template Image3b Imread<Image3b>(const std::string&, int);
template Image4b Imread<Image4b>(const std::string&, int);
template Image1w Imread<Image1w>(const std::string&, int);

void WriteFaceIdAsText(const Image1i& face_id, const std::string& path) {
  std::ofstream ofs(path, std::ios::out);
  for (int y = 0; y < face_id.rows; y++)
    for (int x = 0; x < face_id.cols; x++) ofs << face_id.at<int>(y, x) << "\n";
  ofs.flush();
}

}  // namespace ugu
