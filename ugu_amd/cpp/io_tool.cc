// Host-only driver of the C++ I/O functions (Mesh::LoadObj, Imread / imwrite,
// WriteFaceIdAsText), run by tests/test_io.py to compare them with the Python
// implementation and with independent decoders. No GPU is involved.
//
// usage:
//   ugu_io_tool obj <obj_path> <out_dir>     LoadObj -> raw arrays + meta.txt
//   ugu_io_tool pngread <png> <out.raw> <kind>   kind: 1b 3b 4b 1w; prints "w h"
//                                                (4b reads with IMREAD_UNCHANGED)
//   ugu_io_tool pngwrite <in.raw> <w> <h> <kind> <out.png>
//   ugu_io_tool faceid <in.i32> <w> <h> <out.txt>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "ugu/image_io.h"
#include "ugu/log.h"
#include "ugu/mesh.h"
#include "ugu/util/io_util.h"

namespace {

template <typename T>
void dump(const std::string& path, const std::vector<T>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

std::vector<char> slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

template <typename I>
int png_read(const std::string& in, const std::string& out, int flags = ugu::IMREAD_COLOR) {
  I img = ugu::Imread<I>(in, flags);
  if (img.empty()) return 1;
  std::ofstream f(out, std::ios::binary);
  f.write(reinterpret_cast<const char*>(img.data), std::streamsize(img.total() * img.elemSize()));
  std::printf("%d %d\n", img.cols, img.rows);
  return 0;
}

template <typename I>
int png_write(const std::string& in, int w, int h, const std::string& out) {
  I img = I::zeros(h, w);
  const std::vector<char> d = slurp(in);
  if (d.size() != img.total() * img.elemSize()) return 2;
  std::copy(d.begin(), d.end(), reinterpret_cast<char*>(img.data));
  return ugu::imwrite(out, img) ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string cmd = argv[1];
  if (cmd == "obj" && argc >= 4) {
    ugu::Mesh m;
    const bool ret = m.LoadObj(argv[2]);
    const std::string d = argv[3];
    std::ofstream meta(d + "/meta.txt");
    meta << int(ret) << " " << m.vertices().size() << " " << m.vertex_indices().size() << " "
         << m.normals().size() << " " << m.normal_indices().size() << " " << m.uv().size() << " "
         << m.uv_indices().size() << " " << m.vertex_colors().size() << " "
         << m.materials().size() << "\n";
    for (const auto& mt : m.materials())
      meta << mt.name << " " << mt.diffuse_texname.size() << " " << mt.diffuse_tex.cols << " "
           << mt.diffuse_tex.rows << "\n";
    auto f3 = [](const std::vector<Eigen::Vector3f>& a) {
      std::vector<float> o;
      for (const auto& x : a) o.insert(o.end(), {x[0], x[1], x[2]});
      return o;
    };
    auto i3 = [](const std::vector<Eigen::Vector3i>& a) {
      std::vector<int32_t> o;
      for (const auto& x : a) o.insert(o.end(), {x[0], x[1], x[2]});
      return o;
    };
    std::vector<float> uv;
    for (const auto& x : m.uv()) uv.insert(uv.end(), {x[0], x[1]});
    dump(d + "/verts.f32", f3(m.vertices()));
    dump(d + "/faces.i32", i3(m.vertex_indices()));
    dump(d + "/normals.f32", f3(m.normals()));
    dump(d + "/nidx.i32", i3(m.normal_indices()));
    dump(d + "/fnormals.f32", f3(m.face_normals()));
    dump(d + "/uv.f32", uv);
    dump(d + "/uvidx.i32", i3(m.uv_indices()));
    dump(d + "/colors.f32", f3(m.vertex_colors()));
    dump(d + "/matids.i32", m.material_ids());
    for (size_t i = 0; i < m.materials().size(); i++) {
      const ugu::Image3b& t = m.materials()[i].diffuse_tex;
      std::ofstream tf(d + "/tex" + std::to_string(i) + ".u8", std::ios::binary);
      tf.write(reinterpret_cast<const char*>(t.data), std::streamsize(t.total() * 3));
    }
    return 0;
  }
  if (cmd == "pngread" && argc >= 5) {
    const std::string k = argv[4];
    if (k == "1b") return png_read<ugu::Image1b>(argv[2], argv[3]);
    if (k == "3b") return png_read<ugu::Image3b>(argv[2], argv[3]);
    if (k == "4b") return png_read<ugu::Image4b>(argv[2], argv[3], ugu::IMREAD_UNCHANGED);
    if (k == "1w") return png_read<ugu::Image1w>(argv[2], argv[3]);
    return 2;
  }
  if (cmd == "pngwrite" && argc >= 7) {
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
    const std::string k = argv[5];
    if (k == "1b") return png_write<ugu::Image1b>(argv[2], w, h, argv[6]);
    if (k == "3b") return png_write<ugu::Image3b>(argv[2], w, h, argv[6]);
    if (k == "4b") return png_write<ugu::Image4b>(argv[2], w, h, argv[6]);
    if (k == "1w") return png_write<ugu::Image1w>(argv[2], w, h, argv[6]);
    return 2;
  }
  if (cmd == "faceid" && argc >= 6) {
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
    ugu::Image1i img = ugu::Image1i::zeros(h, w);
    const std::vector<char> d = slurp(argv[2]);
    if (d.size() != img.total() * 4) return 2;
    std::copy(d.begin(), d.end(), reinterpret_cast<char*>(img.data));
    ugu::WriteFaceIdAsText(img, argv[5]);
    return 0;
  }
  return 2;
}
