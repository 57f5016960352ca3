// Host-only driver of the visibility statistics (VertexInfo::CalcStat,
// FaceInfo::CalcStat) and of TextureMapping with kGenerateSimpleTile and
// kGenerateSimpleCharts (the uv types that need no GPU), run by
// tests/test_texturing_host.py against hand-computed values. Nothing here
// touches a GPU.
//
// usage:
//   ugu_stats_tool vertex <records.txt>
//       records: one keyframe per line "kf_id r g b intensity angle distance";
//       prints every statistic as "name value..." (floats in %a)
//   ugu_stats_tool face <records.txt>
//       records: "kf_id area angle distance" per line; prints the indices/ids
//   ugu_stats_tool tile <dir>
//       dir/meta.txt: "n_kf w h n_faces criteria"; dir/ids.txt: keyframe ids;
//       dir/kf<i>.raw: h*w*3 u8; dir/faces.txt: per face "m" then m lines
//       "kf_id area angle distance x0 y0 x1 y1 x2 y2";
//       writes dir/uv.f32, dir/uvi.i32, dir/tex.u8 and prints "rc w h"
//   ugu_stats_tool map <dir>
//       as tile, with dir/meta.txt "n_kf w h n_faces criteria uv_type tex_w
//       tex_h" and the mesh's faces in dir/vi.txt (three vertex ids per face)

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ugu/texturing/texture_mapper.h"
#include "ugu/texturing/visibility_tester.h"

namespace {

// Whitespace-separated numbers of one line, parsed with strtof (which reads
// "inf" and "nan", unlike operator>>).
std::vector<float> numbers(const std::string& line) {
  std::vector<float> v;
  std::istringstream ss(line);
  std::string tok;
  while (ss >> tok) v.push_back(std::strtof(tok.c_str(), nullptr));
  return v;
}

void put(const char* name, float v) { std::printf("%s %a\n", name, static_cast<double>(v)); }
void put(const char* name, int v) { std::printf("%s %d\n", name, v); }
void put(const char* name, const Eigen::Vector3f& v) {
  std::printf("%s %a %a %a\n", name, static_cast<double>(v[0]), static_cast<double>(v[1]),
              static_cast<double>(v[2]));
}

int vertex(const char* path) {
  std::ifstream in(path);
  ugu::VertexInfo vi;
  for (std::string line; std::getline(in, line);) {
    const std::vector<float> x = numbers(line);
    if (x.size() != 7) continue;
    ugu::VertexInfoPerKeyframe k;
    k.kf_id = static_cast<int>(x[0]);
    k.color = Eigen::Vector3f(x[1], x[2], x[3]);
    k.intensity = x[4];
    k.viewing_angle = x[5];
    k.distance = x[6];
    vi.Update(k);
  }
  int custom_calls = 0;
  vi.CalcStat([&](ugu::VertexInfo& v) {
    custom_calls++;
    v.custom_best = v.mode;
  });
  put("mean_color", vi.mean_color);
  put("median_color", vi.median_color);
  put("min_viewing_angle_index", vi.min_viewing_angle_index);
  put("min_viewing_angle_id", vi.min_viewing_angle_id);
  put("min_viewing_angle_color", vi.min_viewing_angle_color);
  put("mean_viewing_angle_color", vi.mean_viewing_angle_color);
  put("median_viewing_angle_color", vi.median_viewing_angle_color);
  put("min_distance_index", vi.min_distance_index);
  put("min_distance_id", vi.min_distance_id);
  put("min_distance_color", vi.min_distance_color);
  put("mean_distance_color", vi.mean_distance_color);
  put("median_distance_color", vi.median_distance_color);
  put("min_intensity", vi.min_intensity);
  put("median_intensity", vi.median_intensity);
  put("median_intensity_color", vi.median_intensity_color);
  put("min_intensity_color", vi.min_intensity_color);
  put("mean_viewing_angle", vi.mean_viewing_angle);
  put("median_viewing_angle", vi.median_viewing_angle);
  put("mean_distance", vi.mean_distance);
  put("median_distance", vi.median_distance);
  put("mode", vi.mode);
  put("mode_frequency", vi.mode_frequency);
  put("mode_viewing_angle_color", vi.mode_viewing_angle_color);
  put("mode_frequency_viewing_angle_color", vi.mode_frequency_viewing_angle_color);
  put("occurrence_size", static_cast<int>(vi.occurrence.size()));
  put("custom_calls", custom_calls);
  put("custom_best", vi.custom_best);
  return 0;
}

int face(const char* path) {
  std::ifstream in(path);
  ugu::FaceInfo fi;
  for (std::string line; std::getline(in, line);) {
    const std::vector<float> x = numbers(line);
    if (x.size() != 4) continue;
    ugu::FaceInfoPerKeyframe k;
    k.kf_id = static_cast<int>(x[0]);
    k.area = x[1];
    k.viewing_angle = x[2];
    k.distance = x[3];
    fi.Update(k);
  }
  fi.CalcStat();
  put("max_area_index", fi.max_area_index);
  put("max_area_id", fi.max_area_id);
  put("max_area", fi.max_area);
  put("min_viewing_angle_index", fi.min_viewing_angle_index);
  put("min_viewing_angle_id", fi.min_viewing_angle_id);
  put("min_viewing_angle", fi.min_viewing_angle);
  put("min_distance_index", fi.min_distance_index);
  put("min_distance_id", fi.min_distance_id);
  put("min_distance", fi.min_distance);
  return 0;
}

template <typename T>
void dump(const std::string& path, const T* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), std::streamsize(n * sizeof(T)));
}

int tile(const std::string& dir, bool with_mesh) {
  std::ifstream meta(dir + "/meta.txt");
  int n = 0, w = 0, h = 0, F = 0, criteria = 0;
  meta >> n >> w >> h >> F >> criteria;
  ugu::TextureMappingOption opt;  // default uv type: kGenerateSimpleTile
  opt.criteria = static_cast<ugu::ViewSelectionCriteria>(criteria);
  if (with_mesh) {
    int uv_type = 0;
    meta >> uv_type >> opt.tex_w >> opt.tex_h;
    opt.uv_type = static_cast<ugu::TexturingOutputUvType>(uv_type);
  }
  std::ifstream ids(dir + "/ids.txt");
  std::vector<std::shared_ptr<ugu::Keyframe>> kfs;
  for (int i = 0; i < n; i++) {
    auto kf = ugu::Keyframe::Create();
    ids >> kf->id;
    kf->color = ugu::Image3b::zeros(h, w);
    std::ifstream raw(dir + "/kf" + std::to_string(i) + ".raw", std::ios::binary);
    raw.read(reinterpret_cast<char*>(kf->color.data), std::streamsize(size_t(w) * h * 3));
    kfs.push_back(kf);
  }
  ugu::Mesh mesh;
  std::vector<Eigen::Vector3i> faces(size_t(F), Eigen::Vector3i(0, 1, 2));
  int n_vertices = 3;
  if (with_mesh) {
    std::ifstream vi(dir + "/vi.txt");
    for (auto& f : faces) {
      vi >> f[0] >> f[1] >> f[2];
      n_vertices = std::max({n_vertices, f[0] + 1, f[1] + 1, f[2] + 1});
    }
  }
  mesh.set_vertices(std::vector<Eigen::Vector3f>(size_t(n_vertices), Eigen::Vector3f(0, 0, 0)));
  mesh.set_vertex_indices(faces);
  ugu::VisibilityInfo info(mesh);
  std::ifstream fin(dir + "/faces.txt");
  for (int f = 0; f < F; f++) {
    int m = 0;
    fin >> m;
    for (int r = 0; r < m; r++) {
      ugu::FaceInfoPerKeyframe k;
      k.face_id = f;
      fin >> k.kf_id >> k.area >> k.viewing_angle >> k.distance;
      for (int j = 0; j < 3; j++) fin >> k.projected_tri[size_t(j)][0] >> k.projected_tri[size_t(j)][1];
      info.Update(f, k);
    }
  }
  info.CalcStatFaceInfo();
  const bool ok = ugu::TextureMapping(kfs, info, &mesh, opt);
  int tw = 0, th = 0;
  if (ok) {
    const ugu::Image3b& tex = mesh.materials()[0].diffuse_tex;
    tw = tex.cols;
    th = tex.rows;
    dump(dir + "/uv.f32", mesh.uv().data()->data(), mesh.uv().size() * 2);
    dump(dir + "/uvi.i32", mesh.uv_indices().data()->data(), mesh.uv_indices().size() * 3);
    dump(dir + "/tex.u8", tex.data, size_t(tw) * size_t(th) * 3);
  }
  std::printf("%d %d %d\n", ok ? 1 : 0, tw, th);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  const std::string cmd = argv[1];
  if (cmd == "vertex") return vertex(argv[2]);
  if (cmd == "face") return face(argv[2]);
  if (cmd == "tile") return tile(argv[2], false);
  if (cmd == "map") return tile(argv[2], true);
  return 2;
}
