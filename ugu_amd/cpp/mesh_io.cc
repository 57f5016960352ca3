// Mesh::LoadObj (src/mesh.cc:500-735) with its own OBJ/MTL reader in place
// of tinyobjloader (syoyo/tinyobjloader, an empty submodule in the reference
// checkout), Mesh::RemoveVertices (src/mesh.cc:1066-1164, 2124-2186),
// set_material_ids and Clear.
//
// The reader keeps the parts of tinyobjloader's LoadObj that the reference
// uses: v (optional r g b, kept only when every vertex has them), vt, vn,
// f with v, v/vt, v//vn, v/vt/vn and negative (relative) indices, usemtl,
// mtllib; faces of 4 vertices are split on the shorter diagonal ([0 1 2]
// [0 2 3] when |v2 - v0|^2 < |v3 - v1|^2, else [0 1 3] [1 2 3]) and larger
// polygons as a fan; MTL newmtl, Ka, Kd, Ks, Ns, d, Tr, illum, map_Kd (the
// last token of the line). Numbers are parsed with strtod and rounded to
// float.

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "ugu/image_io.h"
#include "ugu/log.h"
#include "ugu/mesh.h"

namespace ugu {

namespace {

struct ObjIndex {
  int v = -1, vt = -1, vn = -1;
};

// tinyobjloader's fixIndex: 1-based, negative counts back from the end
bool fix_index(long idx, size_t n, int* out) {
  if (idx > 0) {
    *out = int(idx - 1);
    return true;
  }
  if (idx < 0) {
    *out = int(long(n) + idx);
    return true;
  }
  return false;  // 0 is not a valid index
}

bool parse_vertex_ref(const std::string& tok, size_t nv, size_t nt, size_t nn, ObjIndex* o) {
  const size_t s1 = tok.find('/');
  char* end = nullptr;
  const long v = std::strtol(tok.c_str(), &end, 10);
  if (end == tok.c_str() || !fix_index(v, nv, &o->v)) return false;
  if (s1 == std::string::npos) return true;
  const size_t s2 = tok.find('/', s1 + 1);
  const std::string t = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
  if (!t.empty() && !fix_index(std::strtol(t.c_str(), nullptr, 10), nt, &o->vt)) return false;
  if (s2 != std::string::npos) {
    const std::string n = tok.substr(s2 + 1);
    if (!n.empty() && !fix_index(std::strtol(n.c_str(), nullptr, 10), nn, &o->vn)) return false;
  }
  return true;
}

std::vector<std::string> split_ws(const std::string& line) {
  std::vector<std::string> t;
  std::istringstream is(line);
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

float to_float(const std::string& s) { return static_cast<float>(std::strtod(s.c_str(), nullptr)); }

std::string trim_cr(std::string s) {
  while (!s.empty() && (s.back() == '\r' || s.back() == '\n')) s.pop_back();
  return s;
}

bool load_mtl(const std::string& path, std::vector<ObjMaterial>* mats,
              std::map<std::string, int>* ids) {
  std::ifstream f(path);
  if (!f) return false;
  std::string line;
  ObjMaterial cur;
  bool open = false;
  auto flush = [&]() {
    if (open) {
      (*ids)[cur.name] = int(mats->size());
      mats->push_back(cur);
    }
  };
  while (std::getline(f, line)) {
    const std::vector<std::string> t = split_ws(trim_cr(line));
    if (t.empty() || t[0][0] == '#') continue;
    const std::string& k = t[0];
    if (k == "newmtl") {
      flush();
      cur = ObjMaterial();
      // tinyobjloader's material defaults
      cur.ambient = {0.0f, 0.0f, 0.0f};
      cur.diffuse = {0.0f, 0.0f, 0.0f};
      cur.specular = {0.0f, 0.0f, 0.0f};
      cur.shininess = 1.0f;
      cur.dissolve = 1.0f;
      cur.illum = 0;
      cur.name = t.size() > 1 ? t[1] : "";
      open = true;
    } else if ((k == "Ka" || k == "Kd" || k == "Ks") && t.size() >= 4) {
      std::array<float, 3>& dst = k == "Ka" ? cur.ambient : (k == "Kd" ? cur.diffuse : cur.specular);
      for (int i = 0; i < 3; i++) dst[size_t(i)] = to_float(t[size_t(i) + 1]);
    } else if (k == "Ns" && t.size() >= 2) {
      cur.shininess = to_float(t[1]);
    } else if (k == "d" && t.size() >= 2) {
      cur.dissolve = to_float(t[1]);
    } else if (k == "Tr" && t.size() >= 2) {
      cur.dissolve = 1.0f - to_float(t[1]);
    } else if (k == "illum" && t.size() >= 2) {
      cur.illum = std::atoi(t[1].c_str());
    } else if (k == "map_Kd" && t.size() >= 2) {
      cur.diffuse_texname = t.back();
    }
  }
  flush();
  return true;
}

std::string extract_dir(const std::string& path) {
  const size_t pos = path.find_last_of("/\\");
  return pos == std::string::npos ? "./" : path.substr(0, pos + 1);
}

}  // namespace

void Mesh::Clear() {
  vertices_.clear();
  vertex_colors_.clear();
  vertex_indices_.clear();
  normals_.clear();
  face_normals_.clear();
  normal_indices_.clear();
  uv_.clear();
  uv_indices_.clear();
  materials_.clear();
  material_ids_.clear();
  face_indices_per_material_.clear();
}

bool Mesh::set_material_ids(const std::vector<int>& material_ids) {
  if (material_ids.empty()) {
    LOGE("material id is empty\n");
    return false;
  }
  int max_id = material_ids[0];
  for (int id : material_ids) max_id = std::max(max_id, id);
  if (max_id < 0) {
    LOGE("material id must be positive\n");
    return false;
  }
  material_ids_ = material_ids;
  face_indices_per_material_.assign(size_t(max_id) + 1, {});
  for (int i = 0; i < static_cast<int>(material_ids_.size()); i++)
    if (material_ids_[size_t(i)] >= 0)  // a negative id is read out of bounds by the reference
      face_indices_per_material_[size_t(material_ids_[size_t(i)])].push_back(i);
  return true;
}

int Mesh::RemoveVertices(const std::vector<bool>& valid_vertex_table) {
  if (valid_vertex_table.size() != vertices_.size()) {
    LOGE("valid_vertex_table must be same size to vertices");
    return -1;
  }
  const bool with_uv = !uv_.empty() && !uv_indices_.empty();
  const bool with_vertex_color = !vertex_colors_.empty();
  int num_removed = 0, valid_count = 0;
  std::vector<int> valid_table(vertices_.size(), -1);
  std::vector<Eigen::Vector3f> vv, vc;
  std::vector<Eigen::Vector2f> vuv;
  for (size_t i = 0; i < vertices_.size(); i++) {
    if (valid_vertex_table[i]) {
      valid_table[i] = valid_count++;
      vv.push_back(vertices_[i]);
      // uv is taken per vertex index (RemoveVerticesBase); with fewer uvs
      // than vertices the reference reads out of bounds: 0 here
      if (with_uv) vuv.push_back(i < uv_.size() ? uv_[i] : Eigen::Vector2f(0.0f, 0.0f));
      if (with_vertex_color)
        vc.push_back(i < vertex_colors_.size() ? vertex_colors_[i]
                                               : Eigen::Vector3f(0.0f, 0.0f, 0.0f));
    } else {
      num_removed++;
    }
  }
  std::vector<Eigen::Vector3i> faces;
  std::vector<int> face_table(vertex_indices_.size(), -1);
  for (size_t i = 0; i < vertex_indices_.size(); i++) {
    Eigen::Vector3i f;
    bool ok = true;
    for (int j = 0; j < 3 && ok; j++) {
      const int n = valid_table[size_t(vertex_indices_[i][j])];
      ok = n >= 0;
      f[j] = n;
    }
    if (!ok) continue;
    face_table[i] = static_cast<int>(faces.size());
    faces.push_back(f);
  }
  vertices_ = vv;
  vertex_indices_ = faces;
  if (with_uv) {
    uv_ = vuv;
    uv_indices_ = faces;
  }
  if (with_vertex_color) vertex_colors_ = vc;
  CalcNormal();
  std::vector<int> ids(faces.size(), 0);
  for (size_t i = 0; i < material_ids_.size() && i < face_table.size(); i++)
    if (face_table[i] >= 0) ids[size_t(face_table[i])] = material_ids_[i];
  if (!ids.empty()) set_material_ids(ids);
  return num_removed;
}

bool Mesh::LoadObj(const std::string& obj_path, const std::string& mtl_dir) {
  Clear();
  std::ifstream f(obj_path);
  if (!f) {
    LOGE("Cannot open file [%s]\n", obj_path.c_str());
    return false;
  }
  const std::string dir = mtl_dir.empty() ? extract_dir(obj_path) : mtl_dir;

  std::vector<float> v, vn, vt, vcol;
  bool all_colors = true;
  std::vector<ObjIndex> tri;  // 3 per face
  std::vector<int> tri_mat;
  std::vector<ObjMaterial> mats;
  std::map<std::string, int> mat_ids;
  int cur_mat = -1;
  std::string line;
  int line_no = 0;
  while (std::getline(f, line)) {
    line_no++;
    const std::vector<std::string> t = split_ws(trim_cr(line));
    if (t.empty() || t[0][0] == '#') continue;
    const std::string& k = t[0];
    if (k == "v") {
      float x[6] = {0, 0, 0, 1, 1, 1};
      for (size_t i = 1; i < t.size() && i <= 6; i++) x[i - 1] = to_float(t[i]);
      v.insert(v.end(), x, x + 3);
      all_colors = all_colors && t.size() >= 7;
      vcol.insert(vcol.end(), x + 3, x + 6);
    } else if (k == "vn") {
      for (size_t i = 1; i <= 3; i++) vn.push_back(i < t.size() ? to_float(t[i]) : 0.0f);
    } else if (k == "vt") {
      for (size_t i = 1; i <= 2; i++) vt.push_back(i < t.size() ? to_float(t[i]) : 0.0f);
    } else if (k == "f") {
      std::vector<ObjIndex> poly;
      for (size_t i = 1; i < t.size(); i++) {
        ObjIndex o;
        if (!parse_vertex_ref(t[i], v.size() / 3, vt.size() / 2, vn.size() / 3, &o)) {
          LOGE("Failed parse `f' line (e.g. zero value for face index. line %d.)\n", line_no);
          return false;
        }
        poly.push_back(o);
      }
      if (poly.size() < 3) continue;
      auto emit = [&](int a, int b, int c) {
        tri.push_back(poly[size_t(a)]);
        tri.push_back(poly[size_t(b)]);
        tri.push_back(poly[size_t(c)]);
        tri_mat.push_back(cur_mat);
      };
      if (poly.size() == 3) {
        emit(0, 1, 2);
      } else if (poly.size() == 4) {
        auto P = [&](int i, int c) {
          const size_t vi = size_t(poly[size_t(i)].v) * 3 + size_t(c);
          return vi < v.size() ? v[vi] : 0.0f;
        };
        float s02 = 0.0f, s13 = 0.0f;
        for (int c = 0; c < 3; c++) {
          const float e02 = P(2, c) - P(0, c), e13 = P(3, c) - P(1, c);
          s02 += e02 * e02;
          s13 += e13 * e13;
        }
        if (s02 < s13) {
          emit(0, 1, 2);
          emit(0, 2, 3);
        } else {
          emit(0, 1, 3);
          emit(1, 2, 3);
        }
      } else {
        for (size_t i = 1; i + 1 < poly.size(); i++) emit(0, int(i), int(i + 1));
      }
    } else if (k == "usemtl") {
      const std::string name = t.size() > 1 ? t[1] : "";
      const auto it = mat_ids.find(name);
      cur_mat = it == mat_ids.end() ? -1 : it->second;
    } else if (k == "mtllib") {
      bool found = false;
      for (size_t i = 1; i < t.size() && !found; i++) found = load_mtl(dir + t[i], &mats, &mat_ids);
      if (!found) LOGW("Material file [ %s ] not found.\n", t.size() > 1 ? t[1].c_str() : "");
    }
  }

  const size_t nv = v.size() / 3, nn = vn.size() / 3, nt = vt.size() / 2;
  const size_t face_num = tri.size() / 3;
  for (const ObjIndex& o : tri) {
    const bool bad = size_t(o.v) >= nv || o.v < 0 || (nn > 0 && (o.vn < 0 || size_t(o.vn) >= nn)) ||
                     (nt > 0 && (o.vt < 0 || size_t(o.vt) >= nt));
    if (bad) {
      // the reference indexes the attribute arrays with these values
      LOGE("face index out of range\n");
      return false;
    }
  }
  vertex_indices_.assign(face_num, Eigen::Vector3i(0, 0, 0));
  uv_indices_.assign(face_num, Eigen::Vector3i(0, 0, 0));
  normal_indices_.assign(face_num, Eigen::Vector3i(0, 0, 0));
  material_ids_.assign(face_num, 0);
  vertices_.assign(nv, Eigen::Vector3f(0, 0, 0));
  normals_.assign(nn, Eigen::Vector3f(0, 0, 0));
  uv_.assign(nt, Eigen::Vector2f(0, 0));
  if (all_colors && nv > 0) vertex_colors_.assign(nv, Eigen::Vector3f(0, 0, 0));
  std::vector<bool> valid(nv, false);
  const bool no_face = face_num == 0;
  if (no_face) {
    for (size_t i = 0; i < nv; i++) {
      vertices_[i] = Eigen::Vector3f(v[i * 3], v[i * 3 + 1], v[i * 3 + 2]);
      if (i < nn) {
        normals_[i] = Eigen::Vector3f(vn[i * 3], vn[i * 3 + 1], vn[i * 3 + 2]);
        normals_[i].normalize();
      }
      if (i < nt) uv_[i] = Eigen::Vector2f(vt[i * 2], vt[i * 2 + 1]);
      if (!vertex_colors_.empty())
        vertex_colors_[i] = Eigen::Vector3f(vcol[i * 3] * 255.f, vcol[i * 3 + 1] * 255.f,
                                            vcol[i * 3 + 2] * 255.f);
    }
  } else {
    for (size_t fi = 0; fi < face_num; fi++) {
      material_ids_[fi] = tri_mat[fi];
      for (int c = 0; c < 3; c++) {
        const ObjIndex& o = tri[fi * 3 + size_t(c)];
        const size_t vi = size_t(o.v);
        vertex_indices_[fi][c] = o.v;
        vertices_[vi] = Eigen::Vector3f(v[vi * 3], v[vi * 3 + 1], v[vi * 3 + 2]);
        valid[vi] = true;
        if (nn > 0) {
          const size_t ni = size_t(o.vn);
          normal_indices_[fi][c] = o.vn;
          normals_[ni] = Eigen::Vector3f(vn[ni * 3], vn[ni * 3 + 1], vn[ni * 3 + 2]);
          normals_[ni].normalize();
        }
        if (nt > 0) {
          const size_t ti = size_t(o.vt);
          uv_indices_[fi][c] = o.vt;
          uv_[ti] = Eigen::Vector2f(vt[ti * 2], vt[ti * 2 + 1]);
        }
        if (!vertex_colors_.empty())
          vertex_colors_[vi] = Eigen::Vector3f(vcol[vi * 3] * 255.f, vcol[vi * 3 + 1] * 255.f,
                                               vcol[vi * 3 + 2] * 255.f);
      }
    }
  }
  if (nn == 0) normal_indices_.clear();
  if (nt == 0) uv_indices_.clear();

  CalcFaceNormal();
  if (normals_.empty()) CalcNormal();

  bool ret = true;
  if (mats.empty()) {
    materials_.assign(1, ObjMaterial());
    material_ids_.assign(face_num, 0);
    LOGW("Default material was added because material did not find on input obj\n");
  } else {
    materials_ = mats;
    for (ObjMaterial& m : materials_) {
      m.diffuse_texpath = dir + m.diffuse_texname;
      // as the reference: an empty texture name leaves the directory itself,
      // which opens and then fails to decode (LoadObj returns false)
      std::ifstream ifs(m.diffuse_texpath);
      if (ifs.is_open()) {
        m.diffuse_tex = Imread<Image3b>(m.diffuse_texpath);
        ret = !m.diffuse_tex.empty();
      } else {
        LOGW("diffuse texture doesn't exist %s\n", m.diffuse_texpath.c_str());
      }
    }
  }
  face_indices_per_material_.assign(materials_.size(), {});
  for (int i = 0; i < static_cast<int>(material_ids_.size()); i++) {
    const int id = material_ids_[size_t(i)];
    face_indices_per_material_[size_t(id >= 0 ? id : 0)].push_back(i);
  }
  if (!no_face) {
    size_t cnt = 0;
    for (bool b : valid) cnt += b ? 1 : 0;
    if (cnt != valid.size()) RemoveVertices(valid);
  }
  return ret;
}

}  // namespace ugu
