// Subset of ugu/mesh.h (include/ugu/mesh.h:81-199 in UGU): the arrays the
// renderer reads, their setters, CalcFaceNormal / CalcNormal, LoadObj and
// RemoveVertices.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "ugu/eigen_min.h"
#include "ugu/image.h"

namespace ugu {

// ObjMaterial (include/ugu/mesh.h:26-50 in UGU), the fields LoadObj fills
struct ObjMaterial {
  std::string name = "default_mat";
  std::array<float, 3> ambient{0.117647f, 0.117647f, 0.117647f};   // Ka
  std::array<float, 3> diffuse{0.752941f, 0.752941f, 0.752941f};   // Kd
  std::array<float, 3> specular{0.752941f, 0.752941f, 0.752941f};  // Ks
  float shininess{8.0f};                                           // Ns
  float dissolve{1.0f};                                            // d (1 - Tr)
  int illum{1};
  std::string diffuse_texname;
  std::string diffuse_texpath;
  Image3b diffuse_tex;
};

class Mesh;
using MeshPtr = std::shared_ptr<Mesh>;

class Mesh {
 public:
  Mesh() {}
  static MeshPtr Create() { return std::make_shared<Mesh>(); }

  void CalcNormal();
  void CalcFaceNormal();
  void Clear();

  // Mesh::LoadObj (src/mesh.cc:500-735): triangles (quads split on the
  // shorter diagonal, larger polygons as a fan), per-face materials from the
  // MTL file, diffuse textures read with Imread<Image3b>; mtl_dir defaults to
  // the OBJ's directory.
  bool LoadObj(const std::string& obj_path, const std::string& mtl_dir = "");
  // Mesh::RemoveVertices (src/mesh.cc:1066-1164): keeps the vertices whose
  // entry is true and the faces that use only those; returns the number
  // removed or -1.
  int RemoveVertices(const std::vector<bool>& valid_vertex_table);
  const std::vector<std::vector<int>>& face_indices_per_material() const {
    return face_indices_per_material_;
  }

  const std::vector<Eigen::Vector3f>& vertices() const { return vertices_; }
  const std::vector<Eigen::Vector3f>& vertex_colors() const { return vertex_colors_; }
  const std::vector<Eigen::Vector3i>& vertex_indices() const { return vertex_indices_; }
  const std::vector<Eigen::Vector3f>& normals() const { return normals_; }
  const std::vector<Eigen::Vector3f>& face_normals() const { return face_normals_; }
  const std::vector<Eigen::Vector3i>& normal_indices() const { return normal_indices_; }
  const std::vector<Eigen::Vector2f>& uv() const { return uv_; }
  const std::vector<Eigen::Vector3i>& uv_indices() const { return uv_indices_; }
  const std::vector<int>& material_ids() const { return material_ids_; }
  const std::vector<ObjMaterial>& materials() const { return materials_; }

  bool set_vertices(const std::vector<Eigen::Vector3f>& v) { vertices_ = v; return true; }
  bool set_vertex_colors(const std::vector<Eigen::Vector3f>& c) { vertex_colors_ = c; return true; }
  bool set_vertex_indices(const std::vector<Eigen::Vector3i>& f) { vertex_indices_ = f; return true; }
  bool set_normals(const std::vector<Eigen::Vector3f>& n) { normals_ = n; return true; }
  bool set_face_normals(const std::vector<Eigen::Vector3f>& n) { face_normals_ = n; return true; }
  bool set_normal_indices(const std::vector<Eigen::Vector3i>& i) { normal_indices_ = i; return true; }
  bool set_uv(const std::vector<Eigen::Vector2f>& uv) { uv_ = uv; return true; }
  bool set_uv_indices(const std::vector<Eigen::Vector3i>& i) { uv_indices_ = i; return true; }
  bool set_material_ids(const std::vector<int>& ids);
  bool set_materials(const std::vector<ObjMaterial>& m) { materials_ = m; return true; }
  bool set_single_material(const ObjMaterial& m) {
    materials_.assign(1, m);
    material_ids_.assign(vertex_indices_.size(), 0);
    return true;
  }

 private:
  std::vector<Eigen::Vector3f> vertices_;
  std::vector<Eigen::Vector3f> vertex_colors_;
  std::vector<Eigen::Vector3i> vertex_indices_;
  std::vector<Eigen::Vector3f> normals_;
  std::vector<Eigen::Vector3f> face_normals_;
  std::vector<Eigen::Vector3i> normal_indices_;
  std::vector<Eigen::Vector2f> uv_;
  std::vector<Eigen::Vector3i> uv_indices_;
  std::vector<ObjMaterial> materials_;
  std::vector<int> material_ids_;
  std::vector<std::vector<int>> face_indices_per_material_;
};

}  // namespace ugu
