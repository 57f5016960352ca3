// Minimal fixed-size subset of Eigen's API, enough for the types that appear
// in the public signatures of the renderer (Eigen::Vector{2,3}{f,i,d},
// Eigen::Matrix{3,4}{f,d}, Eigen::Affine3d). This image has no Eigen; when
// the drop-in API is compiled inside UGU's tree, define UGU_AMD_SYSTEM_EIGEN
// and the real <Eigen/Core>/<Eigen/Geometry> are used instead (the adapter
// only calls members that exist in both).
//
// Storage is column-major as in Eigen. Fixed-size reductions are evaluated
// x0 + (x1 + x2) for 3 terms, the order of Eigen's unrolled redux.
#pragma once

#ifdef UGU_AMD_SYSTEM_EIGEN
#include <Eigen/Core>
#include <Eigen/Geometry>
#else

#include <cmath>
#include <cstddef>
#include <initializer_list>

namespace Eigen {

using Index = std::ptrdiff_t;

namespace internal {
template <typename T, int N>
inline T redux_sum(const T* v) {
  // Eigen's redux_novec_unroller splits a length-N sum as N/2 + (N - N/2)
  if constexpr (N == 1) {
    return v[0];
  } else {
    constexpr int h = N / 2;
    return redux_sum<T, h>(v) + redux_sum<T, N - h>(v + h);
  }
}
}  // namespace internal

template <typename T, int R, int C>
class Matrix {
 public:
  using Scalar = T;
  static constexpr int RowsAtCompileTime = R;
  static constexpr int ColsAtCompileTime = C;

  Matrix() {
    for (int i = 0; i < R * C; i++) m_[i] = T(0);
  }
  Matrix(T x, T y) {
    static_assert(R * C == 2, "2 coefficients");
    m_[0] = x;
    m_[1] = y;
  }
  Matrix(T x, T y, T z) {
    static_assert(R * C == 3, "3 coefficients");
    m_[0] = x;
    m_[1] = y;
    m_[2] = z;
  }
  Matrix(T x, T y, T z, T w) {
    static_assert(R * C == 4, "4 coefficients");
    m_[0] = x;
    m_[1] = y;
    m_[2] = z;
    m_[3] = w;
  }

  static Matrix Zero() { return Matrix(); }
  static Matrix Ones() {
    Matrix r;
    for (int i = 0; i < R * C; i++) r.m_[i] = T(1);
    return r;
  }
  static Matrix Identity() {
    Matrix r;
    for (int i = 0; i < (R < C ? R : C); i++) r(i, i) = T(1);
    return r;
  }
  void setZero() { *this = Zero(); }
  void setIdentity() { *this = Identity(); }

  static constexpr Index rows() { return R; }
  static constexpr Index cols() { return C; }
  static constexpr Index size() { return R * C; }

  T& operator()(Index r, Index c) { return m_[c * R + r]; }
  const T& operator()(Index r, Index c) const { return m_[c * R + r]; }
  T& operator()(Index i) { return m_[i]; }
  const T& operator()(Index i) const { return m_[i]; }
  T& operator[](Index i) { return m_[i]; }
  const T& operator[](Index i) const { return m_[i]; }
  T& coeffRef(Index r, Index c) { return (*this)(r, c); }
  T coeff(Index r, Index c) const { return (*this)(r, c); }
  T* data() { return m_; }
  const T* data() const { return m_; }

  T& x() { return m_[0]; }
  T& y() { return m_[1]; }
  T& z() { return m_[2]; }
  T& w() { return m_[3]; }
  const T& x() const { return m_[0]; }
  const T& y() const { return m_[1]; }
  const T& z() const { return m_[2]; }
  const T& w() const { return m_[3]; }

  Matrix operator+(const Matrix& o) const {
    Matrix r;
    for (int i = 0; i < R * C; i++) r.m_[i] = m_[i] + o.m_[i];
    return r;
  }
  Matrix operator-(const Matrix& o) const {
    Matrix r;
    for (int i = 0; i < R * C; i++) r.m_[i] = m_[i] - o.m_[i];
    return r;
  }
  Matrix operator-() const {
    Matrix r;
    for (int i = 0; i < R * C; i++) r.m_[i] = -m_[i];
    return r;
  }
  Matrix operator*(T s) const {
    Matrix r;
    for (int i = 0; i < R * C; i++) r.m_[i] = m_[i] * s;
    return r;
  }
  friend Matrix operator*(T s, const Matrix& a) {
    Matrix r;
    for (int i = 0; i < R * C; i++) r.m_[i] = s * a.m_[i];
    return r;
  }
  Matrix operator/(T s) const {
    Matrix r;
    for (int i = 0; i < R * C; i++) r.m_[i] = m_[i] / s;
    return r;
  }
  Matrix& operator+=(const Matrix& o) { return *this = *this + o; }
  Matrix& operator-=(const Matrix& o) { return *this = *this - o; }
  Matrix& operator*=(T s) { return *this = *this * s; }
  Matrix& operator/=(T s) { return *this = *this / s; }
  bool operator==(const Matrix& o) const {
    for (int i = 0; i < R * C; i++)
      if (m_[i] != o.m_[i]) return false;
    return true;
  }

  // matrix product, each coefficient a fixed-size redux over the inner dim
  template <int C2>
  Matrix<T, R, C2> operator*(const Matrix<T, C, C2>& o) const {
    Matrix<T, R, C2> r;
    for (int i = 0; i < R; i++)
      for (int j = 0; j < C2; j++) {
        T p[C];
        for (int k = 0; k < C; k++) p[k] = (*this)(i, k) * o(k, j);
        r(i, j) = internal::redux_sum<T, C>(p);
      }
    return r;
  }

  Matrix<T, C, R> transpose() const {
    Matrix<T, C, R> r;
    for (int i = 0; i < R; i++)
      for (int j = 0; j < C; j++) r(j, i) = (*this)(i, j);
    return r;
  }

  T dot(const Matrix& o) const {
    T p[R * C];
    for (int i = 0; i < R * C; i++) p[i] = m_[i] * o.m_[i];
    return internal::redux_sum<T, R * C>(p);
  }
  T squaredNorm() const { return dot(*this); }
  T norm() const { return std::sqrt(squaredNorm()); }
  void normalize() {
    T z = squaredNorm();
    if (z > T(0)) *this /= std::sqrt(z);
  }
  Matrix normalized() const {
    Matrix r = *this;
    r.normalize();
    return r;
  }
  Matrix cross(const Matrix& b) const {
    static_assert(R * C == 3, "cross of 3-vectors");
    return Matrix(m_[1] * b.m_[2] - m_[2] * b.m_[1], m_[2] * b.m_[0] - m_[0] * b.m_[2],
                  m_[0] * b.m_[1] - m_[1] * b.m_[0]);
  }
  T determinant() const {
    static_assert(R == 3 && C == 3, "3x3 determinant");
    const Matrix& m = *this;
    return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) -
           m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
           m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
  }

  template <typename U>
  Matrix<U, R, C> cast() const {
    Matrix<U, R, C> r;
    for (int i = 0; i < R * C; i++) r.data()[i] = static_cast<U>(m_[i]);
    return r;
  }

  template <int BR, int BC>
  Matrix<T, BR, BC> block(Index r0, Index c0) const {
    Matrix<T, BR, BC> r;
    for (int i = 0; i < BR; i++)
      for (int j = 0; j < BC; j++) r(i, j) = (*this)(r0 + i, c0 + j);
    return r;
  }
  template <int BR, int BC>
  void setBlock(Index r0, Index c0, const Matrix<T, BR, BC>& b) {
    for (int i = 0; i < BR; i++)
      for (int j = 0; j < BC; j++) (*this)(r0 + i, c0 + j) = b(i, j);
  }
  Matrix<T, R, 1> col(Index j) const { return block<R, 1>(0, j); }
  Matrix<T, 1, C> row(Index i) const { return block<1, C>(i, 0); }
  void setCol(Index j, const Matrix<T, R, 1>& v) { setBlock<R, 1>(0, j, v); }

 private:
  T m_[R * C];
};

using Vector2f = Matrix<float, 2, 1>;
using Vector3f = Matrix<float, 3, 1>;
using Vector4f = Matrix<float, 4, 1>;
using Vector2d = Matrix<double, 2, 1>;
using Vector3d = Matrix<double, 3, 1>;
using Vector3i = Matrix<int, 3, 1>;
using Vector2i = Matrix<int, 2, 1>;
using Matrix3f = Matrix<float, 3, 3>;
using Matrix3d = Matrix<double, 3, 3>;
using Matrix4f = Matrix<float, 4, 4>;
using Matrix4d = Matrix<double, 4, 4>;

enum TransformTraits { Isometry = 1, Affine = 2 };

// Transform<T, 3, Affine>: a 4x4 matrix whose last row is (0, 0, 0, 1).
template <typename T, int Dim, int Mode>
class Transform {
  static_assert(Dim == 3, "3D transforms only");

 public:
  using MatrixType = Matrix<T, 4, 4>;
  using LinearType = Matrix<T, 3, 3>;
  using VectorType = Matrix<T, 3, 1>;

  Transform() : m_(MatrixType::Identity()) {}
  explicit Transform(const MatrixType& m) : m_(m) {}
  static Transform Identity() { return Transform(); }

  MatrixType& matrix() { return m_; }
  const MatrixType& matrix() const { return m_; }
  LinearType linear() const { return m_.template block<3, 3>(0, 0); }
  VectorType translation() const { return m_.template block<3, 1>(0, 3); }
  void setLinear(const LinearType& l) { m_.template setBlock<3, 3>(0, 0, l); }
  void setTranslation(const VectorType& t) { m_.template setBlock<3, 1>(0, 3, t); }

  Transform operator*(const Transform& o) const { return Transform(m_ * o.m_); }
  VectorType operator*(const VectorType& p) const { return linear() * p + translation(); }

  // Transform::inverse(Affine): linear^-1 by cofactors, -linear^-1 * t
  Transform inverse() const {
    const LinearType l = linear();
    LinearType inv;
    auto cof = [&](int i, int j) {
      const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      return l(i1, j1) * l(i2, j2) - l(i1, j2) * l(i2, j1);
    };
    const T c0[3] = {cof(0, 0), cof(1, 0), cof(2, 0)};
    const T det = c0[0] * l(0, 0) + (c0[1] * l(1, 0) + c0[2] * l(2, 0));
    const T invdet = T(1) / det;
    for (int i = 0; i < 3; i++) {
      inv(0, i) = c0[i] * invdet;
      inv(1, i) = cof(i, 1) * invdet;
      inv(2, i) = cof(i, 2) * invdet;
    }
    Transform r;
    r.setLinear(inv);
    r.setTranslation(-(inv * translation()));
    return r;
  }

  // rotation(): the rotation factor of the polar decomposition (SVD based,
  // with the last singular direction flipped when U V^T is a reflection).
  LinearType rotation() const;

 private:
  MatrixType m_;
};

using Affine3d = Transform<double, 3, Affine>;
using Affine3f = Transform<float, 3, Affine>;

namespace internal {
// Jacobi eigen-decomposition of a symmetric 3x3 matrix: a = v diag(w) v^T.
template <typename T>
void jacobi_eigen3(Matrix<T, 3, 3> a, Matrix<T, 3, 3>* v, T w[3]) {
  *v = Matrix<T, 3, 3>::Identity();
  for (int sweep = 0; sweep < 50; sweep++) {
    T off = std::abs(a(0, 1)) + std::abs(a(0, 2)) + std::abs(a(1, 2));
    if (off == T(0)) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (a(p, q) == T(0)) continue;
        const T theta = (a(q, q) - a(p, p)) / (T(2) * a(p, q));
        const T t = (theta >= 0 ? T(1) : T(-1)) /
                    (std::abs(theta) + std::sqrt(theta * theta + T(1)));
        const T c = T(1) / std::sqrt(t * t + T(1)), s = t * c;
        for (int k = 0; k < 3; k++) {
          const T akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < 3; k++) {
          const T apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; k++) {
          const T vkp = (*v)(k, p), vkq = (*v)(k, q);
          (*v)(k, p) = c * vkp - s * vkq;
          (*v)(k, q) = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 3; i++) w[i] = a(i, i);
}
}  // namespace internal

template <typename T, int Dim, int Mode>
typename Transform<T, Dim, Mode>::LinearType Transform<T, Dim, Mode>::rotation() const {
  const LinearType l = linear();
  LinearType V;
  T w[3];
  internal::jacobi_eigen3<T>(l.transpose() * l, &V, w);
  // order singular values descending (JacobiSVD's convention)
  int idx[3] = {0, 1, 2};
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      if (w[idx[j]] > w[idx[i]]) {
        int t = idx[i];
        idx[i] = idx[j];
        idx[j] = t;
      }
  LinearType Vs, U;
  for (int j = 0; j < 3; j++) Vs.setCol(j, V.col(idx[j]));
  for (int j = 0; j < 2; j++) {
    VectorType u = l * Vs.col(j);
    U.setCol(j, u / u.norm());
  }
  // third left singular vector completes a right-handed basis, then the sign
  // follows det(U V^T) as in Transform::computeRotationScaling
  VectorType u2 = U.col(0).cross(U.col(1));
  VectorType lv2 = l * Vs.col(2);
  if (u2.dot(lv2) < T(0)) u2 = -u2;
  U.setCol(2, u2);
  if ((U * Vs.transpose()).determinant() < T(0)) U.setCol(2, -U.col(2));
  return U * Vs.transpose();
}

}  // namespace Eigen

#endif  // UGU_AMD_SYSTEM_EIGEN
