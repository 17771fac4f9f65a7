// Geometry value types of the planner C++ API.  With Eigen available these are Eigen's own
// types, so the API signatures are exactly the reference's; without Eigen (this image has none)
// a minimal value type with the same accessors is used.
#pragma once

#if __has_include(<Eigen/Geometry>)
#include <Eigen/Geometry>
namespace mpx {
using Isometry3d = Eigen::Isometry3d;
using Vector3d = Eigen::Vector3d;
inline double mat(const Isometry3d& T, int r, int c) { return T.matrix()(r, c); }
inline double vec(const Vector3d& v, int i) { return v(i); }
}  // namespace mpx
#else
#include <array>
namespace mpx {
struct Matrix4dView {
    std::array<double, 16> m;
    double operator()(int r, int c) const { return m[4 * r + c]; }
    double& operator()(int r, int c) { return m[4 * r + c]; }
};
struct Isometry3d {
    Matrix4dView M{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}};
    static Isometry3d Identity() { return Isometry3d(); }
    const Matrix4dView& matrix() const { return M; }
    Matrix4dView& matrix() { return M; }
};
struct Vector3d {
    std::array<double, 3> v{{0, 0, 0}};
    Vector3d() = default;
    Vector3d(double x, double y, double z) : v{{x, y, z}} {}
    static Vector3d Zero() { return Vector3d(); }
    double operator()(int i) const { return v[i]; }
    double& operator()(int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
    double& operator[](int i) { return v[i]; }
};
inline double mat(const Isometry3d& T, int r, int c) { return T.matrix()(r, c); }
inline double vec(const Vector3d& v, int i) { return v(i); }
}  // namespace mpx
#endif
