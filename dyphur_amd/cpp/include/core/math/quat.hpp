#pragma once
// Quat<T> (core/include/core/math/quat.hpp): w = scalar part.
#include "vec3.hpp"

namespace dyphur {

template <typename T>
struct Quat {
    T w{T{1}};
    Vec3<T> xyz{};
    constexpr Quat() = default;
    constexpr Quat(T w_, Vec3<T> v) : w(w_), xyz(v) {}
    constexpr Quat(T w_, T x, T y, T z) : w(w_), xyz{x, y, z} {}
    static Quat from_axis_angle(const Vec3<T> &axis, T angle) {
        const T h = angle * T{0.5};
        return {std::cos(h), axis * std::sin(h)};
    }
    // Hamilton product.
    constexpr Quat operator*(const Quat &o) const {
        return {w * o.w - xyz.dot(o.xyz), xyz.cross(o.xyz) + o.xyz * w + xyz * o.w};
    }
    constexpr Quat conjugate() const { return {w, -xyz}; }
    constexpr Quat operator-() const { return {-w, -xyz}; }
    constexpr T norm_sq() const { return w * w + xyz.norm_sq(); }
    T norm() const { return std::sqrt(norm_sq()); }
    Quat normalized() const { const T n = norm(); return {w / n, xyz / n}; }
    // q v q* for a unit q.
    constexpr Vec3<T> rotate(const Vec3<T> &v) const {
        const Vec3<T> t = xyz.cross(v) * T{2};
        return v + t * w + xyz.cross(t);
    }
    static constexpr Quat identity() { return {T{1}, Vec3<T>::zero()}; }
};

using Quatf = Quat<float>;
using Quatd = Quat<double>;

}  // namespace dyphur
