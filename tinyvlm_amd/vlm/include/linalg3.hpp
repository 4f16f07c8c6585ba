// linalg3.hpp -- the small fixed-size vector/matrix types the Backend interface passes by
// reference (the reference uses linalg::float3 / double3 / mat<T,4,4>, headeronly/linalg.h).
// Only what the interface and the drivers need.
#pragma once

#include <cmath>

namespace linalg {

template <class T>
struct vec3 {
    T x{}, y{}, z{};
};
template <class T>
inline vec3<T> operator+(vec3<T> a, vec3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline vec3<T> operator-(vec3<T> a, vec3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline vec3<T> operator*(T s, vec3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T>
inline vec3<T> operator/(vec3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T>
inline T dot(vec3<T> a, vec3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
inline vec3<T> cross(vec3<T> a, vec3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
inline T length2(vec3<T> a) { return dot(a, a); }
template <class T>
inline T length(vec3<T> a) { return std::sqrt(length2(a)); }
template <class T>
inline vec3<T> normalize(vec3<T> a) { return a / length(a); }

using float3 = vec3<float>;
using double3 = vec3<double>;

// 4x4 homogeneous transform, column-major storage m[col * 4 + row] like linalg::mat::store.
template <class T>
struct mat4 {
    T m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    static mat4 translation(T x, T y, T z) {
        mat4 r;
        r.m[12] = x;
        r.m[13] = y;
        r.m[14] = z;
        return r;
    }
    vec3<T> apply(vec3<T> p) const {
        return {m[0] * p.x + m[4] * p.y + m[8] * p.z + m[12], m[1] * p.x + m[5] * p.y + m[9] * p.z + m[13],
                m[2] * p.x + m[6] * p.y + m[10] * p.z + m[14]};
    }
};

}  // namespace linalg
