// vec3.h -- three-float vector with the memory layout of glm::vec3 (12 bytes, x y z).
//
// The reference's public API passes glm::vec3 (src/particle.h:19-34, src/sph_engine.h:133-134);
// GLM is not a dependency of this library, so the same layout is provided here.  When GLM
// has not been included first, `glm::vec3` names this type so code written against the
// reference compiles unchanged.
#pragma once

#include <cmath>
#include <cstddef>

namespace sph {

struct vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    constexpr vec3() = default;
    constexpr explicit vec3(float s) : x(s), y(s), z(s) {}
    constexpr vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    vec3& operator+=(const vec3& o) {
        x = x + o.x;
        y = y + o.y;
        z = z + o.z;
        return *this;
    }
    vec3& operator-=(const vec3& o) {
        x = x - o.x;
        y = y - o.y;
        z = z - o.z;
        return *this;
    }
    vec3& operator*=(float s) {
        x = x * s;
        y = y * s;
        z = z * s;
        return *this;
    }
};

inline vec3 operator+(const vec3& a, const vec3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline vec3 operator-(const vec3& a, const vec3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline vec3 operator-(const vec3& a) { return {-a.x, -a.y, -a.z}; }
inline vec3 operator*(const vec3& a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline vec3 operator*(float s, const vec3& a) { return {s * a.x, s * a.y, s * a.z}; }
inline vec3 operator*(const vec3& a, const vec3& b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline vec3 operator/(const vec3& a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline bool operator==(const vec3& a, const vec3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// glm::dot for vec3 sums the component products left to right: (x*x + y*y) + z*z.
inline float dot(const vec3& a, const vec3& b) {
    vec3 t = a * b;
    return (t.x + t.y) + t.z;
}
inline float length(const vec3& a) { return std::sqrt(dot(a, a)); }

static_assert(sizeof(vec3) == 12, "vec3 must match glm::vec3's layout");

}  // namespace sph

#ifndef GLM_SETUP_INCLUDED
namespace glm {
using vec3 = ::sph::vec3;
using ::sph::dot;
using ::sph::length;
}  // namespace glm
#endif
