// Host-side vector helpers for scene building.  Operation order follows
// glam's Vec3A as the reference uses it: dot = (x*x' + y*y') + z*z',
// normalize = v * (1 / sqrt(dot)), cross = (y z' - z y', z x' - x z', x y' - y x').
#pragma once
#include <cmath>
#include <cstdint>

namespace ag {

struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3& o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator*(const V3& o) const { return V3(x * o.x, y * o.y, z * o.z); }
    V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(float s) const { return V3(x / s, y / s, z / s); }
    V3 operator-() const { return V3(-x, -y, -z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length_sq(V3 a) { return dot(a, a); }
inline float length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a * (1.0f / length(a)); }
inline V3 vmin(V3 a, V3 b) { return V3(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return V3(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }
inline V3 reflect(V3 wo, V3 n) { return -wo + n * (2.0f * dot(wo, n)); }

// rot_y (geom.rs:94-97): deg.to_radians() is deg * (pi / 180) in f32.
inline V3 rot_y(V3 v, float deg) {
    const float k = 3.14159265358979323846f / 180.0f;
    float r = deg * k;
    float s = std::sin(r), c = std::cos(r);
    return V3(c * v.x + s * v.z, v.y, -s * v.x + c * v.z);
}

}  // namespace ag
