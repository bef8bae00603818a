// TEST INFRASTRUCTURE ONLY -- part of the CPU oracle, never linked into the
// product library.  Small fp32 vector helpers with the exact operation order
// of the reference's Vec2f/Vec3f/Vec4f operators (include/shiro/render/
// Types.h:10-111), so that sums and products round identically.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace orc {

struct V2 {
    float x = 0.0f, y = 0.0f;
};

struct V3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

struct V4 {
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f;
};

// Types.h:30-48 -- component-wise, left to right.
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return a * s; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V4 operator+(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline V4 operator/(V4 a, float s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }

// Types.h:70-72: (x*x' + y*y') + z*z'.
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Types.h:74-80.
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

inline float length(V3 a) { return std::sqrt(dot(a, a)); }

// Types.h:84-87: zero-length vectors normalise to zero.
inline V3 normalize(V3 a) {
    const float len = length(a);
    return len > 0.0f ? a / len : V3{};
}

inline float clampf(float v, float lo, float hi) { return std::clamp(v, lo, hi); }

inline V3 clamp3(V3 v, float lo, float hi) {
    return {clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)};
}

// Types.h:101-103: a*(1-t) + b*t.
inline V3 lerp(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }

inline float max3(V3 v) { return std::max(v.x, std::max(v.y, v.z)); }

// Types.h:109-111: i - (2*dot(i,n))*n.
inline V3 reflect(V3 i, V3 n) { return i - 2.0f * dot(i, n) * n; }

inline V3 v3(const float* p) { return {p[0], p[1], p[2]}; }

}  // namespace orc
