// fp32 vector helpers for the gfx950 kernels.  The operation order of every
// helper equals that of Shiro's Vec3f operators (reference
// include/shiro/render/Types.h:30-111) so that, with contraction disabled
// (-ffp-contract=off), sums and products round like the CPU path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DEV __device__ __forceinline__
// Large functions called from several places in one kernel are kept out of
// line so that the shading kernels stay within the instruction cache.
#define DEV_CALL __device__ __noinline__

namespace shiro_hip {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kInvPi = 1.0f / kPi;
constexpr float kInvFourPi = 1.0f / (4.0f * kPi);
constexpr float kRayEps = 1.0e-4f;
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kInf = __builtin_huge_valf();

struct V2 {
    float x, y;
};

struct V3 {
    float x, y, z;
};

DEV V3 mk(float x, float y, float z) { return V3{x, y, z}; }
DEV V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }
DEV V3 xyz(float4 a) { return V3{a.x, a.y, a.z}; }

DEV V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DEV V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEV V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
DEV V3 operator*(float s, V3 a) { return a * s; }
DEV V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
DEV V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }

DEV float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEV V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
DEV float len(V3 a) { return sqrtf(dot(a, a)); }

// Zero-length vectors normalise to zero (Types.h:84-87).
DEV V3 normalize(V3 a) {
    const float l = len(a);
    return l > 0.0f ? a / l : V3{0.0f, 0.0f, 0.0f};
}

DEV V3 safe_normalize(V3 a, V3 fallback) {
    const float l = len(a);
    return l > 0.0f ? a / l : fallback;
}

// std::clamp(v, lo, hi) semantics: v < lo ? lo : (hi < v ? hi : v).
DEV float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
DEV V3 clamp3(V3 v, float lo, float hi) { return V3{clampf(v.x, lo, hi), clampf(v.y, lo, hi), clampf(v.z, lo, hi)}; }

// std::max / std::min semantics (first argument returned on ties/NaN order).
DEV float fmaxs(float a, float b) { return a < b ? b : a; }
DEV float fmins(float a, float b) { return b < a ? b : a; }

DEV V3 lerp(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
DEV float max3(V3 v) { return fmaxs(v.x, fmaxs(v.y, v.z)); }
DEV V3 reflect(V3 i, V3 n) { return i - 2.0f * dot(i, n) * n; }
DEV float sq(float v) { return v * v; }
DEV float luminance(V3 v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }

// Clamp the largest component to `limit` by uniform scaling
// (CpuPathTracer.cpp:1224-1230).
DEV V3 clamp_max_component(V3 v, float limit) {
    const float c = max3(v);
    if (c <= limit || c <= 0.0f) return v;
    return v * (limit / c);
}

}  // namespace shiro_hip
