// Dome-light environment maps on the device: bilinear lat-long / angular
// lookup and importance sampling through the marginal (row) + conditional
// (column) CDFs built on the host.  Follows the reference's EnvironmentMap
// (src/render/EnvironmentMap.cpp:94-330): lower_bound search semantics,
// pixel-centre sample directions, and pdfs from the pixel mass.
#pragma once

#include "dev_math.h"
#include "layout.h"

namespace shiro_hip {

template <bool COUNT>
struct EnvCount {
    unsigned long long* taps;
    unsigned long long* steps;
    DEV void tap(int n) const {
        if (COUNT) *taps += n;
    }
    DEV void step(int n) const {
        if (COUNT) *steps += n;
    }
};

DEV float wrap01(float v) {
    v = fmodf(v, 1.0f);
    return v < 0.0f ? v + 1.0f : v;
}

// EnvironmentMap.cpp:266-300: u wraps, v clamps, half-texel offset.
template <bool COUNT>
DEV V3 env_bilinear(const DevEnv& e, float u, float v, const EnvCount<COUNT>& c) {
    u = wrap01(u) * static_cast<float>(e.w) - 0.5f;
    v = clampf(v, 0.0f, 1.0f) * static_cast<float>(e.h) - 0.5f;
    const int x0 = static_cast<int>(floorf(u));
    const int y0 = static_cast<int>(floorf(v));
    const int x1 = x0 + 1;
    const int H = static_cast<int>(e.h), W = static_cast<int>(e.w);
    const int y1 = y0 + 1 < H - 1 ? y0 + 1 : H - 1;
    const float fx = u - static_cast<float>(x0);
    const float fy = v - static_cast<float>(y0);
    auto wx = [&](int x) {
        int r = x % W;
        return r < 0 ? r + W : r;
    };
    auto cy = [&](int y) { return y < 0 ? 0 : (y > H - 1 ? H - 1 : y); };
    const size_t r0 = static_cast<size_t>(cy(y0)) * e.w, r1 = static_cast<size_t>(cy(y1)) * e.w;
    const int a = wx(x0), b = wx(x1);
    const V3 c00 = xyz(e.texels[r0 + a]);
    const V3 c10 = xyz(e.texels[r0 + b]);
    const V3 c01 = xyz(e.texels[r1 + a]);
    const V3 c11 = xyz(e.texels[r1 + b]);
    c.tap(4);
    return lerp(lerp(c00, c10, fx), lerp(c01, c11, fx), fy);
}

// The four texels and bilinear weights env_bilinear blends at (u, v):
// d env_bilinear / d texel[idx[k]] = w[k] (a clamped row may repeat a texel;
// its weights then add up).
struct EnvTaps {
    uint32_t idx[4];
    float w[4];
};

DEV EnvTaps env_bilinear_taps(const DevEnv& e, float u, float v) {
    u = wrap01(u) * static_cast<float>(e.w) - 0.5f;
    v = clampf(v, 0.0f, 1.0f) * static_cast<float>(e.h) - 0.5f;
    const int x0 = static_cast<int>(floorf(u));
    const int y0 = static_cast<int>(floorf(v));
    const int H = static_cast<int>(e.h), W = static_cast<int>(e.w);
    const int y1 = y0 + 1 < H - 1 ? y0 + 1 : H - 1;
    const float fx = u - static_cast<float>(x0);
    const float fy = v - static_cast<float>(y0);
    auto wx = [&](int x) {
        int r = x % W;
        return r < 0 ? r + W : r;
    };
    auto cy = [&](int y) { return y < 0 ? 0 : (y > H - 1 ? H - 1 : y); };
    const uint32_t r0 = static_cast<uint32_t>(cy(y0)) * e.w, r1 = static_cast<uint32_t>(cy(y1)) * e.w;
    const uint32_t a = static_cast<uint32_t>(wx(x0)), b = static_cast<uint32_t>(wx(x0 + 1));
    EnvTaps t;
    t.idx[0] = r0 + a;
    t.idx[1] = r0 + b;
    t.idx[2] = r1 + a;
    t.idx[3] = r1 + b;
    t.w[0] = (1.0f - fx) * (1.0f - fy);
    t.w[1] = fx * (1.0f - fy);
    t.w[2] = (1.0f - fx) * fy;
    t.w[3] = fx * fy;
    return t;
}

// Taps of env_eval at a map-space direction; false where the map adds
// nothing (outside an angular map's disc).
DEV bool env_taps(const DevEnv& e, V3 dir, EnvTaps* t) {
    const V3 d = normalize(dir);
    if (e.layout == SHIRO_HIP_ENV_ANGULAR) {
        const float radial = sqrtf(d.x * d.x + d.y * d.y);
        if (radial <= 1.0e-6f) {
            *t = env_bilinear_taps(e, 0.5f, 0.5f);
            return true;
        }
        const float r = acosf(clampf(d.z, -1.0f, 1.0f)) / kPi;
        const float u = 0.5f + 0.5f * r * (d.x / radial);
        const float v = 0.5f + 0.5f * r * (d.y / radial);
        if (u < 0.0f || u > 1.0f || v < 0.0f || v > 1.0f) return false;
        *t = env_bilinear_taps(e, u, v);
        return true;
    }
    const float phi = atan2f(d.x, d.z);
    const float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    *t = env_bilinear_taps(e, wrap01(phi / kTwoPi), clampf(theta / kPi, 0.0f, 1.0f));
    return true;
}

template <bool COUNT>
DEV V3 env_eval(const DevEnv& e, V3 dir, const EnvCount<COUNT>& c) {
    const V3 d = normalize(dir);
    if (e.layout == SHIRO_HIP_ENV_ANGULAR) {
        const float radial = sqrtf(d.x * d.x + d.y * d.y);
        if (radial <= 1.0e-6f) return env_bilinear(e, 0.5f, 0.5f, c);
        const float r = acosf(clampf(d.z, -1.0f, 1.0f)) / kPi;
        const float u = 0.5f + 0.5f * r * (d.x / radial);
        const float v = 0.5f + 0.5f * r * (d.y / radial);
        if (u < 0.0f || u > 1.0f || v < 0.0f || v > 1.0f) return V3{0.0f, 0.0f, 0.0f};
        return env_bilinear(e, u, v, c);
    }
    const float phi = atan2f(d.x, d.z);
    const float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    return env_bilinear(e, wrap01(phi / kTwoPi), clampf(theta / kPi, 0.0f, 1.0f), c);
}

// std::lower_bound over cdf[begin+1 .. end] of the clamped sample; returns
// the bucket relative to begin and its mass (EnvironmentMap.cpp:315-330).
// The bucket is clamped to the table so that a malformed table (NaN
// weights) cannot index past it.
// `guide` (G + 1 positions, build_cdf_guide) narrows the search to the
// positions of the keys (g - 1) / G and (g + 2) / G around the key's
// bucket g (one bucket of slack on each side for the rounding of key * G),
// which bracket the key's lower_bound; the result is that of the full
// search.
template <bool COUNT>
DEV uint32_t env_sample_cdf(const float* cdf, uint32_t begin, uint32_t end, float u, float* mass,
                            const uint32_t* guide, uint32_t G, const EnvCount<COUNT>& c) {
    const float key = clampf(u, 0.0f, 0x1.fffffep-1f);
    uint32_t g = static_cast<uint32_t>(key * static_cast<float>(G));
    g = g < G - 1u ? g : G - 1u;
    const uint32_t lo = guide[g > 0u ? g - 1u : 0u];
    const uint32_t hi = guide[g + 2u < G ? g + 2u : G];
    uint32_t first = lo;
    uint32_t count = hi - lo;
    int steps = 0;
    while (count > 0u) {
        const uint32_t half = count >> 1;
        const uint32_t mid = first + half;
        ++steps;
        if (cdf[mid] < key) {
            first = mid + 1u;
            count -= half + 1u;
        } else {
            count = half;
        }
    }
    c.step(steps + 2);
    uint32_t idx = first - 1u;
    if (idx > end - 1u) idx = end - 1u;
    *mass = cdf[idx + 1u] - cdf[idx];
    return idx - begin;
}

DEV V3 uniform_sphere(V2 u) {
    const float y = 1.0f - 2.0f * u.y;
    const float r = sqrtf(fmaxs(0.0f, 1.0f - y * y));
    const float phi = kTwoPi * u.x;
    return V3{sinf(phi) * r, y, cosf(phi) * r};
}

// EnvironmentMap::Sample (EnvironmentMap.cpp:109-163).
template <bool COUNT>
DEV V3 env_sample(const DevEnv& e, V2 u, V3* dirOut, float* pdfOut, const EnvCount<COUNT>& c) {
    if (e.layout == SHIRO_HIP_ENV_LATLONG && e.hasTable) {
        float rowMass = 0.0f, colMass = 0.0f;
        const uint32_t row = env_sample_cdf(e.rowCdf, 0u, e.h, u.y, &rowMass, e.rowGuide, kEnvRowGuide, c);
        const uint32_t off = row * (e.w + 1u);
        const uint32_t col = env_sample_cdf(e.condCdf, off, off + e.w, u.x, &colMass,
                                            e.condGuide + static_cast<size_t>(row) * (kEnvColGuide + 1u),
                                            kEnvColGuide, c);
        const float uu = (static_cast<float>(col) + 0.5f) / static_cast<float>(e.w);
        const float vv = (static_cast<float>(row) + 0.5f) / static_cast<float>(e.h);
        const float theta = vv * kPi;
        const float phi = uu * kTwoPi;
        const float st = sinf(theta);
        const V3 d = V3{sinf(phi) * st, cosf(theta), cosf(phi) * st};
        *dirOut = d;
        *pdfOut = st > 1.0e-6f ? rowMass * colMass * static_cast<float>(e.w * e.h) / (2.0f * kPi * kPi * st) : 0.0f;
        return env_eval(e, d, c);
    }
    const V3 d = uniform_sphere(u);
    *dirOut = d;
    *pdfOut = kInvFourPi;
    return env_eval(e, d, c);
}

// EnvironmentMap::Pdf (EnvironmentMap.cpp:165-183, 302-313).
template <bool COUNT>
DEV float env_pdf(const DevEnv& e, V3 dir, const EnvCount<COUNT>& c) {
    if (e.layout != SHIRO_HIP_ENV_LATLONG || !e.hasTable) return kInvFourPi;
    const V3 d = normalize(dir);
    const float phi = atan2f(d.x, d.z);
    const float theta = acosf(clampf(d.y, -1.0f, 1.0f));
    const float u = wrap01(phi / kTwoPi);
    const float v = clampf(theta / kPi, 0.0f, 1.0f);
    uint32_t x = static_cast<uint32_t>(u * static_cast<float>(e.w));
    uint32_t y = static_cast<uint32_t>(v * static_cast<float>(e.h));
    x = x < e.w - 1u ? x : e.w - 1u;
    y = y < e.h - 1u ? y : e.h - 1u;
    const float st = sinf(theta);
    if (st <= 1.0e-6f) return 0.0f;
    const float rowMass = e.rowCdf[y + 1u] - e.rowCdf[y];
    const size_t o = static_cast<size_t>(y) * (e.w + 1u);
    const float colMass = e.condCdf[o + x + 1u] - e.condCdf[o + x];
    c.step(4);
    return rowMass * colMass * static_cast<float>(e.w * e.h) / (2.0f * kPi * kPi * st);
}

}  // namespace shiro_hip
