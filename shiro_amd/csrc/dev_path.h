// The path integrator of Shiro's CPU backend on the device: TracePath and
// DirectLighting (reference src/backend/cpu/CpuPathTracer.cpp:1511-2017),
// camera rays (:1633-1648) and the environment terms (:315-384, 1497-1509).
// Same control flow, sampler dimensions, clamps, MIS weights, Russian
// roulette and depth caps as the CPU path; intersection is the BVH4 traversal
// of dev_bvh.h.
#pragma once

#include "dev_bsdf.h"
#include "dev_bvh.h"
#include "dev_env.h"
#include "dev_sampler.h"
#include "dev_texture.h"

namespace shiro_hip {

constexpr float kDirectClamp = 24.0f;
constexpr float kThroughputClamp = 12.0f;
constexpr float kDeltaRoughness = 0.02f;
constexpr uint32_t kBranchesPerDepth = 12u;

template <bool COUNT>
struct Counters {
    unsigned long long rays = 0, shadowRays = 0, nodes = 0, tris = 0, taps = 0, steps = 0, hits = 0;
    unsigned long long shadowNodes = 0, shadowTris = 0;
    unsigned long long primaryRays = 0, primaryNodes = 0, primaryTris = 0;  // the depth-0 part of rays/nodes/tris
    DEV TravCount<COUNT> trav() { return TravCount<COUNT>{&nodes, &tris}; }
    DEV TravCount<COUNT> shadow_trav() { return TravCount<COUNT>{&shadowNodes, &shadowTris}; }
    DEV EnvCount<COUNT> env() { return EnvCount<COUNT>{&taps, &steps}; }
};

template <bool COUNT>
DEV bool occluded(const DevScene& S, V3 o, V3 d, float maxDist, Counters<COUNT>& c) {
    if (COUNT) c.shadowRays += 1;
    Ray r;
    r.o = o;
    r.d = d;
    r.tmin = kRayEps;
    r.tmax = maxDist;
    return intersect_any(S, r, c.shadow_trav());
}

// A wave-uniform record in device memory, read into scalar registers
// (every lane asks for the same address).
template <typename T>
DEV T uniform_load(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "record of whole words");
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
    uint32_t v[sizeof(T) / 4];
#pragma unroll
    for (size_t k = 0; k < sizeof(T) / 4; ++k) v[k] = __builtin_amdgcn_readfirstlane(w[k]);
    T out;
    __builtin_memcpy(&out, v, sizeof(T));
    return out;
}

// Light records: the first ones from the kernel arguments, the rest from
// the scene's device arrays.  Every index here is wave-uniform.
DEV DevDirLight dir_light(const DevScene& S, uint32_t j) {
    return j < kInlineDirLights ? S.dirIn[j] : uniform_load(S.dirs + j);
}
DEV DevDome dome_light(const DevScene& S, uint32_t i) {
    return i < kInlineDomes ? S.domeIn[i] : uniform_load(S.domes + i);
}
DEV DevEnv env_map(const DevScene& S, int32_t e) {
    const uint32_t i = static_cast<uint32_t>(e);
    return i < kInlineDomes ? S.envIn[i] : uniform_load(S.envs + i);
}

DEV V3 world_to_dome(const DevDome& d, V3 v) {
    return normalize(V3{dot(v, ld3(d.right)), dot(v, ld3(d.up)), dot(v, ld3(d.forward))});
}

DEV V3 dome_to_world(const DevDome& d, V3 v) {
    return normalize(ld3(d.right) * v.x + ld3(d.up) * v.y + ld3(d.forward) * v.z);
}

// EnvironmentRadiance (CpuPathTracer.cpp:334-345).
template <bool COUNT>
DEV V3 environment_radiance(const DevScene& S, V3 dir, Counters<COUNT>& c) {
    if (S.domeCount > 0) {
        V3 sum{0.0f, 0.0f, 0.0f};
        for (uint32_t i = 0; i < S.domeCount; ++i) {
            const DevDome d = dome_light(S, i);
            if (d.env >= 0) {
                sum = sum + env_eval(env_map(S, d.env), world_to_dome(d, dir), c.env()) * ld3(d.radiance);
            } else {
                sum = sum + ld3(d.radiance);
            }
        }
        return sum;
    }
    const float t = clampf(dir.y * 0.5f + 0.5f, 0.0f, 1.0f);
    return lerp(ld3(S.envBottom), ld3(S.envTop), t);
}

template <bool COUNT>
DEV float env_pdf_sq_sum(const DevScene& S, V3 dir, Counters<COUNT>& c) {
    float acc = 0.0f;
    for (uint32_t i = 0; i < S.domeCount; ++i) {
        const DevDome d = dome_light(S, i);
        const float p = d.env >= 0 ? env_pdf(env_map(S, d.env), world_to_dome(d, dir), c.env()) : kInvFourPi;
        acc += p * p;
    }
    return acc;
}

// SampleDomeLight (CpuPathTracer.cpp:354-384).
template <bool COUNT>
DEV bool dome_sample(const DevScene& S, const DevDome& d, V2 u, V3* dir, V3* rad, float* pdf, Counters<COUNT>& c,
                     V3* localOut = nullptr) {
    V3 ld;
    float lp;
    V3 lr;
    if (d.env >= 0) {
        lr = env_sample(env_map(S, d.env), u, &ld, &lp, c.env());
    } else {
        ld = uniform_sphere(u);
        lp = kInvFourPi;
        lr = V3{1.0f, 1.0f, 1.0f};
    }
    if (localOut) *localOut = ld;
    if (lp <= 0.0f) return false;
    *dir = dome_to_world(d, ld);
    *rad = lr * ld3(d.radiance);
    *pdf = lp;
    return true;
}

// GenerateCameraRay (CpuPathTracer.cpp:1633-1648).
DEV Ray camera_ray(const DevCamera& c, const DevSettings& st, uint32_t x, uint32_t y, V2 j) {
    const float fx = ((static_cast<float>(x) + j.x) / static_cast<float>(st.width)) * 2.0f - 1.0f;
    const float fy = ((static_cast<float>(y) + j.y) / static_cast<float>(st.height)) * 2.0f - 1.0f;
    const float th = c.tanHalfFov;
    const V3 d = ld3(c.forward) + ld3(c.right) * (fx * c.aspect * th) + ld3(c.up) * (fy * th);
    Ray r;
    r.o = ld3(c.position);
    r.d = normalize(d);
    r.tmin = kRayEps;
    r.tmax = kFltMax;
    return r;
}

// Surface data of a closest hit (CpuPathTracer.cpp:126-168, 446-454).
struct SurfaceHit {
    V3 p, ng, ns;
    uint32_t material;
    uint32_t prim;    // global triangle index
    uint32_t light;   // emissive-list index + 1, 0 if the triangle is not a light
};

DEV SurfaceHit fetch_hit(const DevScene& S, const Ray& ray, const HitRec& h) {
    SurfaceHit s;
    s.p = ray.o + ray.d * h.t;
    const float4* tp = S.tris + static_cast<size_t>(h.slot) * 3u;
    const V3 p0 = xyz(tp[0]), p1 = xyz(tp[1]), p2 = xyz(tp[2]);
    s.ng = normalize(cross(p1 - p0, p2 - p0));
    const float4* hd = S.hitdata + static_cast<size_t>(h.slot) * 3u;
    const float4 h0 = hd[0], h1 = hd[1], h2 = hd[2];
    s.material = __float_as_uint(h0.w);
    s.prim = __float_as_uint(tp[0].w);
    s.light = __float_as_uint(h2.w);
    s.ns = s.ng;
    if (__float_as_uint(h1.w) & kTriHasNormals) {
        const float b0 = 1.0f - h.b1 - h.b2;
        const V3 n = xyz(h0) * b0 + xyz(h1) * h.b1 + xyz(h2) * h.b2;
        if (len(n) > 0.0f) s.ns = normalize(n);
    }
    return s;
}

enum BounceType : uint32_t { kBounceNone = 0, kBounceDiffuse, kBounceReflective, kBounceTransmission };

// Per-path state carried between bounces (TracePath's locals,
// CpuPathTracer.cpp:1659-1670) plus the first-hit AOVs.
struct PathState {
    V3 o, d;          // current ray (tMin = 1e-4, tMax = FLT_MAX)
    V3 T;             // throughput
    V3 L;             // radiance gathered so far
    V3 ext;           // interior medium extinction
    float prevPdf;
    uint32_t bits;    // see kBits*
    V3 albedo, normal;
    float depth;      // first-hit distance, +inf if none
};

// PathState::bits layout.
constexpr uint32_t kBitsTypeMask = 3u;        // previous BounceType
constexpr uint32_t kBitsPrevDelta = 4u;
constexpr uint32_t kBitsInside = 8u;
constexpr uint32_t kBitsMediumOn = 16u;
constexpr uint32_t kBitsDiffuseShift = 8u;    // diffuse bounce count (8 bits)
constexpr uint32_t kBitsSpecularShift = 16u;  // specular bounce count (8 bits)

DEV PathState init_path(const DevCamera& cam, const DevSettings& st, uint32_t x, uint32_t y, const Sampler& smp) {
    PathState ps;
    const Ray r = camera_ray(cam, st, x, y, smp.next2(0u));
    ps.o = r.o;
    ps.d = r.d;
    ps.T = V3{1.0f, 1.0f, 1.0f};
    ps.L = V3{0.0f, 0.0f, 0.0f};
    ps.ext = V3{0.0f, 0.0f, 0.0f};
    ps.prevPdf = 0.0f;
    ps.bits = kBitsPrevDelta;  // previousBounceDelta starts true
    ps.albedo = ps.normal = V3{0.0f, 0.0f, 0.0f};
    ps.depth = kInf;
    return ps;
}

DEV Ray path_ray(const PathState& ps) {
    Ray r;
    r.o = ps.o;
    r.d = ps.d;
    r.tmin = kRayEps;
    r.tmax = kFltMax;
    return r;
}

// Environment seen by a ray that left the scene (CpuPathTracer.cpp:1675-1693).
// `factor` (optional) receives Le times the MIS weight: the radiance added
// per unit of throughput.
template <bool COUNT>
DEV void handle_miss(const DevScene& S, const DevSettings& st, uint32_t depth, PathState& ps, Counters<COUNT>& c,
                     V3* factor = nullptr, float* misOut = nullptr) {
    if (misOut) *misOut = 0.0f;
    if (depth > 0 || st.backgroundVisible) {
        const V3 Le = environment_radiance(S, ps.d, c);
        const uint32_t prevType = ps.bits & kBitsTypeMask;
        V3 dL;
        if (S.domeCount == 0 || depth == 0 || (ps.bits & kBitsPrevDelta) || prevType == kBounceTransmission) {
            dL = ps.T * Le;
            if (factor) *factor = Le;
            if (misOut) *misOut = 1.0f;
        } else {
            const float sq2 = env_pdf_sq_sum(S, ps.d, c);
            const float b2 = ps.prevPdf * ps.prevPdf;
            const float w = b2 / fmaxs(b2 + sq2, 1.0e-6f);
            dL = ps.T * Le * w;
            if (factor) *factor = Le * w;
            if (misOut) *misOut = w;
        }
        ps.L = ps.L + dL;
    } else if (factor) {
        *factor = V3{0.0f, 0.0f, 0.0f};
    }
}

// Emissive-triangle lights (CpuPathTracer.cpp:1443-1495, 1578-1628): entry
// i of the list is S.emissive[5 i .. 5 i + 4].
struct EmissiveLight {
    V3 p0, p1, p2, n, radiance;
    float area, weight;
    uint32_t prim;
};

DEV EmissiveLight load_light(const DevScene& S, uint32_t i) {
    const float4* q = S.emissive + static_cast<size_t>(i) * 5u;
    const float4 a = q[0], b = q[1], c = q[2], d = q[3], e = q[4];
    return EmissiveLight{xyz(a), xyz(b), xyz(c), xyz(d), xyz(e), a.w, b.w, __float_as_uint(c.w)};
}

// PickEmissiveTriangle: the first entry whose cumulative weight reaches
// clamp(u, 0, 1 - 2^-24) * sum (lower_bound), the last one if none does.
DEV uint32_t pick_light(const DevScene& S, float u) {
    const float target = clampf(u, 0.0f, 0.99999994f) * S.emissiveSum;
    uint32_t lo = 0, n = S.emissiveCount;
    while (n > 0) {
        const uint32_t half = n >> 1;
        if (S.emissiveCdf[lo + half] < target) {
            lo += half + 1u;
            n -= half + 1u;
        } else {
            n = half;
        }
    }
    return lo < S.emissiveCount ? lo : S.emissiveCount - 1u;
}

// EmissiveLightPdf: solid-angle pdf of reaching `pos` on light `light`
// (list index + 1) from `origin`; 0 for triangles that are not lights.
DEV float emissive_pdf(const DevScene& S, uint32_t light, V3 origin, V3 pos) {
    if (light == 0u || S.emissiveSum <= 0.0f) return 0.0f;
    const EmissiveLight e = load_light(S, light - 1u);
    if (e.area <= 0.0f) return 0.0f;
    const V3 d = pos - origin;
    const float d2 = dot(d, d);
    if (d2 <= 1.0e-10f) return 0.0f;
    const V3 ld = normalize(d);
    const float lc = fmaxs(fabsf(dot(e.n, ld * -1.0f)), 1.0e-6f);
    const float sel = e.weight / S.emissiveSum;
    return sel / e.area * d2 / lc;
}

// Shading frame at a closest hit.
struct Surface {
    V3 p, ng, n, cn, v;
    float t;
    uint32_t material;
    uint32_t prim;

    V2 uv;            // interpolated UV (textured materials only)
    Mat M;
};

// Everything of a bounce up to direct lighting (CpuPathTracer.cpp:1695-1748):
// medium attenuation, normal orientation, opacity, first-hit AOVs, emission.
// Returns false when the opacity test lets the ray pass through (the ray is
// advanced and the bounce ends without lighting, as the reference's
// `continue`).  TEX: the scene has textured materials; their per-hit
// constants are derived into *ts, which sf.M then points to.  `emitted`
// (optional) receives the radiance added per unit of throughput.
template <bool TEX>
DEV bool shade_hit(const DevScene& S, uint32_t depth, const Sampler& smp, const Ray& ray, const HitRec& h,
                   PathState& ps, Surface& sf, TexState* ts, V3* emitted = nullptr) {
    const SurfaceHit hit = fetch_hit(S, ray, h);
    if (ps.bits & kBitsMediumOn) {
        ps.T = ps.T * V3{expf(-ps.ext.x * h.t), expf(-ps.ext.y * h.t), expf(-ps.ext.z * h.t)};
    }
    const shiro_hip_material& m = S.materials[hit.material];
    sf.M = Mat{&m, S.derived + hit.material, S.derived + hit.material};
    sf.uv = V2{0.0f, 0.0f};
    if (TEX && (S.derived[hit.material].flags & kMatTextured)) {
        sf.uv = hit_uv(S, h.slot, h.b1, h.b2);
        texture_material(S, m, S.derived[hit.material], sf.uv, *ts);
        sf.M.d = &ts->eval;
        sf.M.ds = &ts->samp;
    }
    const MaterialDerived& md = *sf.M.d;
    const bool front = dot(hit.ng, ray.d) < 0.0f;
    const V3 ng = front ? hit.ng : hit.ng * -1.0f;
    V3 nbase = front ? hit.ns : hit.ns * -1.0f;
    if (dot(nbase, ng) < 0.0f) nbase = nbase * -1.0f;
    V3 n = nbase;
    if (md.flags & kMatNormalOverride) n = resolve_normal(m, nbase, false);
    if (dot(n, ng) < 0.0f) n = n * -1.0f;
    V3 cn = n;
    if (md.flags & kMatCoatNormalOverride) cn = resolve_normal(m, n, true);
    if (md.flags & kMatOpacity) {
        const float opacity = clampf(m.opacity, 0.0f, 1.0f);
        if (smp.next1(depth * kBranchesPerDepth) > opacity) {
            ps.o = hit.p + ray.d * kRayEps;
            return false;
        }
        ps.T = ps.T / fmaxs(opacity, 1.0e-4f);
    }
    if (depth == 0) {
        ps.albedo = ld3(md.diffuseColor);
        ps.normal = n * 0.5f + V3{0.5f, 0.5f, 0.5f};
        ps.depth = h.t;
    }
    if (emitted) *emitted = V3{0.0f, 0.0f, 0.0f};
    if (md.flags & kMatEmissive) {
        // Emission (CpuPathTracer.cpp:1728-1748): in full when seen directly
        // or through a delta or transmission bounce, otherwise MIS-weighted
        // against the emissive-triangle light sampling.
        const uint32_t prevType = ps.bits & kBitsTypeMask;
        const V3 Le = ld3(md.emission);
        if (depth == 0 || (ps.bits & kBitsPrevDelta) || prevType == kBounceTransmission) {
            ps.L = ps.L + ps.T * Le;
            if (emitted) *emitted = Le;
        } else {
            const float lp = emissive_pdf(S, hit.light, ray.o, hit.p);
            const float w = lp > 0.0f ? power_heuristic(ps.prevPdf, lp) : 1.0f;
            ps.L = ps.L + ps.T * Le * w;
            if (emitted) *emitted = Le * w;
        }
    }
    sf.p = hit.p;
    sf.ng = ng;
    sf.n = n;
    sf.cn = cn;
    sf.v = ray.d * -1.0f;
    sf.t = h.t;
    sf.material = hit.material;
    sf.prim = hit.prim;
    return true;
}

// Number of light samples of DirectLighting at one hit: each directional
// light once, each dome domeLightSamples times, one emissive triangle
// (CpuPathTracer.cpp:1529-1631).  The wavefront's shadow-record capacity
// per path (enqueue_wavefront) uses the same count.
DEV uint32_t direct_sample_count(const DevScene& S, const DevSettings& st) {
    const uint32_t K = st.domeLightSamples > 1u ? st.domeLightSamples : 1u;
    return S.dirCount + S.domeCount * K + (S.emissiveCount > 0u && S.emissiveSum > 0.0f ? 1u : 0u);
}

// Light sample j of DirectLighting: the shadow ray (from p + n*eps, up to
// *maxDist) and the contribution it carries if unoccluded.  group 0 =
// directional light or emissive triangle (added as is), 1 + i = dome i.
// Returns false when the sample needs no shadow ray.  `fOut` / `lfOut`
// (optional) receive the BSDF value and the light factor,
// contrib = clamp_max_component(f * lf, 24).
// `domeOut` (optional) receives, for a dome sample, the map-space
// direction the radiance was looked up at and the factor s of
// lf = env radiance * dome radiance * s.
struct DomeSampleInfo {
    V3 local;
    float s;
};

template <bool COUNT, bool PLAIN = false>
DEV bool direct_sample(const DevScene& S, const DevSettings& st, const Surface& sf, const Mat& M, const Sampler& smp,
                       uint32_t depth, uint32_t j, V3* dirOut, V3* contrib, uint32_t* group, float* maxDist,
                       Counters<COUNT>& c, V3* fOut = nullptr, V3* lfOut = nullptr,
                       DomeSampleInfo* domeOut = nullptr) {
    *maxDist = kFltMax;
    if (j < S.dirCount) {
        const DevDirLight dl = dir_light(S, j);
        const V3 ld = normalize(ld3(dl.direction) * -1.0f);
        const float nl = clampf(dot(sf.n, ld), 0.0f, 1.0f);
        if (nl <= 0.0f) return false;
        *dirOut = ld;
        const V3 f = bsdf_eval<PLAIN>(M, sf.n, sf.cn, sf.v, ld);
        *contrib = clamp_max_component(f * ld3(dl.radiance) * nl, kDirectClamp);
        if (fOut) *fOut = f;
        if (lfOut) *lfOut = ld3(dl.radiance) * nl;
        *group = 0u;
        return true;
    }
    const uint32_t K = st.domeLightSamples > 1u ? st.domeLightSamples : 1u;
    const uint32_t jj = j - S.dirCount;
    if (jj >= S.domeCount * K) {
        // Emissive triangle (CpuPathTracer.cpp:1578-1628).
        const uint32_t base = 8192u + depth * 8u;
        const EmissiveLight e = load_light(S, pick_light(S, smp.next1(base)));
        if (e.prim == sf.prim) return false;
        const V2 u = smp.next2(base + 1u);
        const float su = sqrtf(clampf(u.x, 0.0f, 1.0f));
        const float b0 = 1.0f - su;
        const float b1 = clampf(u.y, 0.0f, 1.0f) * su;
        const float b2 = 1.0f - b0 - b1;
        const V3 lpos = e.p0 * b0 + e.p1 * b1 + e.p2 * b2;
        const V3 tl = lpos - sf.p;
        const float d2 = dot(tl, tl);
        if (!(d2 > 1.0e-10f)) return false;
        const float dist = sqrtf(d2);
        const V3 ld = tl / dist;
        const float nl = clampf(dot(sf.n, ld), 0.0f, 1.0f);
        const float lc = fabsf(dot(e.n, ld * -1.0f));
        if (nl <= 0.0f || lc <= 0.0f) return false;
        const float sel = e.weight / S.emissiveSum;
        const float pa = sel / fmaxs(e.area, 1.0e-6f);
        const float lp = pa * d2 / fmaxs(lc, 1.0e-6f);
        if (!(lp > 0.0f)) return false;
        const float bp = bsdf_pdf<PLAIN>(M, sf.n, sf.cn, sf.v, ld);
        const float w = power_heuristic(lp, bp);
        const V3 f = bsdf_eval<PLAIN>(M, sf.n, sf.cn, sf.v, ld);
        *dirOut = ld;
        *contrib = clamp_max_component(f * e.radiance * (nl * w / lp), kDirectClamp);
        if (fOut) *fOut = f;
        if (lfOut) *lfOut = e.radiance * (nl * w / lp);
        *group = 0u;
        *maxDist = fmaxs(dist - 2.0f * kRayEps, kRayEps);
        return true;
    }
    const uint32_t li = jj / K, k = jj - li * K;
    V3 ld, lr;
    float lp = 0.0f;
    const uint32_t branch = 256u + depth * 512u + li * 128u + k * 2u;
    if (!dome_sample(S, dome_light(S, li), smp.next2(branch), &ld, &lr, &lp, c, domeOut ? &domeOut->local : nullptr) ||
        lp <= 0.0f)
        return false;
    const float nl = clampf(dot(sf.n, ld), 0.0f, 1.0f);
    if (nl <= 0.0f) return false;
    const float bp = bsdf_pdf<PLAIN>(M, sf.n, sf.cn, sf.v, ld);
    const float w = power_heuristic(lp, bp);
    *dirOut = ld;
    const V3 f = bsdf_eval<PLAIN>(M, sf.n, sf.cn, sf.v, ld);
    *contrib = clamp_max_component(f * lr * (nl * w / lp), kDirectClamp);
    if (fOut) *fOut = f;
    if (lfOut) *lfOut = lr * (nl * w / lp);
    if (domeOut) domeOut->s = nl * w / lp;
    *group = 1u + li;
    return true;
}

// Sums visible light samples in DirectLighting's order: directional lights
// add directly; the samples of dome i are summed and divided by K when the
// next group starts.
struct DirectSum {
    V3 L, acc;
    int dome;
    float K;
    DEV void init(const DevSettings& st) {
        L = acc = V3{0.0f, 0.0f, 0.0f};
        dome = -1;
        K = static_cast<float>(st.domeLightSamples > 1u ? st.domeLightSamples : 1u);
    }
    DEV void flush() {
        if (dome >= 0) L = L + acc / K;
        acc = V3{0.0f, 0.0f, 0.0f};
        dome = -1;
    }
    DEV void add(uint32_t group, V3 c, bool visible) {
        if (group == 0u) {
            flush();
            if (visible) L = L + c;
            return;
        }
        const int d = static_cast<int>(group) - 1;
        if (d != dome) {
            flush();
            dome = d;
        }
        if (visible) acc = acc + c;
    }
};

// What a bounce did to the throughput, for the adjoint:
// T' = clamp_max_component(T * f * c, 12), then T' / q when roulette ran.
struct BounceRec {
    V3 f;            // BSDF value of the sampled direction
    float c;         // cos / pdf
    V3 next;         // sampled direction
    bool reflective; // f is bsdf_eval(v, next) (not a transmission lobe)
    bool rr;         // roulette ran (and the path survived)
};

// Lobe choice, BSDF sampling, throughput update and Russian roulette
// (CpuPathTracer.cpp:1750-2013).  Returns false when the path ends.  `rec`
// (optional) receives the throughput factors of a surviving bounce.
template <bool PLAIN = false>
DEV bool sample_bounce(const DevScene& S, const DevSettings& st, uint32_t depth, const Sampler& smp,
                       const Surface& sf, PathState& ps, BounceRec* rec = nullptr) {
    const Mat& M = sf.M;
    const shiro_hip_material& m = *M.m;
    const MaterialDerived& md = *M.d;
    const MaterialDerived& ms = *M.ds;
    const uint32_t b0 = depth * kBranchesPerDepth;
    const V3 n = sf.n, cn = sf.cn, ng = sf.ng, v = sf.v;
    const V3 rd = sf.v * -1.0f;
    bool inside = (ps.bits & kBitsInside) != 0u;
    bool mediumOn = (ps.bits & kBitsMediumOn) != 0u;
    uint32_t nDiffuse = (ps.bits >> kBitsDiffuseShift) & 0xffu;
    uint32_t nSpecular = (ps.bits >> kBitsSpecularShift) & 0xffu;

    const float wd = ms.wDiffuse;
    float ws = ms.wBaseSpecular;
    const float wc = ms.wCoat;
    const float wtRaw = ms.wTransmission;
    float wt = wtRaw;
    if (wtRaw > 0.0f) {
        const float ei = inside ? fmaxs(m.ior, 1.0f) : 1.0f;
        const float et = inside ? 1.0f : fmaxs(m.ior, 1.0f);
        const float R = fresnel_dielectric(fabsf(dot(ng, v)), ei, et);
        const float pass = m.thinWalled ? pane_transmittance(R) : (1.0f - R);
        wt = wtRaw * pass;
        ws += wtRaw * (1.0f - pass);
    }
    const float wr = wd + ws + wc;
    const float wsum = wr + wt;
    if (wsum <= 0.0f) return false;
    const float sel = smp.next1(b0 + 1u) * wsum;
    const V2 du = smp.next2(b0 + 2u);

    V3 next{0.0f, 0.0f, 0.0f}, f{0.0f, 0.0f, 0.0f};
    float pdf = 1.0f;
    uint32_t type = kBounceDiffuse;
    bool delta = false, pickT = false, pickS = false, pickC = false;
    bool needEval = false, needPdf = false;
    if (sel < wt) {
        pickT = true;
    } else {
        const float rs = sel - wt;
        if (rs < wd) {
        } else if (rs < wd + ws) {
            pickS = true;
        } else {
            pickC = true;
        }
    }

    if (pickT) {
        if (nSpecular >= st.specularDepth) return false;
        ++nSpecular;
        type = kBounceTransmission;
        const float pT = fmaxs(wt / wsum, 1.0e-6f);
        const float rT = m.thinWalled ? 0.0f : ms.sampleRoughnessTrans;
        const float mt = clampf(m.transmission, 0.0f, 1.0f) * (1.0f - clampf(m.metallic, 0.0f, 1.0f));
        if (m.thinWalled) {
            next = rd;
            const float R = fresnel_dielectric(fabsf(dot(ng, v)), 1.0f, fmaxs(m.ior, 1.0f));
            const float pane = pane_transmittance(R);
            const V3 scale = ld3(md.transmissionTint) * coat_transmission_scale(M, cn, v, next) * mt;
            const float ac = fmaxs(fabsf(dot(ng, next)), 1.0e-4f);
            f = scale * pane / ac;
            pdf = pT;
            delta = true;
        } else {
            const float ei = inside ? fmaxs(m.ior, 1.0f) : 1.0f;
            const float et = inside ? 1.0f : fmaxs(m.ior, 1.0f);
            const float etaRatio = ei / et;
            const float relEta = et / fmaxs(ei, 1.0f);
            const float R = fresnel_dielectric(fabsf(dot(ng, v)), ei, et);
            V3 mn = ng;
            if (rT > kDeltaRoughness) {
                mn = sample_ggx_half(ng, du, rT);
                if (dot(mn, v) < 0.0f) mn = mn * -1.0f;
            }
            if (!refract_dir(rd, mn, etaRatio, &next)) {
                next = normalize(reflect(rd, mn));
                if (wr <= 0.0f) return false;
                needEval = needPdf = true;
                type = kBounceReflective;
                delta = rT <= kDeltaRoughness;
                pickT = false;
                pickS = true;
            } else {
                if (rT <= kDeltaRoughness) {
                    const V3 scale = ld3(md.transmissionTint) * coat_transmission_scale(M, cn, v, next) * mt;
                    const float ac = fmaxs(fabsf(dot(ng, next)), 1.0e-4f);
                    const float transport = 1.0f / fmaxs(relEta * relEta, 1.0e-6f);
                    f = scale * (1.0f - R) * transport / ac;
                    pdf = pT;
                    delta = true;
                } else {
                    f = transmission_lobe(M, ng, cn, v, next, ei, et, rT);
                    pdf = fmaxs(pT * transmission_pdf(ng, v, next, et / ei, rT), 1.0e-6f);
                    delta = false;
                }
                inside = !inside;
                if (inside) {
                    ps.ext = ld3(md.mediumExtinction);
                    mediumOn = m.transmissionDepth > 0.0f || max3(ld3(m.transmissionScatter)) > 0.0f;
                } else {
                    mediumOn = false;
                }
            }
        }
    } else if (pickS || pickC) {
        if (nSpecular >= st.specularDepth) return false;
        ++nSpecular;
        type = kBounceReflective;
        const V3 sn = pickC ? cn : n;
        const float sr = pickC ? ms.sampleRoughnessCoat : ms.sampleRoughnessBase;
        const V3 hv = sample_ggx_half(sn, du, sr);
        next = normalize(reflect(rd, hv));
        const float rc = clampf(dot(sn, next), 0.0f, 1.0f);
        if (rc <= 0.0f) return false;
        needEval = needPdf = true;
        delta = sr <= kDeltaRoughness;
    } else {
        if (nDiffuse >= st.diffuseDepth) return false;
        ++nDiffuse;
        // Guide-free diffuse sampling: cosine pdf times the reflective share
        // (CpuPathTracer.cpp:1948-1988 with no guide snapshot).
        next = sample_cosine(n, du);
        const float cpdf = clampf(dot(n, next), 0.0f, 1.0f) * kInvPi;
        pdf = fmaxs((wr / wsum) * (1.0f * cpdf + 0.0f * 0.0f), 1.0e-6f);
        needEval = true;
    }
    // The reflective lobes evaluate the BSDF (and its mixture pdf) once,
    // after the lobe branches, so that a wave whose lanes picked different
    // lobes runs these calls once rather than once per branch.
    if (needPdf) pdf = fmaxs((wr / wsum) * bsdf_pdf<PLAIN>(M, n, cn, v, next), 1.0e-6f);
    if (needEval) f = bsdf_eval<PLAIN>(M, n, cn, v, next);

    const float cosT = pickT ? fmaxs(fabsf(dot(ng, next)), 1.0e-4f) : clampf(dot(n, next), 0.0f, 1.0f);
    if (cosT <= 0.0f) return false;
    if (rec) {
        rec->f = f;
        rec->c = cosT / pdf;
        rec->next = next;
        rec->reflective = !pickT;
        rec->rr = false;
    }
    ps.T = clamp_max_component(ps.T * f * (cosT / pdf), kThroughputClamp);
    const V3 rn = pickT ? ng : n;
    const float side = dot(rn, next) >= 0.0f ? 1.0f : -1.0f;
    ps.o = sf.p + rn * (kRayEps * side);
    ps.d = next;
    ps.prevPdf = pdf;
    ps.bits = type | (delta ? kBitsPrevDelta : 0u) | (inside ? kBitsInside : 0u) | (mediumOn ? kBitsMediumOn : 0u) |
              (nDiffuse << kBitsDiffuseShift) | (nSpecular << kBitsSpecularShift);
    if (depth >= st.rrDepth) {
        const float q = clampf(max3(ps.T), 0.05f, 0.95f);
        if (smp.next1(b0 + 4u) > q) return false;
        ps.T = ps.T / q;
        if (rec) rec->rr = true;
    }
    return true;
}

// TracePath (CpuPathTracer.cpp:1650-2017) in one thread: the megakernel
// form.  Shadow rays are traced as soon as they are generated.
template <bool COUNT, bool TEX>
DEV PathState trace_path(const DevScene& S, const DevSettings& st, const DevCamera& cam, uint32_t x, uint32_t y,
                         const Sampler& smp, Counters<COUNT>& c) {
    PathState ps = init_path(cam, st, x, y, smp);
    const uint32_t nLight = direct_sample_count(S, st);
    for (uint32_t depth = 0; depth < st.maxDepth; ++depth) {
        const Ray ray = path_ray(ps);
        if (COUNT) c.rays += 1;
        const unsigned long long nodes0 = c.nodes, tris0 = c.tris;
        const HitRec h = intersect_closest(S, ray, c.trav());
        if (COUNT && depth == 0u) {
            c.primaryRays += 1;
            c.primaryNodes += c.nodes - nodes0;
            c.primaryTris += c.tris - tris0;
        }
        if (h.slot == kNoHit) {
            handle_miss(S, st, depth, ps, c);
            break;
        }
        if (COUNT) c.hits += 1;
        Surface sf;
        TexState ts;
        if (!shade_hit<TEX>(S, depth, smp, ray, h, ps, sf, &ts)) continue;
        const Mat& M = sf.M;
        DirectSum ds;
        ds.init(st);
        const V3 origin = sf.p + sf.n * kRayEps;
        for (uint32_t j = 0; j < nLight; ++j) {
            V3 ld, contrib;
            uint32_t group;
            float maxDist;
            if (!direct_sample(S, st, sf, M, smp, depth, j, &ld, &contrib, &group, &maxDist, c)) continue;
            ds.add(group, contrib, !occluded(S, origin, ld, maxDist, c));
        }
        ds.flush();
        ps.L = ps.L + ps.T * ds.L;
        if (!sample_bounce(S, st, depth, smp, sf, ps)) break;
    }
    return ps;
}

}  // namespace shiro_hip
