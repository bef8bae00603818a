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

namespace shiro_hip {

constexpr float kDirectClamp = 24.0f;
constexpr float kThroughputClamp = 12.0f;
constexpr float kDeltaRoughness = 0.02f;
constexpr uint32_t kBranchesPerDepth = 12u;

template <bool COUNT>
struct Counters {
    unsigned long long rays = 0, shadowRays = 0, nodes = 0, tris = 0, taps = 0, steps = 0, hits = 0;
    DEV TravCount<COUNT> trav() { return TravCount<COUNT>{&nodes, &tris}; }
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
    return intersect_any(S, r, c.trav());
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
            const DevDome& d = S.domes[i];
            if (d.env >= 0) {
                sum = sum + env_eval(S.envs[d.env], world_to_dome(d, dir), c.env()) * ld3(d.radiance);
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
        const DevDome& d = S.domes[i];
        const float p = d.env >= 0 ? env_pdf(S.envs[d.env], world_to_dome(d, dir), c.env()) : kInvFourPi;
        acc += p * p;
    }
    return acc;
}

// SampleDomeLight (CpuPathTracer.cpp:354-384).
template <bool COUNT>
DEV bool dome_sample(const DevScene& S, const DevDome& d, V2 u, V3* dir, V3* rad, float* pdf, Counters<COUNT>& c) {
    V3 ld;
    float lp;
    V3 lr;
    if (d.env >= 0) {
        lr = env_sample(S.envs[d.env], u, &ld, &lp, c.env());
    } else {
        ld = uniform_sphere(u);
        lp = kInvFourPi;
        lr = V3{1.0f, 1.0f, 1.0f};
    }
    if (lp <= 0.0f) return false;
    *dir = dome_to_world(d, ld);
    *rad = lr * ld3(d.radiance);
    *pdf = lp;
    return true;
}

// DirectLighting (CpuPathTracer.cpp:1511-1631) for directional and dome
// lights.
template <bool COUNT>
DEV V3 direct_lighting(const DevScene& S, const DevSettings& st, V3 p, V3 n, V3 cn, const Mat& M, V3 v,
                       const Sampler& smp, uint32_t depth, Counters<COUNT>& c) {
    V3 L{0.0f, 0.0f, 0.0f};
    const V3 origin = p + n * kRayEps;
    for (uint32_t i = 0; i < S.dirCount; ++i) {
        const DevDirLight& dl = S.dirs[i];
        const V3 ld = normalize(ld3(dl.direction) * -1.0f);
        const float nl = clampf(dot(n, ld), 0.0f, 1.0f);
        if (nl <= 0.0f || occluded(S, origin, ld, kFltMax, c)) continue;
        L = L + clamp_max_component(bsdf_eval(M, n, cn, v, ld) * ld3(dl.radiance) * nl, kDirectClamp);
    }
    const uint32_t K = st.domeLightSamples > 1u ? st.domeLightSamples : 1u;
    for (uint32_t li = 0; li < S.domeCount; ++li) {
        V3 acc{0.0f, 0.0f, 0.0f};
        for (uint32_t k = 0; k < K; ++k) {
            V3 ld, lr;
            float lp = 0.0f;
            const uint32_t branch = 256u + depth * 512u + li * 128u + k * 2u;
            if (!dome_sample(S, S.domes[li], smp.next2(branch), &ld, &lr, &lp, c) || lp <= 0.0f) continue;
            const float nl = clampf(dot(n, ld), 0.0f, 1.0f);
            if (nl <= 0.0f || occluded(S, origin, ld, kFltMax, c)) continue;
            const float bp = bsdf_pdf(M, n, cn, v, ld);
            const float w = power_heuristic(lp, bp);
            acc = acc + clamp_max_component(bsdf_eval(M, n, cn, v, ld) * lr * (nl * w / lp), kDirectClamp);
        }
        L = L + acc / static_cast<float>(K);
    }
    return L;
}

// GenerateCameraRay (CpuPathTracer.cpp:1633-1648).
DEV Ray camera_ray(const DevCamera& c, const DevSettings& st, uint32_t x, uint32_t y, V2 j) {
    const float fx = ((static_cast<float>(x) + j.x) / static_cast<float>(st.width)) * 2.0f - 1.0f;
    const float fy = ((static_cast<float>(y) + j.y) / static_cast<float>(st.height)) * 2.0f - 1.0f;
    const float th = tanf(c.vfov * 0.5f * kPi / 180.0f);
    const V3 d = ld3(c.forward) + ld3(c.right) * (fx * c.aspect * th) + ld3(c.up) * (fy * th);
    Ray r;
    r.o = ld3(c.position);
    r.d = normalize(d);
    r.tmin = kRayEps;
    r.tmax = kFltMax;
    return r;
}

struct PathResult {
    V3 radiance, albedo, normal;
    float depth;
};

// Surface data of a closest hit (CpuPathTracer.cpp:126-168, 446-454).
struct SurfaceHit {
    V3 p, ng, ns;
    uint32_t material;
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
    s.ns = s.ng;
    if (__float_as_uint(h1.w) & kTriHasNormals) {
        const float b0 = 1.0f - h.b1 - h.b2;
        const V3 n = xyz(h0) * b0 + xyz(h1) * h.b1 + xyz(h2) * h.b2;
        if (len(n) > 0.0f) s.ns = normalize(n);
    }
    return s;
}

enum BounceType : uint32_t { kBounceNone = 0, kBounceDiffuse, kBounceReflective, kBounceTransmission };

// TracePath (CpuPathTracer.cpp:1650-2017) without the path guide and
// emissive-triangle lights (the uploader reports scenes with emissive
// surfaces as unsupported, so the caller falls back to the CPU backend).
template <bool COUNT>
DEV PathResult trace_path(const DevScene& S, const DevSettings& st, Ray ray, const Sampler& smp, Counters<COUNT>& c) {
    PathResult res;
    res.radiance = res.albedo = res.normal = V3{0.0f, 0.0f, 0.0f};
    res.depth = kInf;
    V3 T{1.0f, 1.0f, 1.0f};
    uint32_t nDiffuse = 0, nSpecular = 0;
    uint32_t prevType = kBounceNone;
    bool prevDelta = true;
    float prevPdf = 0.0f;
    bool inside = false, mediumOn = false;
    V3 ext{0.0f, 0.0f, 0.0f};
    for (uint32_t depth = 0; depth < st.maxDepth; ++depth) {
        const uint32_t b0 = depth * kBranchesPerDepth;
        if (COUNT) c.rays += 1;
        const HitRec h = intersect_closest(S, ray, c.trav());
        if (h.slot == kNoHit) {
            if (depth > 0 || st.backgroundVisible) {
                const V3 Le = environment_radiance(S, ray.d, c);
                V3 dL;
                if (S.domeCount == 0 || depth == 0 || prevDelta || prevType == kBounceTransmission) {
                    dL = T * Le;
                } else {
                    const float sq2 = env_pdf_sq_sum(S, ray.d, c);
                    const float b2 = prevPdf * prevPdf;
                    dL = T * Le * (b2 / fmaxs(b2 + sq2, 1.0e-6f));
                }
                res.radiance = res.radiance + dL;
            }
            break;
        }
        if (COUNT) c.hits += 1;
        const SurfaceHit hit = fetch_hit(S, ray, h);
        if (mediumOn) {
            T = T * V3{expf(-ext.x * h.t), expf(-ext.y * h.t), expf(-ext.z * h.t)};
        }
        const Mat M{S.materials + hit.material, S.derived + hit.material};
        const shiro_hip_material& m = *M.m;
        const MaterialDerived& md = *M.d;
        const bool front = dot(hit.ng, ray.d) < 0.0f;
        const V3 ng = front ? hit.ng : hit.ng * -1.0f;
        V3 nbase = front ? hit.ns : hit.ns * -1.0f;
        if (dot(nbase, ng) < 0.0f) nbase = nbase * -1.0f;
        V3 n = nbase;
        if (md.flags & kMatNormalOverride) n = resolve_normal(m, nbase, false);
        if (dot(n, ng) < 0.0f) n = n * -1.0f;
        V3 cn = n;
        if (md.flags & kMatCoatNormalOverride) cn = resolve_normal(m, n, true);
        const V3 v = ray.d * -1.0f;
        if (md.flags & kMatOpacity) {
            const float opacity = clampf(m.opacity, 0.0f, 1.0f);
            if (smp.next1(b0) > opacity) {
                ray.o = hit.p + ray.d * kRayEps;
                continue;
            }
            T = T / fmaxs(opacity, 1.0e-4f);
        }
        if (depth == 0) {
            res.albedo = ld3(md.diffuseColor);
            res.normal = n * 0.5f + V3{0.5f, 0.5f, 0.5f};
            res.depth = h.t;
        }
        if (md.flags & kMatEmissive) {
            // Emission seen directly, through a delta or transmission
            // bounce (CpuPathTracer.cpp:1728-1748).
            if (depth == 0 || prevDelta || prevType == kBounceTransmission) {
                res.radiance = res.radiance + T * ld3(md.emission);
            }
        }
        res.radiance = res.radiance + T * direct_lighting(S, st, hit.p, n, cn, M, v, smp, depth, c);

        const float wd = md.wDiffuse;
        float ws = md.wBaseSpecular;
        const float wc = md.wCoat;
        const float wtRaw = md.wTransmission;
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
        if (wsum <= 0.0f) break;
        const float sel = smp.next1(b0 + 1u) * wsum;
        const V2 du = smp.next2(b0 + 2u);

        V3 next{0.0f, 0.0f, 0.0f}, f{0.0f, 0.0f, 0.0f};
        float pdf = 1.0f;
        uint32_t type = kBounceDiffuse;
        bool delta = false, pickT = false, pickS = false, pickC = false;
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
            if (nSpecular >= st.specularDepth) break;
            ++nSpecular;
            type = kBounceTransmission;
            const float pT = fmaxs(wt / wsum, 1.0e-6f);
            const float rT = m.thinWalled ? 0.0f : md.sampleRoughnessTrans;
            const float mt = clampf(m.transmission, 0.0f, 1.0f) * (1.0f - clampf(m.metallic, 0.0f, 1.0f));
            if (m.thinWalled) {
                next = ray.d;
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
                bool refracted;
                V3 mn = ng;
                if (rT > kDeltaRoughness) {
                    mn = sample_ggx_half(ng, du, rT);
                    if (dot(mn, v) < 0.0f) mn = mn * -1.0f;
                }
                refracted = refract_dir(ray.d, mn, etaRatio, &next);
                if (!refracted) {
                    next = normalize(reflect(ray.d, mn));
                    if (wr <= 0.0f) break;
                    f = bsdf_eval(M, n, cn, v, next);
                    pdf = fmaxs((wr / wsum) * bsdf_pdf(M, n, cn, v, next), 1.0e-6f);
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
                        ext = ld3(md.mediumExtinction);
                        mediumOn = m.transmissionDepth > 0.0f || max3(ld3(m.transmissionScatter)) > 0.0f;
                    } else {
                        mediumOn = false;
                    }
                }
            }
        } else if (pickS || pickC) {
            if (nSpecular >= st.specularDepth) break;
            ++nSpecular;
            type = kBounceReflective;
            const V3 sn = pickC ? cn : n;
            const float sr = pickC ? md.sampleRoughnessCoat : md.sampleRoughnessBase;
            const V3 hv = sample_ggx_half(sn, du, sr);
            next = normalize(reflect(ray.d, hv));
            const float rc = clampf(dot(sn, next), 0.0f, 1.0f);
            if (rc <= 0.0f) break;
            pdf = fmaxs((wr / wsum) * bsdf_pdf(M, n, cn, v, next), 1.0e-6f);
            f = bsdf_eval(M, n, cn, v, next);
            delta = sr <= kDeltaRoughness;
        } else {
            if (nDiffuse >= st.diffuseDepth) break;
            ++nDiffuse;
            next = sample_cosine(n, du);
            const float cpdf = clampf(dot(n, next), 0.0f, 1.0f) * kInvPi;
            pdf = fmaxs((wr / wsum) * (1.0f * cpdf + 0.0f * 0.0f), 1.0e-6f);
            f = bsdf_eval(M, n, cn, v, next);
        }

        const float cosT = pickT ? fmaxs(fabsf(dot(ng, next)), 1.0e-4f) : clampf(dot(n, next), 0.0f, 1.0f);
        if (cosT <= 0.0f) break;
        T = clamp_max_component(T * f * (cosT / pdf), kThroughputClamp);
        const V3 rn = pickT ? ng : n;
        const float side = dot(rn, next) >= 0.0f ? 1.0f : -1.0f;
        ray.o = hit.p + rn * (kRayEps * side);
        ray.d = next;
        prevType = type;
        prevDelta = delta;
        prevPdf = pdf;
        if (depth >= st.rrDepth) {
            const float q = clampf(max3(T), 0.05f, 0.95f);
            if (smp.next1(b0 + 4u) > q) break;
            T = T / q;
        }
    }
    return res;
}

}  // namespace shiro_hip
