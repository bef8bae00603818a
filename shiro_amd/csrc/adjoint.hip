// Path-replay adjoint of the render with respect to texture texels and
// dome-map texels (texture extension; no counterpart in the reference).
//
// The forward estimate of one sample is L = sum_j T_j * C_j, where C_j is
// what vertex j adds per unit of throughput (emission, the clamped and
// averaged light samples of DirectLighting, or the environment on a miss)
// and T_{j+1} = RR(clamp_max_component(T_j * f_j * c_j, 12)) with the
// roulette division RR and c_j = cos / pdf.  Sampling is detached: lobe
// choices, directions, pdfs and MIS weights are constants (they read the
// sampling copy of the texels), so the derivative is that of the
// frozen-sampling render, whose finite differences it reproduces.
//
// One thread replays one (pixel, sample) path with the forward code
// (closest hit, shade_hit<true>, direct_sample, shadow rays, sample_bounce:
// the same decisions as the textured render; when the forward wavefront
// render of the same samples left a replay log, its hits and light-sample
// visibilities are read instead of tracing the rays again), records each
// vertex (throughput, attenuation, C_j, bounce factors, d f_j / d theta at
// the sampled direction), adds the light-sample gradients on the way, then
// walks the vertices backwards propagating dLoss/dT_j through the clamp and
// the roulette division.  Texel gradients are scattered with atomics through
// the bilinear footprint of each hit.  The dome maps enter only through the
// radiance an escaped ray or a dome light sample reads (the importance
// tables are sampling, hence detached), so their gradient is scattered where
// that radiance is looked up, weighted by what the lookup adds to the pixel:
// throughput x MIS weight on a miss, and the clamp's vector-Jacobian product
// x BSDF x (cos x MIS / pdf) / K for a visible light sample.
#include "kernels.h"

#include "dev_path.h"
#include "dev_ring.h"

namespace shiro_hip {

namespace {

enum VtxKind : uint32_t { kVtxBounce = 0, kVtxEnd = 1, kVtxPass = 2, kVtxMiss = 3 };

struct Vtx {
    V3 T;          // throughput at the vertex after medium attenuation and opacity
    V3 att;        // T / throughput on arrival
    V3 C;          // radiance added per unit of T
    V3 f;          // bounce: BSDF value
    V3 dfa, dfr;   // bounce: d f / d base colour (per channel), d f / d roughness
    V2 uv;
    float c;       // bounce: cos / pdf
    uint32_t material;
    uint32_t kind;
    uint32_t flags;  // 1: roulette ran, 2: textured bounce gradient recorded
};

// Vector-Jacobian product of clamp_max_component(v, lim).
DEV V3 clamp_max_vjp(V3 v, float lim, V3 g) {
    const float c = max3(v);
    if (c <= lim || c <= 0.0f) return g;
    const float s = lim / c;
    V3 r = g * s;
    const float corr = dot(g, v) * lim / (c * c);
    // max3 returns the first of equal maxima in x, y, z order.
    if (!(v.x < fmaxs(v.y, v.z))) {
        r.x -= corr;
    } else if (!(v.y < v.z)) {
        r.y -= corr;
    } else {
        r.z -= corr;
    }
    return r;
}

// Vector-Jacobian product of y / clamp(max3(y), 0.05, 0.95).
DEV V3 roulette_vjp(V3 y, V3 g) {
    const float m = max3(y);
    const float q = clampf(m, 0.05f, 0.95f);
    V3 r = g / q;
    if (!(m < 0.05f) && !(0.95f < m)) {
        const float corr = dot(g, y) / (q * q);
        if (!(y.x < fmaxs(y.y, y.z))) {
            r.x -= corr;
        } else if (!(y.y < y.z)) {
            r.y -= corr;
        } else {
            r.z -= corr;
        }
    }
    return r;
}

// Per-wave sums of texel gradients in LDS.  A wave's dome light samples are
// importance-sampled onto the same few bright texels, the escaped rays of
// its 8x8 tile read neighbouring map texels and its hits neighbouring
// texture texels, so adding them in LDS first leaves one global atomic per
// distinct texel and channel instead of one per lookup, which contended on
// the hot texels.  Open addressing over kGradCacheSlots keys (array <<
// kTexelBits | texel), where array < kMaxDomes names dome `array`'s map (3
// floats per texel) and kMaxDomes + t texture t (its channel count per
// texel); a lookup that finds no slot in kGradCacheProbes probes adds to
// global memory directly.
constexpr uint32_t kGradCacheSlots = 256;
constexpr int kGradCacheProbes = 4;
constexpr uint32_t kTexelBits = 27;

struct GradCache {
    uint32_t key[kGradCacheSlots];
    float val[kGradCacheSlots * 3];
};

// Every lane of the block (one wave) calls this before any adds.
DEV void cache_clear(GradCache& c) {
    for (uint32_t s = threadIdx.x; s < kGradCacheSlots; s += 64u) {
        c.key[s] = kEmpty;
        c.val[3u * s] = c.val[3u * s + 1u] = c.val[3u * s + 2u] = 0.0f;
    }
    __syncthreads();
}

// Adds channels 0..n-1 of v to texel `texel` of gradient array `array`.
DEV void add_global(const AdjointParams& P, uint32_t array, uint32_t texel, V3 v, uint32_t n) {
    float* q;
    if (array < static_cast<uint32_t>(kMaxDomes)) {
        q = P.envGrad[array] + static_cast<size_t>(texel) * 3u;
    } else {
        const uint32_t t = array - kMaxDomes;
        q = P.grad[t] + static_cast<size_t>(texel) * P.scene.textures[t].channels;
    }
    if (v.x != 0.0f) atomicAdd(q, v.x);
    if (n > 1u && v.y != 0.0f) atomicAdd(q + 1, v.y);
    if (n > 2u && v.z != 0.0f) atomicAdd(q + 2, v.z);
}

DEV void cache_add(GradCache& c, const AdjointParams& P, uint32_t array, uint32_t texel, V3 v, uint32_t n) {
    const uint32_t key = (array << kTexelBits) | texel;
    uint32_t slot = (key * 2654435761u) >> 24;  // 8-bit multiplicative hash
    for (int probe = 0; probe < kGradCacheProbes; ++probe) {
        const uint32_t prev = atomicCAS(&c.key[slot], kEmpty, key);
        if (prev == kEmpty || prev == key) {
            float* q = &c.val[3u * slot];
            if (v.x != 0.0f) atomicAdd(q, v.x);
            if (n > 1u && v.y != 0.0f) atomicAdd(q + 1, v.y);
            if (n > 2u && v.z != 0.0f) atomicAdd(q + 2, v.z);
            return;
        }
        slot = (slot + 1u) & (kGradCacheSlots - 1u);
    }
    add_global(P, array, texel, v, n);
}

// Every lane of the block calls this after its last add.
DEV void cache_flush(GradCache& c, const AdjointParams& P) {
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < kGradCacheSlots; s += 64u) {
        const uint32_t key = c.key[s];
        if (key == kEmpty) continue;
        add_global(P, key >> kTexelBits, key & ((1u << kTexelBits) - 1u),
                   V3{c.val[3u * s], c.val[3u * s + 1u], c.val[3u * s + 2u]}, 3u);
    }
}

// Adds g[0..n-1] per unit of the bilinear lookup of texture `index` at uv
// to its texel gradients.
DEV void scatter(const AdjointParams& P, GradCache& cache, uint32_t index, V2 uv, V3 g, uint32_t n) {
    const DevTexture& t = P.scene.textures[index];
    const TexTap tap = tex_tap(t.w, t.h, uv);
    for (int k = 0; k < 4; ++k) {
        if (tap.w[k] == 0.0f) continue;
        cache_add(cache, P, kMaxDomes + index, tap.i[k], g * tap.w[k], n);
    }
}

// Adds dLoss/dbaseColor and dLoss/droughness at a hit to the texel
// gradients of the material's textures.
DEV void scatter_grad(const AdjointParams& P, GradCache& cache, uint32_t material, V2 uv, V3 gBase, float gRough) {
    const DevScene& S = P.scene;
    const shiro_hip_material& m = S.materials[material];
    if (m.baseColorTexture >= 0 && P.grad[m.baseColorTexture]) {
        const uint32_t t = static_cast<uint32_t>(m.baseColorTexture);
        if (S.textures[t].channels >= 3u) {
            scatter(P, cache, t, uv, gBase, 3u);
        } else {
            scatter(P, cache, t, uv, V3{gBase.x + gBase.y + gBase.z, 0.0f, 0.0f}, 1u);
        }
    }
    if (m.roughnessTexture >= 0 && P.grad[m.roughnessTexture])
        scatter(P, cache, static_cast<uint32_t>(m.roughnessTexture), uv, V3{gRough, 0.0f, 0.0f}, 1u);
}

// Adds g (per channel, per unit of the map's radiance at map-space
// direction `local`) to the texel gradients of dome i's map through the
// bilinear footprint of the lookup.
DEV void scatter_env(const AdjointParams& P, GradCache& cache, uint32_t i, V3 local, V3 g) {
    const DevScene& S = P.scene;
    if (i >= static_cast<uint32_t>(kMaxDomes) || !P.envGrad[i]) return;
    const DevDome d = dome_light(S, i);
    if (d.env < 0) return;
    EnvTaps t;
    if (!env_taps(env_map(S, d.env), local, &t)) return;
    const V3 gr = g * ld3(d.radiance);
    for (int k = 0; k < 4; ++k) {
        if (t.w[k] == 0.0f) continue;
        cache_add(cache, P, i, t.idx[k], gr * t.w[k], 3u);
    }
}

DEV V3 ratio(V3 a, V3 b) {
    return V3{b.x != 0.0f ? a.x / b.x : 0.0f, b.y != 0.0f ? a.y / b.y : 0.0f, b.z != 0.0f ? a.z / b.z : 0.0f};
}

}  // namespace

// Replays the path of pixel (x, y), sample s with pixel adjoint a (the
// method described at the top of this file): adds its light-sample and
// texel gradients, and its dome-map ones to `cache`.  MAXD bounds the path
// length, sizing the vertex tape.  LOG: read hits and visibilities from the
// replay log (P.hitLog, P.visLog) instead of tracing.  PLAIN: the scene's
// materials need no anisotropy, thin film, sheen or subsurface wrap
// (DevScene::plainBsdf), and the BSDF code leaves those terms out, as the
// wavefront's shade kernels do; the values are those of the general code.
template <uint32_t MAXD, bool LOG, bool PLAIN>
DEV void adjoint_path(const AdjointParams& P, GradCache& cache, uint32_t x, uint32_t y, uint32_t s, V3 a) {
    const DevSettings& st = P.settings;
    const DevScene& S = P.scene;
    const Sampler smp{x, y, P.sampleStart + s};
    PathState ps = init_path(P.camera, st, x, y, smp);
    const uint32_t nLight = direct_sample_count(S, st);
    Counters<false> cnt;
    // Traversal stacks: the newest entries in the block's LDS ring
    // (dev_ring.h), older ones here; same traversal and results as
    // intersect_closest / intersect_any.
    const uint32_t lane = threadIdx.x;
    uint32_t ovNode[kStackSize];
    float ovT[kStackSize];
    Vtx tape[MAXD];
    uint32_t nv = 0;
    for (uint32_t depth = 0; depth < st.maxDepth && depth < MAXD; ++depth) {
        const Ray ray = path_ray(ps);
        const size_t li = replay_index(st.width, st.height, P.sampleCount, x, y, s, depth);
        HitRec h;
        if constexpr (LOG) {
            // The forward pass's hit of this vertex (its record as the
            // wavefront stores it; the triangle index follows from the slot).
            const float4 hh = P.hitLog[li];
            h.t = hh.x;
            h.b1 = hh.y;
            h.b2 = hh.z;
            h.slot = __float_as_uint(hh.w);
            h.prim = kNoHit;
        } else {
            h = intersect_closest_ring(S, ray, lane, ovNode, ovT);
        }
        Vtx& V = tape[nv++];
        V.flags = 0u;
        if (h.slot == kNoHit) {
            V.kind = kVtxMiss;
            V.T = ps.T;
            V.att = V3{1.0f, 1.0f, 1.0f};
            float mis = 0.0f;
            handle_miss(S, st, depth, ps, cnt, &V.C, &mis);
            // The escaped ray adds T * mis * sum_i env_i(dir) * radiance_i.
            if (P.envWanted && mis != 0.0f) {
                const V3 g = a * V.T * mis;
                for (uint32_t i = 0; i < S.domeCount; ++i)
                    scatter_env(P, cache, i, world_to_dome(dome_light(S, i), ps.d), g);
            }
            break;
        }
        const V3 Tin = ps.T;
        Surface sf;
        TexState ts;
        V3 emitted;
        const bool lit = shade_hit<true>(S, depth, smp, ray, h, ps, sf, &ts, &emitted);
        V.T = ps.T;
        V.att = ratio(ps.T, Tin);
        V.C = V3{0.0f, 0.0f, 0.0f};
        if (!lit) {
            V.kind = kVtxPass;
            continue;
        }
        const Mat& M = sf.M;
        const bool textured = (S.derived[sf.material].flags & kMatTextured) != 0u;
        V.C = emitted;
        // Light samples: forward sum and the gradient of each visible one.
        DirectSum ds;
        ds.init(st);
        const V3 origin = sf.p + sf.n * kRayEps;
        V3 gBase{0.0f, 0.0f, 0.0f};
        float gRough = 0.0f;
        const V3 aT = a * ps.T;
        const uint32_t visBits = LOG ? P.visLog[li] : 0u;
        uint32_t record = 0;
        for (uint32_t j = 0; j < nLight; ++j) {
            V3 ld, contrib, f, lf;
            uint32_t group;
            float maxDist;
            DomeSampleInfo dome;
            if (!direct_sample<false, PLAIN>(S, st, sf, M, smp, depth, j, &ld, &contrib, &group, &maxDist, cnt, &f,
                                             &lf, &dome))
                continue;
            bool visible;
            if constexpr (LOG) {
                // Light samples that need a shadow ray are numbered in order,
                // as the forward pass's shadow records are.
                visible = ((visBits >> (record & 31u)) & 1u) != 0u;
            } else {
                Ray sr;
                sr.o = origin;
                sr.d = ld;
                sr.tmin = kRayEps;
                sr.tmax = maxDist;
                visible = !intersect_any_ring(S, sr, lane, ovNode);
            }
            ++record;
            ds.add(group, contrib, visible);
            if (visible && (textured || (P.envWanted && group > 0u))) {
                const V3 gc = group == 0u ? aT : aT / ds.K;
                const V3 gy = clamp_max_vjp(f * lf, kDirectClamp, gc);
                if (textured) {
                    const V3 gf = gy * lf;
                    const BsdfGrad bg = bsdf_grad(M, sf.n, sf.cn, sf.v, ld);
                    gBase = gBase + gf * bg.dBase;
                    gRough += dot(gf, bg.dRough);
                }
                // lf = env(local) * radiance * s for the sample of dome group - 1.
                if (P.envWanted && group > 0u) scatter_env(P, cache, group - 1u, dome.local, gy * f * dome.s);
            }
        }
        ds.flush();
        ps.L = ps.L + ps.T * ds.L;
        V.C = V.C + ds.L;
        if (textured) scatter_grad(P, cache, sf.material, sf.uv, gBase, gRough);
        BounceRec br;
        const bool alive = sample_bounce<PLAIN>(S, st, depth, smp, sf, ps, &br);
        if (!alive) {
            V.kind = kVtxEnd;
            break;
        }
        V.kind = kVtxBounce;
        V.f = br.f;
        V.c = br.c;
        V.flags = br.rr ? 1u : 0u;
        if (textured && br.reflective) {
            const BsdfGrad bg = bsdf_grad(M, sf.n, sf.cn, sf.v, br.next);
            V.dfa = bg.dBase;
            V.dfr = bg.dRough;
            V.uv = sf.uv;
            V.material = sf.material;
            V.flags |= 2u;
        }
    }
    // Reverse sweep: g = dLoss / dT of the vertex after the current one.
    V3 g{0.0f, 0.0f, 0.0f};
    for (uint32_t k = nv; k-- > 0;) {
        const Vtx& V = tape[k];
        V3 gT{0.0f, 0.0f, 0.0f};
        if (V.kind == kVtxBounce) {
            const V3 y = V.T * V.f * V.c;
            V3 gy = g;
            if (V.flags & 1u) gy = roulette_vjp(clamp_max_component(y, kThroughputClamp), gy);
            gy = clamp_max_vjp(y, kThroughputClamp, gy);
            gT = gy * V.f * V.c;
            if (V.flags & 2u) {
                const V3 gf = gy * V.T * V.c;
                scatter_grad(P, cache, V.material, V.uv, gf * V.dfa, dot(gf, V.dfr));
            }
        } else if (V.kind == kVtxPass) {
            gT = g;
        }
        gT = gT + a * V.C;
        g = gT * V.att;
    }
}

// Grid: one 64-thread block per group of G = 2^groupLog samples of an
// 8/sqrt(G)-pixel square of an 8x8 tile, laid out as the wavefront's path
// groups (block w: samples (w / (tiles * G)) * G + [0, G), tile (w / G) %
// tiles, square w % G; thread = sample in group * R * R + pixel in square):
// the lanes of a wave replay nearly identical paths of one pixel or a few
// neighbouring ones, and their texel gradients land on the same entries of
// the block's cache.  MAXD bounds the path length, sizing the per-thread vertex tape.
// Register caps (c5, measured): the tracing replay at 128 VGPRs (4 waves
// per SIMD; it spills some registers there, and still runs faster than at
// 2 or 3 waves); the replay from the log at 168 (3 waves: 2.16 ms, against
// 2.35 at 4, 2.23 at 2 and 2.90 at 6).  PLAIN (adjoint_path) is instantiated
// for the replay from the log only.
template <uint32_t MAXD, bool LOG, bool PLAIN>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(LOG ? 3 : 4))) void k_adjoint(AdjointParams P) {
    __shared__ GradCache cache;
    cache_clear(cache);
    const uint32_t gl = P.groupLog, rl = 3u - (gl >> 1u), ql = 3u - rl;
    const uint32_t perGroup = P.tiles << gl;
    const uint32_t sg = blockIdx.x / perGroup;
    const uint32_t rem = blockIdx.x - sg * perGroup;
    const uint32_t tile = rem >> gl, q = rem & ((1u << gl) - 1u);
    const uint32_t lane = threadIdx.x;
    const uint32_t sub = lane >> (2u * rl), j = lane & ((1u << (2u * rl)) - 1u);
    const uint32_t s = (sg << gl) + sub;
    const uint32_t x = (tile % P.tilesX) * 8u + ((q & ((1u << ql) - 1u)) << rl) + (j & ((1u << rl) - 1u));
    const uint32_t y = (tile / P.tilesX) * 8u + ((q >> ql) << rl) + (j >> rl);
    const DevSettings& st = P.settings;
    if (x < st.width && y < st.height && s < P.sampleCount) {
        const float4 db = P.dBeauty[static_cast<size_t>(y) * st.width + x];
        const float inv = 1.0f / static_cast<float>(P.sampleCount);
        const V3 a{db.x * inv, db.y * inv, db.z * inv};
        if (a.x != 0.0f || a.y != 0.0f || a.z != 0.0f) adjoint_path<MAXD, LOG, PLAIN>(P, cache, x, y, s, a);
    }
    cache_flush(cache, P);
}

hipError_t launch_adjoint(const AdjointParams& P, hipStream_t s) {
    const uint64_t blocks = static_cast<uint64_t>(P.tiles) * P.sampleCount;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    const dim3 grid(static_cast<uint32_t>(blocks));
    const uint32_t md = P.settings.maxDepth;
    if (md > kAdjointMaxDepth) return hipErrorInvalidValue;
    const bool log = P.hitLog != nullptr && P.visLog != nullptr;
    const bool plain = log && P.scene.plainBsdf != 0u;
    const uint32_t v = (md <= 4u ? 0u : (md <= 8u ? 3u : 6u)) + (plain ? 2u : (log ? 1u : 0u));
    switch (v) {
        case 0: hipLaunchKernelGGL((k_adjoint<4, false, false>), grid, dim3(64), 0, s, P); break;
        case 1: hipLaunchKernelGGL((k_adjoint<4, true, false>), grid, dim3(64), 0, s, P); break;
        case 2: hipLaunchKernelGGL((k_adjoint<4, true, true>), grid, dim3(64), 0, s, P); break;
        case 3: hipLaunchKernelGGL((k_adjoint<8, false, false>), grid, dim3(64), 0, s, P); break;
        case 4: hipLaunchKernelGGL((k_adjoint<8, true, false>), grid, dim3(64), 0, s, P); break;
        case 5: hipLaunchKernelGGL((k_adjoint<8, true, true>), grid, dim3(64), 0, s, P); break;
        case 6: hipLaunchKernelGGL((k_adjoint<kAdjointMaxDepth, false, false>), grid, dim3(64), 0, s, P); break;
        case 7: hipLaunchKernelGGL((k_adjoint<kAdjointMaxDepth, true, false>), grid, dim3(64), 0, s, P); break;
        default: hipLaunchKernelGGL((k_adjoint<kAdjointMaxDepth, true, true>), grid, dim3(64), 0, s, P); break;
    }
    return hipGetLastError();
}

}  // namespace shiro_hip
