// gfx950 kernels of the path-tracing backend and their host launchers.
//
//  k_render        the megakernel: one thread per pixel runs a slice of the
//                  batch's samples through the full TracePath and keeps the
//                  per-pixel running sums (beauty + hit count, albedo,
//                  normal, depth minimum) in HBM between slice launches; the
//                  last slice writes the batch means like
//                  CpuPathTracer::RenderSampleBatch (CpuPathTracer.cpp:2088-2137).
//                  One 64-lane wave covers an 8x8 pixel tile; tiles are
//                  dealt to the eight XCDs in contiguous ranges so that
//                  neighbouring tiles share an L2.
//  k_render_many   the batched form: persistent blocks take (item, tile)
//                  units of many scenes' batches from one cost-sorted list.
//  k_accum_*       FrameAccumulator on the device (Renderer.cpp:47-110).
//  k_qmc_fill, k_intersect, k_occluded, k_trace_primary, k_eval_bsdf,
//  k_env_sample, k_env_eval   single stages of the path for parity tests.
#include <hip/hip_runtime.h>

#include "dev_path.h"
#include "dev_ring.h"
#include "kernels.h"

namespace shiro_hip {

// Tile index for a block: hardware deals consecutive blocks round-robin to
// the 8 XCDs; remap so that XCD k gets the k-th contiguous range of tiles.
DEV uint32_t xcd_tile(uint32_t bid, uint32_t nblocks) {
    const uint32_t per = (nblocks + 7u) / 8u;
    const uint32_t xcd = bid % 8u, slot = bid / 8u;
    const uint32_t full = nblocks % 8u == 0u ? 8u : nblocks % 8u;  // XCDs with `per` blocks
    // Blocks [0, full*per) belong to XCDs that get `per` tiles, the rest get per-1.
    uint32_t base;
    if (xcd < full) {
        base = xcd * per;
    } else {
        base = full * per + (xcd - full) * (per - 1u);
    }
    return base + slot;
}

// The end of a launch for one pixel: the running sums kept for the next
// slice, or on the last slice the batch means, beauty divided by the count
// and albedo/normal scaled by its reciprocal (CpuPathTracer.cpp:2132-2137).
DEV void store_pixel(const RenderParams& P, size_t pix, float4 beauty, float4 albedo, float4 normal, float depth) {
    if (P.sliceBegin + P.sliceCount < P.batchCount) {
        P.sums.beauty[pix] = beauty;
        P.sums.albedo[pix] = albedo;
        P.sums.normal[pix] = normal;
        P.sums.depth[pix] = depth;
        return;
    }
    const float n = static_cast<float>(P.batchCount);
    const float inv = 1.0f / n;
    float* b = P.frame.beauty + pix * 4u;
    b[0] = beauty.x / n;
    b[1] = beauty.y / n;
    b[2] = beauty.z / n;
    b[3] = beauty.w / n;
    float* a = P.frame.albedo + pix * 3u;
    a[0] = albedo.x * inv;
    a[1] = albedo.y * inv;
    a[2] = albedo.z * inv;
    float* nn = P.frame.normal + pix * 3u;
    nn[0] = normal.x * inv;
    nn[1] = normal.y * inv;
    nn[2] = normal.z * inv;
    P.frame.depth[pix] = depth;
}

// The megakernel's work for one 8x8 tile of one batch (one lane per pixel).
template <bool COUNT, bool TEX>
DEV void render_tile(const RenderParams& P, uint32_t tile, uint32_t lane) {
    const uint32_t x = (tile % P.tilesX) * 8u + (lane & 7u);
    const uint32_t y = (tile / P.tilesX) * 8u + (lane >> 3);
    const DevSettings& st = P.settings;
    if (x >= st.width || y >= st.height) return;
    const size_t pix = static_cast<size_t>(y) * st.width + x;

    float4 beauty, albedo, normal;
    float depth;
    if (P.sliceBegin == 0u) {
        beauty = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        albedo = normal = beauty;
        depth = kInf;
    } else {
        beauty = P.sums.beauty[pix];
        albedo = P.sums.albedo[pix];
        normal = P.sums.normal[pix];
        depth = P.sums.depth[pix];
    }
    Counters<COUNT> cnt;
    for (uint32_t i = 0; i < P.sliceCount; ++i) {
        const Sampler smp{x, y, P.sampleStart + P.sliceBegin + i};
        const PathState pr = trace_path<COUNT, TEX>(P.scene, st, P.camera, x, y, smp, cnt);
        beauty.x = beauty.x + pr.L.x;
        beauty.y = beauty.y + pr.L.y;
        beauty.z = beauty.z + pr.L.z;
        beauty.w = beauty.w + (pr.depth < kInf ? 1.0f : 0.0f);
        albedo.x = albedo.x + pr.albedo.x;
        albedo.y = albedo.y + pr.albedo.y;
        albedo.z = albedo.z + pr.albedo.z;
        normal.x = normal.x + pr.normal.x;
        normal.y = normal.y + pr.normal.y;
        normal.z = normal.z + pr.normal.z;
        if (pr.depth < depth) depth = pr.depth;
    }
    if (COUNT) {
        DevCounters* c = P.counters;
        atomicAdd(&c->samples, static_cast<unsigned long long>(P.sliceCount));
        atomicAdd(&c->rays, cnt.rays);
        atomicAdd(&c->shadowRays, cnt.shadowRays);
        atomicAdd(&c->nodeFetches, cnt.nodes);
        atomicAdd(&c->triFetches, cnt.tris);
        atomicAdd(&c->envTaps, cnt.taps);
        atomicAdd(&c->cdfSteps, cnt.steps);
        atomicAdd(&c->hitFetches, cnt.hits);
        atomicAdd(&c->shadowNodeFetches, cnt.shadowNodes);
        atomicAdd(&c->shadowTriFetches, cnt.shadowTris);
        atomicAdd(&c->primaryRays, cnt.primaryRays);
        atomicAdd(&c->primaryNodeFetches, cnt.primaryNodes);
        atomicAdd(&c->primaryTriFetches, cnt.primaryTris);
    }
    store_pixel(P, pix, beauty, albedo, normal, depth);
}

template <bool COUNT, bool TEX>
__global__ __launch_bounds__(64) void k_render(RenderParams P) {
    render_tile<COUNT, TEX>(P, xcd_tile(blockIdx.x, gridDim.x), threadIdx.x);
}

// The megakernel with path regeneration ("megakernel+", the transport
// SHIRO_HIP_TRANSPORT_MEGAKERNEL_PLUS): persistent one-wave blocks whose
// lanes each take a pixel (in 8x8-tile order, from one counter, 64 at a time
// per wave), run its samples one after the other and take the next pixel
// when they are done.  The loop advances every lane by one bounce per
// iteration, so a lane whose path ends starts its next sample (or pixel)
// at once instead of idling until the wave's longest path ends, and the
// traversal stacks keep their newest kRing entries in LDS as the
// wavefront's traversal kernels do.  Every sample runs trace_path's
// bounce body (dev_path.h) with the same functions in the same order, and
// each pixel sums its samples in order, so frames equal k_render's bit for
// bit.
template <bool TEX>
__global__ __launch_bounds__(64) void k_render_plus(RenderParams P, uint32_t* counter) {
    const DevSettings& st = P.settings;
    const DevScene& S = P.scene;
    const uint32_t lane = threadIdx.x;
    const uint32_t total = P.tilesX * ((st.height + 7u) / 8u) * 64u;
    const uint32_t nLight = direct_sample_count(S, st);
    uint32_t ovNode[kStackSize];
    float ovT[kStackSize];
    Counters<false> c;
    bool active = false, exhausted = false;
    uint32_t x = 0, y = 0, i = 0, depth = 0;
    size_t pix = 0;
    float4 beauty{}, albedo{}, normal{};
    float dmin = kInf;
    PathState ps;
    for (;;) {
        // Lanes without a pixel take the next ones: one atomic per wave.
        const bool want = !active && !exhausted;
        const unsigned long long wm = __ballot(want);
        if (wm) {
            const uint32_t cnt = static_cast<uint32_t>(__popcll(wm));
            const int leader = __ffsll(static_cast<long long>(wm)) - 1;
            uint32_t first = 0;
            if (static_cast<int>(lane) == leader) first = atomicAdd(counter, cnt);
            first = __shfl(first, leader);
            if (want) {
                const uint32_t idx = first + static_cast<uint32_t>(__popcll(wm & ((1ull << lane) - 1ull)));
                if (idx >= total) {
                    exhausted = true;
                } else {
                    const uint32_t tile = idx >> 6, l = idx & 63u;
                    x = (tile % P.tilesX) * 8u + (l & 7u);
                    y = (tile / P.tilesX) * 8u + (l >> 3);
                    if (x < st.width && y < st.height) {
                        pix = static_cast<size_t>(y) * st.width + x;
                        if (P.sliceBegin == 0u) {
                            beauty = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            albedo = normal = beauty;
                            dmin = kInf;
                        } else {
                            beauty = P.sums.beauty[pix];
                            albedo = P.sums.albedo[pix];
                            normal = P.sums.normal[pix];
                            dmin = P.sums.depth[pix];
                        }
                        i = 0;
                        depth = 0;
                        ps = init_path(P.camera, st, x, y, Sampler{x, y, P.sampleStart + P.sliceBegin});
                        active = P.sliceCount > 0u;
                    }
                }
            }
        }
        if (__all(exhausted)) break;
        if (!active) continue;
        // One bounce of TracePath (dev_path.h:trace_path's loop body).
        const Sampler smp{x, y, P.sampleStart + P.sliceBegin + i};
        bool ended = depth >= st.maxDepth;
        if (!ended) {
            const Ray ray = path_ray(ps);
            const HitRec h = intersect_closest_ring(S, ray, lane, ovNode, ovT);
            if (h.slot == kNoHit) {
                handle_miss(S, st, depth, ps, c);
                ended = true;
            } else {
                Surface sf;
                TexState ts;
                if (shade_hit<TEX>(S, depth, smp, ray, h, ps, sf, &ts)) {
                    const Mat& M = sf.M;
                    DirectSum ds;
                    ds.init(st);
                    const V3 origin = sf.p + sf.n * kRayEps;
                    for (uint32_t j = 0; j < nLight; ++j) {
                        V3 ld, contrib;
                        uint32_t group;
                        float maxDist;
                        if (!direct_sample(S, st, sf, M, smp, depth, j, &ld, &contrib, &group, &maxDist, c)) continue;
                        Ray sr;
                        sr.o = origin;
                        sr.d = ld;
                        sr.tmin = kRayEps;
                        sr.tmax = maxDist;
                        ds.add(group, contrib, !intersect_any_ring(S, sr, lane, ovNode));
                    }
                    ds.flush();
                    ps.L = ps.L + ps.T * ds.L;
                    if (!sample_bounce(S, st, depth, smp, sf, ps)) ended = true;
                }
            }
            ++depth;
            if (depth >= st.maxDepth) ended = true;
        }
        if (!ended) continue;
        // The sample is done: add it in order, then the next sample.
        beauty.x = beauty.x + ps.L.x;
        beauty.y = beauty.y + ps.L.y;
        beauty.z = beauty.z + ps.L.z;
        beauty.w = beauty.w + (ps.depth < kInf ? 1.0f : 0.0f);
        albedo.x = albedo.x + ps.albedo.x;
        albedo.y = albedo.y + ps.albedo.y;
        albedo.z = albedo.z + ps.albedo.z;
        normal.x = normal.x + ps.normal.x;
        normal.y = normal.y + ps.normal.y;
        normal.z = normal.z + ps.normal.z;
        if (ps.depth < dmin) dmin = ps.depth;
        if (++i < P.sliceCount) {
            depth = 0;
            ps = init_path(P.camera, st, x, y, Sampler{x, y, P.sampleStart + P.sliceBegin + i});
            continue;
        }
        store_pixel(P, pix, beauty, albedo, normal, dmin);
        active = false;
    }
}

// Batched multi-scene megakernel: persistent one-wave blocks take work
// units (batch item, tile) from a list sorted by decreasing estimated cost
// with one atomic per unit, so large units start first and the tail is
// short.  Each unit is exactly the k_render tile of its item (all samples of
// the batch in one pass), so the frames equal single-scene renders bit for
// bit.
template <bool TEX>
__global__ __launch_bounds__(64) void k_render_many(const RenderParams* __restrict__ items,
                                                    const uint2* __restrict__ work, uint32_t workCount,
                                                    uint32_t* counter) {
    for (;;) {
        uint32_t w = 0u;
        if (threadIdx.x == 0u) w = atomicAdd(counter, 1u);
        w = __shfl(w, 0);
        if (w >= workCount) return;
        const uint2 u = work[w];
        render_tile<false, TEX>(items[u.x], u.y, threadIdx.x);
    }
}

// VALU issue-rate probe: eight independent FMA chains per thread, so that
// a wave always has an FMA ready and the SIMDs issue VALU instructions
// back to back.  kValuProbeChains FMAs per iteration.
constexpr uint32_t kValuProbeChains = 8;
__global__ __launch_bounds__(256) void k_valu_probe(float* out, uint32_t iters, float seed) {
    float a[kValuProbeChains];
#pragma unroll
    for (uint32_t k = 0; k < kValuProbeChains; ++k) a[k] = seed + static_cast<float>(threadIdx.x + k);
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (uint32_t k = 0; k < kValuProbeChains; ++k) a[k] = __builtin_fmaf(a[k], 0.999999f, 1.0e-7f);
    }
    float s = 0.0f;
#pragma unroll
    for (uint32_t k = 0; k < kValuProbeChains; ++k) s += a[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

hipError_t launch_valu_probe(float* out, uint32_t blocks, uint32_t iters, hipStream_t s) {
    hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, s, out, iters, 1.0f);
    return hipGetLastError();
}

__global__ void k_clear_u32(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// FrameAccumulator::Accumulate (Renderer.cpp:60-83).
__global__ void k_accum_add(AccumParams A, DevFrame f, uint32_t count, uint32_t np) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float w = static_cast<float>(count);
#pragma unroll
    for (int c = 0; c < 4; ++c) A.beauty[p * 4 + c] = A.beauty[p * 4 + c] + f.beauty[p * 4 + c] * w;
#pragma unroll
    for (int c = 0; c < 3; ++c) A.albedo[p * 3 + c] = A.albedo[p * 3 + c] + f.albedo[p * 3 + c] * w;
#pragma unroll
    for (int c = 0; c < 3; ++c) A.normal[p * 3 + c] = A.normal[p * 3 + c] + f.normal[p * 3 + c] * w;
    const float d = f.depth[p];
    if (d < kInf) {
        A.depthSum[p] += d * w;
        A.depthCount[p] += count;
    }
}

// FrameAccumulator::Resolve (Renderer.cpp:85-110).
__global__ void k_accum_resolve(AccumParams A, DevFrame f, uint32_t total, uint32_t np) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    if (total == 0u) {
        for (int c = 0; c < 4; ++c) f.beauty[p * 4 + c] = 0.0f;
        for (int c = 0; c < 3; ++c) f.albedo[p * 3 + c] = 0.0f;
        for (int c = 0; c < 3; ++c) f.normal[p * 3 + c] = 0.0f;
        f.depth[p] = kInf;
        return;
    }
    const float inv = 1.0f / static_cast<float>(total);
#pragma unroll
    for (int c = 0; c < 4; ++c) f.beauty[p * 4 + c] = A.beauty[p * 4 + c] * inv;
#pragma unroll
    for (int c = 0; c < 3; ++c) f.albedo[p * 3 + c] = A.albedo[p * 3 + c] * inv;
#pragma unroll
    for (int c = 0; c < 3; ++c) f.normal[p * 3 + c] = A.normal[p * 3 + c] * inv;
    f.depth[p] = A.depthCount[p] > 0u ? A.depthSum[p] / static_cast<float>(A.depthCount[p]) : kInf;
}

__global__ void k_qmc_fill(uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t s, uint32_t dim0,
                           uint32_t dimCount, float* out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t total = static_cast<uint64_t>(w) * h * dimCount;
    if (i >= total) return;
    const uint32_t d = static_cast<uint32_t>(i % dimCount);
    const uint64_t p = i / dimCount;
    const uint32_t x = static_cast<uint32_t>(p % w), y = static_cast<uint32_t>(p / w);
    out[i] = qmc_value(x0 + x, y0 + y, s, dim0 + d);
}

__global__ void k_intersect(DevScene S, const float* rays, uint32_t n, uint32_t* prim, float* t, float* uv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = rays + static_cast<size_t>(i) * 8u;
    Ray r;
    r.o = V3{q[0], q[1], q[2]};
    r.d = V3{q[3], q[4], q[5]};
    r.tmin = q[6];
    r.tmax = q[7];
    Counters<false> c;
    const HitRec h = intersect_closest(S, r, c.trav());
    prim[i] = h.prim;
    t[i] = h.slot == kNoHit ? kFltMax : h.t;
    uv[2 * i] = h.b1;
    uv[2 * i + 1] = h.b2;
}

__global__ void k_occluded(DevScene S, const float* rays, uint32_t n, uint8_t* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = rays + static_cast<size_t>(i) * 8u;
    Ray r;
    r.o = V3{q[0], q[1], q[2]};
    r.d = V3{q[3], q[4], q[5]};
    r.tmin = q[6];
    r.tmax = q[7];
    Counters<false> c;
    out[i] = intersect_any(S, r, c.trav()) ? 1u : 0u;
}

__global__ __launch_bounds__(64) void k_trace_primary(DevScene S, DevCamera cam, DevSettings st, uint32_t s,
                                                      uint32_t* prim, float* t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= st.width * st.height) return;
    const uint32_t x = i % st.width, y = i / st.width;
    const Sampler smp{x, y, s};
    const Ray r = camera_ray(cam, st, x, y, smp.next2(0u));
    Counters<false> c;
    const HitRec h = intersect_closest(S, r, c.trav());
    prim[i] = h.prim;
    t[i] = h.slot == kNoHit ? kFltMax : h.t;
}

__global__ void k_eval_bsdf(DevScene S, const uint32_t* mi, const float* n, const float* cn, const float* v,
                            const float* l, uint32_t count, float* value, float* pdf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t m = mi[i];
    const Mat M{S.materials + m, S.derived + m, S.derived + m};
    const V3 N = ld3(n + 3 * i), CN = ld3(cn + 3 * i), Vv = ld3(v + 3 * i), L = ld3(l + 3 * i);
    const V3 f = bsdf_eval(M, N, CN, Vv, L);
    value[3 * i] = f.x;
    value[3 * i + 1] = f.y;
    value[3 * i + 2] = f.z;
    pdf[i] = bsdf_pdf(M, N, CN, Vv, L);
}

__global__ void k_env_sample(DevEnv e, const float* u, uint32_t n, float* dir, float* pdf, float* rgb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 d;
    float p;
    unsigned long long a = 0, b = 0;
    const V3 c = env_sample(e, V2{u[2 * i], u[2 * i + 1]}, &d, &p, EnvCount<false>{&a, &b});
    dir[3 * i] = d.x;
    dir[3 * i + 1] = d.y;
    dir[3 * i + 2] = d.z;
    pdf[i] = p;
    rgb[3 * i] = c.x;
    rgb[3 * i + 1] = c.y;
    rgb[3 * i + 2] = c.z;
}

__global__ void k_env_eval(DevEnv e, const float* dir, uint32_t n, float* pdf, float* rgb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long a = 0, b = 0;
    const EnvCount<false> ec{&a, &b};
    const V3 d = ld3(dir + 3 * i);
    pdf[i] = env_pdf(e, d, ec);
    const V3 c = env_eval(e, d, ec);
    rgb[3 * i] = c.x;
    rgb[3 * i + 1] = c.y;
    rgb[3 * i + 2] = c.z;
}

static inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return static_cast<uint32_t>((n + bs - 1) / bs); }

hipError_t launch_render(const RenderParams& P, bool count, hipStream_t stream) {
    const uint32_t tilesY = (P.settings.height + 7u) / 8u;
    const uint32_t nb = P.tilesX * tilesY;
    if (nb == 0) return hipSuccess;
    const bool tex = P.scene.textured != 0u;
    if (count) {
        if (tex) {
            hipLaunchKernelGGL((k_render<true, true>), dim3(nb), dim3(64), 0, stream, P);
        } else {
            hipLaunchKernelGGL((k_render<true, false>), dim3(nb), dim3(64), 0, stream, P);
        }
    } else {
        if (tex) {
            hipLaunchKernelGGL((k_render<false, true>), dim3(nb), dim3(64), 0, stream, P);
        } else {
            hipLaunchKernelGGL((k_render<false, false>), dim3(nb), dim3(64), 0, stream, P);
        }
    }
    return hipGetLastError();
}

hipError_t launch_render_plus(const RenderParams& P, uint32_t* counter, uint32_t blocks, hipStream_t stream) {
    const uint32_t tilesY = (P.settings.height + 7u) / 8u;
    if (P.tilesX * tilesY == 0 || blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, stream, counter, 1u);
    if (P.scene.textured != 0u) {
        hipLaunchKernelGGL((k_render_plus<true>), dim3(blocks), dim3(64), 0, stream, P, counter);
    } else {
        hipLaunchKernelGGL((k_render_plus<false>), dim3(blocks), dim3(64), 0, stream, P, counter);
    }
    return hipGetLastError();
}

hipError_t launch_render_many(const RenderParams* items, const uint2* work, uint32_t workCount, uint32_t* counter,
                              uint32_t blocks, bool tex, hipStream_t s) {
    if (workCount == 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(64), 0, s, counter, 1u);
    if (tex) {
        hipLaunchKernelGGL(k_render_many<true>, dim3(blocks), dim3(64), 0, s, items, work, workCount, counter);
    } else {
        hipLaunchKernelGGL(k_render_many<false>, dim3(blocks), dim3(64), 0, s, items, work, workCount, counter);
    }
    return hipGetLastError();
}

hipError_t launch_accum_add(const AccumParams& A, const DevFrame& f, uint32_t count, uint32_t np, hipStream_t s) {
    if (np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_accum_add, dim3(blocks_for(np, 256)), dim3(256), 0, s, A, f, count, np);
    return hipGetLastError();
}

hipError_t launch_accum_resolve(const AccumParams& A, const DevFrame& f, uint32_t total, uint32_t np, hipStream_t s) {
    if (np == 0) return hipSuccess;
    hipLaunchKernelGGL(k_accum_resolve, dim3(blocks_for(np, 256)), dim3(256), 0, s, A, f, total, np);
    return hipGetLastError();
}

hipError_t launch_qmc_fill(uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t s, uint32_t dim0,
                           uint32_t dimCount, float* out, hipStream_t st) {
    const uint64_t n = static_cast<uint64_t>(w) * h * dimCount;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_qmc_fill, dim3(blocks_for(n, 256)), dim3(256), 0, st, x0, y0, w, h, s, dim0, dimCount, out);
    return hipGetLastError();
}

hipError_t launch_intersect(const DevScene& S, const float* rays, uint32_t n, uint32_t* prim, float* t, float* uv,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_intersect, dim3(blocks_for(n, 64)), dim3(64), 0, s, S, rays, n, prim, t, uv);
    return hipGetLastError();
}

hipError_t launch_occluded(const DevScene& S, const float* rays, uint32_t n, uint8_t* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_occluded, dim3(blocks_for(n, 64)), dim3(64), 0, s, S, rays, n, out);
    return hipGetLastError();
}

hipError_t launch_trace_primary(const DevScene& S, const DevCamera& cam, const DevSettings& st, uint32_t sample,
                                uint32_t* prim, float* t, hipStream_t s) {
    const uint64_t n = static_cast<uint64_t>(st.width) * st.height;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_primary, dim3(blocks_for(n, 64)), dim3(64), 0, s, S, cam, st, sample, prim, t);
    return hipGetLastError();
}

hipError_t launch_eval_bsdf(const DevScene& S, const uint32_t* mi, const float* n, const float* cn, const float* v,
                            const float* l, uint32_t count, float* value, float* pdf, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_eval_bsdf, dim3(blocks_for(count, 64)), dim3(64), 0, s, S, mi, n, cn, v, l, count, value,
                       pdf);
    return hipGetLastError();
}

hipError_t launch_env_sample(const DevEnv& e, const float* u, uint32_t n, float* dir, float* pdf, float* rgb,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_env_sample, dim3(blocks_for(n, 64)), dim3(64), 0, s, e, u, n, dir, pdf, rgb);
    return hipGetLastError();
}

hipError_t launch_env_eval(const DevEnv& e, const float* dir, uint32_t n, float* pdf, float* rgb, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_env_eval, dim3(blocks_for(n, 64)), dim3(64), 0, s, e, dir, n, pdf, rgb);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Dome maps updated on the device (texture extension): RGB texels into the
// float4 layout, and the importance tables rebuilt from them with the host
// build's operation order (build_env_tables, build_cdf_guide), so that a map
// updated on the device samples exactly like the same map uploaded.

__global__ void k_env_texels(const float* rgb, float4* out, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = make_float4(rgb[3u * i], rgb[3u * i + 1u], rgb[3u * i + 2u], 0.0f);
}

// One thread per row: sequential prefix sum of max(0, luminance) * sin, then
// normalised (uniform where the row is black); the row's total to rowW.
__global__ void k_env_cond(const float4* tex, uint32_t w, uint32_t h, const float* rowSin, float* cond, float* rowW) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= h) return;
    const float st = rowSin[y];
    float* row = cond + static_cast<size_t>(y) * (w + 1u);
    row[0] = 0.0f;
    float sum = 0.0f;
    for (uint32_t x = 0; x < w; ++x) {
        const float4 c = tex[static_cast<size_t>(y) * w + x];
        const float lum = c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
        sum += fmaxs(0.0f, lum) * st;
        row[x + 1u] = sum;
    }
    if (sum > 0.0f) {
        for (uint32_t x = 1; x <= w; ++x) row[x] /= sum;
    } else {
        for (uint32_t x = 1; x <= w; ++x) row[x] = static_cast<float>(x) / static_cast<float>(w);
        sum = 1.0f;
    }
    rowW[y] = sum;
}

// One thread: the row CDF from the row totals, in row order.
__global__ void k_env_rows(const float* rowW, uint32_t h, float* rowCdf) {
    if (blockIdx.x != 0u || threadIdx.x != 0u) return;
    float total = 0.0f;
    rowCdf[0] = 0.0f;
    for (uint32_t y = 0; y < h; ++y) {
        total += rowW[y];
        rowCdf[y + 1u] = total;
    }
    if (total <= 0.0f) {
        for (uint32_t y = 1; y <= h; ++y) rowCdf[y] = static_cast<float>(y) / static_cast<float>(h);
        return;
    }
    for (uint32_t y = 1; y <= h; ++y) rowCdf[y] /= total;
}

// build_cdf_guide on the device: block y < h the guide of row y's
// conditional CDF, block h the row CDF's; the block's threads check the
// table's monotonicity together, then take one guide position each.
__global__ __launch_bounds__(256) void k_env_guides(const float* rowCdf, const float* cond, uint32_t w, uint32_t h,
                                                    uint32_t* rowGuide, uint32_t* condGuide) {
    const uint32_t y = blockIdx.x;
    const bool row = y == h;
    const float* cdf = row ? rowCdf : cond;
    const uint32_t begin = row ? 0u : y * (w + 1u);
    const uint32_t end = row ? h : begin + w;
    const uint32_t G = row ? kEnvRowGuide : kEnvColGuide;
    uint32_t* out = row ? rowGuide : condGuide + static_cast<size_t>(y) * (kEnvColGuide + 1u);
    int ok = 1;
    for (uint32_t i = begin + threadIdx.x; i < end; i += blockDim.x)
        if (!(cdf[i] <= cdf[i + 1u])) ok = 0;
    const bool monotone = __syncthreads_and(ok) != 0;
    for (uint32_t g = threadIdx.x; g <= G; g += blockDim.x) {
        if (!monotone) {
            out[g] = g == G ? end + 1u : begin + 1u;
            continue;
        }
        const float key = static_cast<float>(g) / static_cast<float>(G);
        uint32_t first = begin + 1u, count = end - begin;
        while (count > 0u) {
            const uint32_t half = count >> 1;
            const uint32_t mid = first + half;
            if (cdf[mid] < key) {
                first = mid + 1u;
                count -= half + 1u;
            } else {
                count = half;
            }
        }
        out[g] = first;
    }
}

hipError_t launch_env_update(const EnvUpdate& U, hipStream_t s) {
    const uint32_t n = U.w * U.h;
    hipLaunchKernelGGL(k_env_texels, dim3(blocks_for(n, 256)), dim3(256), 0, s, U.rgb, U.texels, n);
    if (!U.rebuild) return hipGetLastError();
    hipLaunchKernelGGL(k_env_cond, dim3(blocks_for(U.h, 64)), dim3(64), 0, s, U.texels, U.w, U.h, U.rowSin, U.cond,
                       U.rowW);
    hipLaunchKernelGGL(k_env_rows, dim3(1), dim3(64), 0, s, U.rowW, U.h, U.rowCdf);
    hipLaunchKernelGGL(k_env_guides, dim3(U.h + 1u), dim3(256), 0, s, U.rowCdf, U.cond, U.w, U.h, U.rowGuide,
                       U.condGuide);
    return hipGetLastError();
}

}  // namespace shiro_hip
