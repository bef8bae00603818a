// Wavefront transport: the path integrator split into stage kernels over
// path state kept in HBM (structure of arrays), with ray queues compacted by
// wave ballots and block-wide prefix sums.  Per pass (a slice of samples for
// every pixel):
//
//   k_wf_reset     clear the queue counters of the pass
//   per depth d:
//     k_wf_trace<closest>  closest hit of every queued ray; at depth 0 every
//                  (pixel, sample) path of the pass, its camera ray computed
//                  in place (no queue, no stored ray), or k_wf_primary,
//                  the same camera rays traced as 64-ray tile packets
//     k_wf_classify, k_wf_order (depth > 0, optional)  counting sort of
//                  the hit queue by shading class (misses, and materials
//                  that take the same shading branches), so that the lanes
//                  of a shade wave follow the same code path
//     k_wf_shade   surface terms, light samples -> shadow rays, BSDF
//                  sampling and roulette -> next ray queue
//     k_wf_trace<any>  any-hit test of every queued shadow ray
//     k_wf_resolve visible contributions summed per path in DirectLighting's
//                  order and added to the path radiance
//
// The traversal kernels are persistent: a fixed grid whose lanes fetch rays
// from per-wave pools of the queue (one atomic per kPoolChunk rays), and a
// lane whose ray is done fetches the next one at once instead of idling
// until the longest ray of its wave ends (dynamic ray fetch, Aila and Laine
// 2009).
//   k_wf_accum     per-pixel sums over the pass's samples, in sample order
//
// Every stage uses the same device functions as the megakernel
// (dev_path.h), shadow contributions are applied in the reference's light
// order before the next bounce, and pixels accumulate samples in order, so
// both transports produce the same per-sample radiance up to the float
// rounding of identical operations.  Grids are fixed (grid-stride loops read
// the queue lengths from device memory), so no host round trip is needed
// between stages.
#include <hip/hip_runtime.h>

#include <mutex>

#include "dev_path.h"
#include "dev_ring.h"
#include "kernels.h"

namespace shiro_hip {

DEV uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of distinct values of `id` over the active lanes, returned to the
// lowest active lane (0 to the others).
DEV uint32_t wave_distinct(uint32_t id) {
    const unsigned long long active = __ballot(true);
    unsigned long long m = active;
    uint32_t n = 0;
    while (m != 0ull) {
        const int l = __ffsll(static_cast<long long>(m)) - 1;
        const uint32_t v = __shfl(id, l);
        m &= ~__ballot(id == v);
        ++n;
    }
    return static_cast<int>(lane_id()) == __ffsll(static_cast<long long>(active)) - 1 ? n : 0u;
}

// Fetch counters of the measurement instantiations of the traversal kernels
// (ON = true, results in WfParams::wfCount): per lane, the nodes and
// triangles each ray's traversal reads; per wave, the distinct nodes and
// triangles of each load instruction (the lanes of one instruction that
// read the same record need it once: every node of a camera packet, and
// the rays of one path group that walk the same nodes in step).
template <bool ON>
struct WaveCount {
    uint32_t v[kWfCountFields] = {};
    DEV void ray() {
        if (ON) ++v[kWfCountRays];
    }
    DEV void node(uint32_t id) {
        if (ON) {
            ++v[kWfCountNodes];
            v[kWfCountWaveNodes] += wave_distinct(id);
        }
    }
    DEV void tri(uint32_t slot) {
        if (ON) {
            ++v[kWfCountTris];
            v[kWfCountWaveTris] += wave_distinct(slot);
        }
    }
    DEV void add(int field, bool c) {
        if (ON && c) ++v[field];
    }
    // Sums the wave's counters into out[kind * kWfCountFields + field];
    // called by every lane of the wave.
    DEV void flush(unsigned long long* out, uint32_t kind) {
        if (!ON || !out) return;
        for (int f = 0; f < kWfCountFields; ++f) {
            uint32_t x = v[f];
            for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d);
            if (lane_id() == 0u && x) atomicAdd(out + kind * kWfCountFields + f, static_cast<unsigned long long>(x));
        }
    }
};

struct PathCoord {
    uint32_t x, y, s;
};

// Path layout of a pass: path p = 64 * w + lane, where group w holds G =
// 2^groupLog consecutive samples of an R x R pixel square of an 8x8 tile,
// R = 8 / sqrt(G) (G = 1: the whole tile at one sample; 4: a 4x4 quarter at
// four samples; 16: a 2x2 square at sixteen; 64: one pixel at sixty-four).
// Group w covers samples (w / (tiles * G)) * G + [0, G), tile
// (w / G) % tiles and square w % G (squares row-major in the tile); lane =
// (sample in group) * R * R + (pixel in square, row-major).  The samples of
// one pixel differ only by their jitter, so a wave of consecutive paths
// holds nearly identical camera rays and surfaces, and queues filled in
// path order keep them together for the later stages.  sliceCount % G == 0.
DEV PathCoord path_coord(const WfParams& P, uint32_t p) {
    const uint32_t gl = P.groupLog, rl = 3u - (gl >> 1u), ql = 3u - rl;
    const uint32_t w = p >> 6, lane = p & 63u;
    const uint32_t perGroup = P.tiles << gl;
    const uint32_t sg = w / perGroup;
    const uint32_t rem = w - sg * perGroup;
    const uint32_t t = rem >> gl, q = rem & ((1u << gl) - 1u);
    const uint32_t sub = lane >> (2u * rl), j = lane & ((1u << (2u * rl)) - 1u);
    const uint32_t xin = ((q & ((1u << ql) - 1u)) << rl) + (j & ((1u << rl) - 1u));
    const uint32_t yin = ((q >> ql) << rl) + (j >> rl);
    const uint32_t tg = P.tileBase + t;
    return PathCoord{(tg % P.tilesX) * 8u + xin, (tg / P.tilesX) * 8u + yin, (sg << gl) + sub};
}

// Inverse of path_coord: the path of sample s of pixel (xin, yin) of the
// pass's tile t.
DEV uint32_t path_index(const WfParams& P, uint32_t s, uint32_t t, uint32_t xin, uint32_t yin) {
    const uint32_t gl = P.groupLog, rl = 3u - (gl >> 1u), ql = 3u - rl;
    const uint32_t sg = s >> gl, sub = s & ((1u << gl) - 1u);
    const uint32_t q = ((yin >> rl) << ql) + (xin >> rl);
    const uint32_t j = ((yin & ((1u << rl) - 1u)) << rl) + (xin & ((1u << rl) - 1u));
    const uint32_t w = sg * (P.tiles << gl) + (t << gl) + q;
    return (w << 6) + (sub << (2u * rl)) + j;
}

DEV Sampler path_sampler(const WfParams& P, const PathCoord& c) {
    return Sampler{c.x, c.y, P.sampleStart + P.sliceBegin + c.s};
}

// Loads what a bounce reads: ray, throughput, radiance, flags; the medium
// extinction only inside a medium.  First-hit AOVs are written once at depth
// 0 and never read back here.
DEV PathState load_state(const WfBuffers& B, uint32_t p) {
    PathState ps;
    const size_t r = static_cast<size_t>(p) * kPathRecord;
    const float4 o = B.o[r], d = B.d[r], T = B.T[r], L = B.L[p];
    ps.o = xyz(o);
    ps.prevPdf = o.w;
    ps.d = xyz(d);
    ps.bits = __float_as_uint(d.w);
    ps.T = xyz(T);
    ps.L = xyz(L);
    ps.ext = V3{0.0f, 0.0f, 0.0f};
    if (ps.bits & kBitsMediumOn) ps.ext = xyz(B.E[r]);
    ps.albedo = ps.normal = V3{0.0f, 0.0f, 0.0f};
    ps.depth = kInf;
    return ps;
}

// Stores the state after a bounce.  `alive` paths keep their ray and
// throughput; every path keeps its radiance; the AOVs are stored at depth 0;
// the extinction only when it was (re)set by entering a medium.
DEV void store_state(const WfBuffers& B, uint32_t p, const PathState& ps, uint32_t depth, bool alive,
                     bool extChanged) {
    const size_t r = static_cast<size_t>(p) * kPathRecord;
    if (alive) {
        B.o[r] = make_float4(ps.o.x, ps.o.y, ps.o.z, ps.prevPdf);
        B.d[r] = make_float4(ps.d.x, ps.d.y, ps.d.z, __uint_as_float(ps.bits));
        B.T[r] = make_float4(ps.T.x, ps.T.y, ps.T.z, 0.0f);
        if (extChanged) B.E[r] = make_float4(ps.ext.x, ps.ext.y, ps.ext.z, 0.0f);
    }
    B.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
    if (depth == 0u) {
        B.A[p] = make_float4(ps.albedo.x, ps.albedo.y, ps.albedo.z, ps.depth);
        B.N[p] = make_float4(ps.normal.x, ps.normal.y, ps.normal.z, 0.0f);
    }
}

__global__ void k_wf_reset(uint32_t* counts, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) counts[i] = 0u;
}

DEV bool path_valid(const WfParams& P, const PathCoord& c) {
    return c.x < P.settings.width && c.y < P.settings.height;
}

// Number of work items of stage `depth`: every path of the pass at depth 0,
// the length of the ray queue afterwards.
DEV uint32_t stage_items(const WfParams& P, uint32_t depth) {
    return depth == 0u ? P.sliceCount * P.tiles * 64u : P.B.counts[depth];
}

// Work distribution of the persistent kernels: each wave owns a pool of
// queue positions [next, end), a chunk of 64, so lanes refill from the pool
// without a global atomic most of the time.  Chunks come from kPoolStripes
// counters (kPoolStride words apart, one 128-byte line each), counter s
// handing out the chunks of the s-th of kPoolStripes contiguous ranges of
// the queue; a wave starts at stripe blockIdx % kPoolStripes and moves to
// the next stripe when its own is used up.  With one counter for the whole
// grid, its atomics serialised on its address (about 88M per second were
// measured): 64-position chunks from one counter took c2's closest-hit /
// any-hit stages 33.4 / 20.3 ms, 256-position chunks 29.4 / 14.5 but slowed
// c3 (45.8 -> 49.7 ms closest hit: a wave's lanes then hold rays from a
// wider span of the queue, which is less coherent).
struct WavePool {
    uint32_t next = 0, end = 0;
    uint32_t stripe = 0;   // wave-uniform
    uint32_t tries = 0;    // stripes found used up (wave-uniform)
    // Hands the wanting lanes consecutive positions; positions >= n mean the
    // queue is exhausted.  `counters` holds the kPoolStripes counters.
    DEV uint32_t fetch(bool want, uint32_t* counters, uint32_t n) {
        const unsigned long long mask = __ballot(want);
        if (mask == 0ull) return 0xffffffffu;
        const uint32_t lane = lane_id();
        const uint32_t nWant = static_cast<uint32_t>(__popcll(mask));
        const uint32_t avail = end - next;
        uint32_t chunk = 0;
        if (nWant > avail) {
            const int leader = __ffsll(static_cast<long long>(mask)) - 1;
            const uint32_t chunks = (n + 63u) / 64u;
            if (static_cast<int>(lane) == leader) {
                // Chunks [chunks * s / S, chunks * (s + 1) / S) belong to stripe s.
                chunk = chunks * 64u;  // past the queue: exhausted
                while (tries < kPoolStripes) {
                    const uint32_t lo = static_cast<uint32_t>(static_cast<uint64_t>(chunks) * stripe / kPoolStripes);
                    const uint32_t hi =
                        static_cast<uint32_t>(static_cast<uint64_t>(chunks) * (stripe + 1u) / kPoolStripes);
                    const uint32_t c = lo < hi ? atomicAdd(counters + stripe * kPoolStride, 1u) : hi - lo;
                    if (c < hi - lo) {
                        chunk = (lo + c) * 64u;
                        break;
                    }
                    stripe = stripe + 1u == kPoolStripes ? 0u : stripe + 1u;
                    ++tries;
                }
            }
            chunk = __shfl(chunk, leader);
            stripe = __shfl(stripe, leader);
            tries = __shfl(tries, leader);
        }
        const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
        const uint32_t idx = rank < avail ? next + rank : chunk + (rank - avail);
        if (nWant > avail) {
            next = chunk + (nWant - avail);
            end = chunk + 64u;
        } else {
            next += nWant;
        }
        return idx;
    }
};

// Persistent traversal.  ANY = false: closest hit of ray-queue entries (all
// paths at depth 0), result in H.  ANY = true: any hit of shadow-queue
// entries, result in svis.  The traversal itself is the one of dev_bvh.h
// (same node test, child order, tie rule and triangle test) reorganised so
// that each lane can restart with a new ray.
// FETCH: 0 = wave-synchronous rounds (a wave takes new rays only when all
// its lanes are done), 1 = lanes refill as their rays finish (static item
// order), 2 = lanes refill from per-wave pools of the global queue.
// The node phase ends when at most P.leafWait (closest hit) or
// P.shadowLeafWait (any hit) lanes of the wave are still searching for a
// leaf (closest hit, measured: 24 of 64 is 5% faster than waiting for every
// lane on c3, 12-32 are within 1%).
// COUNT: the measurement instantiation (fetch counters in P.wfCount).
// W: tree width (4: `nodes`, 8: `nodes8`).  At most 64 VGPRs for the
// 4-wide tree, so that eight waves fit on each SIMD.
template <bool ANY, int FETCH, int W, bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W == 8 ? (ANY ? 6 : 4) : 8))) void k_wf_trace(WfParams P, uint32_t depth) {
    const uint32_t md = P.settings.maxDepth;
    const uint32_t n = ANY ? P.B.counts[P.B.shadowRayBase + depth] : stage_items(P, depth);
    uint32_t* counter = &P.B.counts[P.B.fetchBase + ((ANY ? md : 0u) + depth) * kPoolStripes * kPoolStride];
    const uint32_t* in = (depth & 1u) ? P.B.qb : P.B.qa;
    const DevScene& S = P.scene;
    if (S.nodeCount == 0) {
        // Empty scene: every ray misses.
        for (uint32_t i = blockIdx.x * 64u + threadIdx.x; i < n; i += gridDim.x * 64u) {
            if (ANY) {
                P.B.svis[P.B.qsr[i]] = 1u;
            } else {
                const uint32_t p = depth == 0u ? i : in[i];
                P.B.H[p] = make_float4(kFltMax, 0.0f, 0.0f, __uint_as_float(kNoHit));
            }
        }
        return;
    }
    uint32_t ovNode[kStackSize];
    float ovT[ANY ? 1 : kStackSize];
    int sp = 0, base = 0;
    const uint32_t lane = threadIdx.x;
    uint32_t cur = kEmpty;
    uint32_t item = 0;
    bool active = false, exhausted = false;
    RayPre r;
    float tmax = kFltMax;
    HitRec best;
    best.t = kFltMax;
    best.b1 = best.b2 = 0.0f;
    best.slot = best.prim = kNoHit;
    WaveCount<COUNT> cnt;
    const uint32_t leafWait = ANY ? P.shadowLeafWait : P.leafWait;
    WavePool pool;
    pool.stripe = blockIdx.x % kPoolStripes;
    // Static assignment: lane i of the grid takes items i, i + G, i + 2G...
    uint32_t nextStatic = blockIdx.x * 64u + threadIdx.x;
    for (;;) {
        const bool want = !active && !exhausted && (FETCH != 0 || !__any(active));
        uint32_t idx;
        if (FETCH == 2) {
            idx = pool.fetch(want, counter, n);
        } else {
            idx = nextStatic;
            if (want) nextStatic += gridDim.x * 64u;
        }
        if (want) {
            if (idx >= n) {
                exhausted = true;
            } else {
                item = idx;
                Ray ray;
                bool valid = true;
                float rayMax = kFltMax;
                if (ANY) {
                    const uint32_t rid = P.B.qsr[idx];
                    const uint32_t p = rid / P.B.maxShadow;
                    ray.o = xyz(P.B.so[p]);
                    const float4 dm = P.B.sr[2u * static_cast<size_t>(rid)];
                    ray.d = xyz(dm);
                    rayMax = dm.w;
                    item = rid;
                } else if (depth == 0u) {
                    const PathCoord c = path_coord(P, idx);
                    valid = path_valid(P, c);
                    if (valid) ray = camera_ray(P.camera, P.settings, c.x, c.y, path_sampler(P, c).next2(0u));
                } else {
                    item = in[idx];
                    ray.o = xyz(P.B.o[static_cast<size_t>(item) * kPathRecord]);
                    ray.d = xyz(P.B.d[static_cast<size_t>(item) * kPathRecord]);
                }
                if (valid) {
                    ray.tmin = kRayEps;
                    ray.tmax = rayMax;
                    r = precompute(ray);
                    tmax = rayMax;
                    best.t = kFltMax;
                    best.b1 = best.b2 = 0.0f;
                    best.slot = best.prim = kNoHit;
                    sp = base = 0;
                    cur = 0u;
                    active = true;
                    cnt.ray();
                }
            }
        }
        // A lane is exhausted once the queue is; an inactive lane that is not
        // exhausted (it fetched an off-image path) fetches again.
        if (__all(exhausted)) break;
        if (!active) continue;
        bool occluded = false;
        // Node phase with postponed leaves: the first leaf a lane reaches is
        // held and the lane keeps descending (from the next stack entry)
        // until at most kLeafWaitLanes lanes are still looking for a leaf;
        // those carry on in the next round instead of holding up the leaf
        // tests of the others.
        uint32_t leaf = kEmpty;
        for (;;) {
            if (cur != kEmpty && !(cur & kLeafBit)) {
                const float4* nd = node_at<W>(S, cur);
                cnt.node(cur);
                float tn[W];
                uint32_t w[W];
                const uint32_t mask = node_test<W>(nd, r, ANY ? tmax : best.t, tn, w);
                if (mask) {
                    cur = ANY ? ring_push_in_order<W>(mask, w, lane, ovNode, sp, base)
                              : ring_push_sorted<true, W>(mask, tn, w, lane, ovNode, ovT, sp, base);
                } else {
                    cur = ring_pop<ANY>(best.t, lane, ovNode, ovT, sp, base);
                }
            }
            if (leaf == kEmpty && cur != kEmpty && (cur & kLeafBit)) {
                leaf = cur;
                cur = ring_pop<ANY>(best.t, lane, ovNode, ovT, sp, base);
            }
            if (__popcll(__ballot(leaf == kEmpty && cur != kEmpty)) <= leafWait) break;
        }
        // Leaf phase.
        if (leaf != kEmpty) {
            if (ANY) {
                uint32_t slot = leaf & ~kLeafBit;
                for (;;) {
                    const float4* tp = S.tris + static_cast<size_t>(slot) * 3u;
                    const float4 a = tp[0], b = tp[1], c = tp[2];
                    cnt.tri(slot);
                    float t, b1, b2;
                    if (tri_test(r, xyz(a), xyz(b), xyz(c), tmax, &t, &b1, &b2)) {
                        occluded = true;
                        break;
                    }
                    if (__float_as_uint(b.w) & kLeafEndFlag) break;
                    ++slot;
                }
            } else {
                leaf_closest(S, r, leaf & ~kLeafBit, best, cnt);
            }
        }
        // The next entry was popped when the leaf was postponed; the stack is
        // empty when it is kEmpty.
        const bool finished = occluded || cur == kEmpty;
        if (finished) {
            if (ANY) {
                P.B.svis[item] = occluded ? 0u : 1u;
                cnt.add(kWfCountOccluded, occluded);
            } else {
                P.B.H[item] = make_float4(best.t, best.b1, best.b2, __uint_as_float(best.slot));
            }
            active = false;
        }
    }
    cnt.flush(P.wfCount, ANY ? 2u : (depth == 0u ? 0u : 1u));
}

// Closest hit of the pass's camera rays as packets
// (intersect_closest_packet, dev_bvh.h): one wave per group of 64
// consecutive paths (path_coord: G samples of an 8/sqrt(G)-pixel square);
// every lane tests each node the packet visits against its own ray and
// keeps its own closest hit.
template <int W, bool COUNT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W == 8 ? 6 : 8))) void k_wf_primary(WfParams P) {
    __shared__ uint32_t wstack[kStackSize];
    WaveCount<COUNT> cnt;
    const uint32_t packets = P.sliceCount * P.tiles;
    for (uint32_t w = blockIdx.x; w < packets; w += gridDim.x) {
        const uint32_t p = w * 64u + threadIdx.x;
        const PathCoord c = path_coord(P, p);
        const bool valid = path_valid(P, c);
        Ray ray{};
        if (valid) {
            ray = camera_ray(P.camera, P.settings, c.x, c.y, path_sampler(P, c).next2(0u));
            ray.tmin = kRayEps;
            ray.tmax = kFltMax;
        }
        if (valid) cnt.ray();
        const HitRec best = intersect_closest_packet<W>(P.scene, ray, valid, wstack, cnt);
        if (valid) P.B.H[p] = make_float4(best.t, best.b1, best.b2, __uint_as_float(best.slot));
    }
    cnt.flush(P.wfCount, 0u);
}

// Shading class of the hit of path p (kShadeMissClass for a miss).
DEV uint32_t hit_class(const WfParams& P, uint32_t p) {
    const uint32_t slot = __float_as_uint(P.B.H[p].w);
    return slot == kNoHit ? kShadeMissClass : P.scene.triClass[slot];
}

// Range of queue positions of block b in the class sort: kOrderBlocks
// contiguous ranges, whole multiples of the block size.
DEV void order_range(uint32_t n, uint32_t b, uint32_t* lo, uint32_t* hi) {
    uint32_t r = (n + kOrderBlocks - 1u) / kOrderBlocks;
    r = (r + kOrderThreads - 1u) / kOrderThreads * kOrderThreads;
    const uint64_t l = static_cast<uint64_t>(b) * r;
    *lo = static_cast<uint32_t>(l < n ? l : n);
    *hi = static_cast<uint32_t>(l + r < n ? l + r : n);
}

// Adds `valid` lanes to the LDS counter of their class, one LDS atomic per
// distinct class of the wave; returns each lane's rank among its wave's
// lanes of the same class, and in *base the counter's value before the add.
DEV uint32_t wave_class_add(bool valid, uint32_t k, uint32_t* counter, uint32_t* base) {
    const uint32_t lane = lane_id();
    bool pending = valid;
    uint32_t rank = 0, b = 0;
    for (;;) {
        const unsigned long long live = __ballot(pending);
        if (live == 0ull) break;
        const int leader = __ffsll(static_cast<long long>(live)) - 1;
        const uint32_t lk = __shfl(k, leader);
        const bool mine = pending && k == lk;
        const unsigned long long same = __ballot(mine);
        uint32_t got = 0;
        if (static_cast<int>(lane) == leader) got = atomicAdd(counter + lk, static_cast<uint32_t>(__popcll(same)));
        got = __shfl(got, leader);
        if (mine) {
            rank = static_cast<uint32_t>(__popcll(same & ((1ull << lane) - 1ull)));
            b = got;
            pending = false;
        }
    }
    *base = b;
    return rank;
}

// Class sort of the depth's hit queue, pass 1: class of every position and
// the per-block class counts.
__global__ __launch_bounds__(kOrderThreads) void k_wf_classify(WfParams P, uint32_t depth) {
    __shared__ uint32_t hist[kShadeClasses];
    const uint32_t n = P.B.counts[depth];
    const uint32_t* in = (depth & 1u) ? P.B.qb : P.B.qa;
    if (threadIdx.x < kShadeClasses) hist[threadIdx.x] = 0u;
    __syncthreads();
    uint32_t lo, hi;
    order_range(n, blockIdx.x, &lo, &hi);
    for (uint32_t base = lo; base < hi; base += kOrderThreads) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < hi;
        uint32_t k = 0;
        if (valid) {
            k = hit_class(P, in[i]);
            P.B.qkey[i] = static_cast<uint8_t>(k);
        }
        uint32_t unused;
        wave_class_add(valid, k, hist, &unused);
    }
    __syncthreads();
    if (threadIdx.x < kShadeClasses) P.B.blockCount[threadIdx.x * kOrderBlocks + blockIdx.x] = hist[threadIdx.x];
}

// Pass 2: block k scans class k's counts over the blocks (first slot of each
// block within the class) and stores the class total.
__global__ __launch_bounds__(kOrderThreads) void k_wf_order_scan(WfParams P) {
    constexpr uint32_t kPer = kOrderBlocks / kOrderThreads;
    __shared__ uint32_t part[kOrderThreads];
    const uint32_t k = blockIdx.x;
    const uint32_t* c = P.B.blockCount + k * kOrderBlocks + threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t sum = 0;
    for (uint32_t j = 0; j < kPer; ++j) {
        v[j] = c[j];
        sum += v[j];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < kOrderThreads; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    uint32_t* out = P.B.blockStart + k * kOrderBlocks + threadIdx.x * kPer;
    for (uint32_t j = 0; j < kPer; ++j) {
        out[j] = run;
        run += v[j];
    }
    if (threadIdx.x == kOrderThreads - 1u) P.B.blockStart[kShadeClasses * kOrderBlocks + k] = part[threadIdx.x];
}

// Pass 3: scatters the queue into class order.  Class k starts after the
// totals of the smaller classes; the block's hits of class k take
// consecutive slots from its scanned start, in the order of the wave
// atomics.  Every path's result is independent of where it is shaded.
__global__ __launch_bounds__(kOrderThreads) void k_wf_order(WfParams P, uint32_t depth) {
    __shared__ uint32_t start[kShadeClasses];
    __shared__ uint32_t cursor[kShadeClasses];
    const uint32_t n = P.B.counts[depth];
    const uint32_t* in = (depth & 1u) ? P.B.qb : P.B.qa;
    if (threadIdx.x < 64u) {
        // Exclusive scan of the class totals, one class per lane of wave 0.
        const uint32_t lane = threadIdx.x;
        const uint32_t total = P.B.blockStart[kShadeClasses * kOrderBlocks + lane];
        uint32_t incl = total;
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off);
            if (lane >= off) incl += t;
        }
        start[lane] = incl - total + P.B.blockStart[lane * kOrderBlocks + blockIdx.x];
        cursor[lane] = 0u;
    }
    __syncthreads();
    uint32_t lo, hi;
    order_range(n, blockIdx.x, &lo, &hi);
    for (uint32_t base = lo; base < hi; base += kOrderThreads) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < hi;
        const uint32_t k = valid ? P.B.qkey[i] : 0u;
        uint32_t got;
        const uint32_t rank = wave_class_add(valid, k, cursor, &got);
        if (valid) P.B.qsorted[start[k] + got + rank] = in[i];
    }
}

// Inclusive prefix sum over the lanes of a wave.
DEV uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
    for (uint32_t d = 1u; d < 64u; d <<= 1) {
        const uint32_t t = __shfl_up(v, d);
        if (lane >= d) v += t;
    }
    return v;
}

// Shade blocks: four waves, whose queue appends are reserved together.
constexpr uint32_t kShadeThreads = 256;
constexpr uint32_t kShadeWaves = kShadeThreads / 64u;

// Reserves c[k] consecutive slots of queue k (k = 0, 1, 2; counters
// count[k]) for every thread of the block and returns the first in pos[k]:
// one global atomic per queue and block.  Appends to one queue counter
// from every wave of the grid serialise on that counter's address (about
// 88M atomics per second on one address were measured), which bounded the
// first bounce's shade stage; a block of four waves needs a quarter of
// them.  Every thread of the block calls this, outside divergent code.
DEV void block_reserve3(const uint32_t c[3], uint32_t* count0, uint32_t* count1, uint32_t* count2, uint32_t pos[3]) {
    __shared__ uint32_t waveTotal[3][kShadeWaves];
    __shared__ uint32_t blockBase[3];
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    uint32_t incl[3];
    for (int k = 0; k < 3; ++k) {
        incl[k] = wave_incl_scan(c[k]);
        if (lane == 63u) waveTotal[k][wave] = incl[k];
    }
    __syncthreads();
    if (threadIdx.x < 3u) {
        const uint32_t k = threadIdx.x;
        uint32_t total = 0;
        for (uint32_t w = 0; w < kShadeWaves; ++w) total += waveTotal[k][w];
        uint32_t* count = k == 0u ? count0 : (k == 1u ? count1 : count2);
        blockBase[k] = total ? atomicAdd(count, total) : 0u;
    }
    __syncthreads();
    for (int k = 0; k < 3; ++k) {
        uint32_t off = blockBase[k];
        for (uint32_t w = 0; w < wave; ++w) off += waveTotal[k][w];
        pos[k] = off + incl[k] - c[k];
    }
    __syncthreads();  // the shared totals are rewritten by the next call
}

// FIRST: the first bounce (depth 0), compiled apart so that the depth
// tests of the shading code fold; `depthArg` is the depth otherwise.
// At most 168 VGPRs, three waves per SIMD (a few registers spill): the
// kernel is bound by its instruction stream, and the third wave hides more
// of its latencies than the spills cost (c3 shade stage 34.4 -> 29.4 ms;
// capped at 128 VGPRs, four waves, it spills 38 registers and loses).
// PLAIN: the scene's materials need no anisotropy, thin-film, sheen or
// subsurface-wrap terms (DevScene::plainBsdf): the BSDF evaluation leaves
// them out.
template <bool TEX, bool FIRST, bool PLAIN>
__global__ __launch_bounds__(kShadeThreads) __attribute__((amdgpu_waves_per_eu(PLAIN && FIRST ? 4 : 3))) void k_wf_shade(WfParams P,
                                                                                                   uint32_t depthArg) {
    const uint32_t depth = FIRST ? 0u : depthArg;
    const uint32_t* in = depth > 0u && P.B.shadeOrder ? P.B.qsorted : ((depth & 1u) ? P.B.qb : P.B.qa);
    uint32_t* out = (depth & 1u) ? P.B.qa : P.B.qb;
    const uint32_t n = stage_items(P, depth);
    const uint32_t nLight = direct_sample_count(P.scene, P.settings);
    const bool more = depth + 1u < P.settings.maxDepth;
    for (uint32_t base = blockIdx.x * kShadeThreads; base < n; base += gridDim.x * kShadeThreads) {
        const uint32_t i = base + threadIdx.x;
        bool alive = false, shadow = false, extChanged = false;
        uint32_t ns = 0;
        uint32_t p = 0;
        bool work = i < n;
        PathCoord pc{};
        if (work) {
            p = depth == 0u ? i : in[i];
            pc = path_coord(P, p);
            work = depth > 0u || path_valid(P, pc);
        }
        if (work) {
            const Sampler smp = path_sampler(P, pc);
            PathState ps = depth == 0u ? init_path(P.camera, P.settings, pc.x, pc.y, smp) : load_state(P.B, p);
            const float4 hh = P.B.H[p];
            if (P.hitLog)
                P.hitLog[replay_index(P.settings.width, P.settings.height, P.batchCount, pc.x, pc.y,
                                      P.sliceBegin + pc.s, depth)] = hh;
            HitRec h;
            h.t = hh.x;
            h.b1 = hh.y;
            h.b2 = hh.z;
            h.slot = __float_as_uint(hh.w);
            Counters<false> c;
            if (h.slot == kNoHit) {
                handle_miss(P.scene, P.settings, depth, ps, c);
            } else {
                const Ray ray = path_ray(ps);
                Surface sf;
                TexState ts;
                if (!shade_hit<TEX>(P.scene, depth, smp, ray, h, ps, sf, &ts)) {
                    alive = true;
                } else {
                    const Mat& M = sf.M;
                    const size_t r0 = static_cast<size_t>(p) * P.B.maxShadow;
                    // nLight <= maxShadow by construction (enqueue_wavefront); the
                    // bound keeps the records inside the path's slots regardless.
                    for (uint32_t j = 0; j < nLight && ns < P.B.maxShadow; ++j) {
                        V3 ld, contrib;
                        uint32_t group;
                        float maxDist;
                        if (!direct_sample<false, PLAIN>(P.scene, P.settings, sf, M, smp, depth, j, &ld, &contrib,
                                                         &group, &maxDist, c))
                            continue;
                        P.B.sr[2u * (r0 + ns)] = make_float4(ld.x, ld.y, ld.z, maxDist);
                        P.B.sr[2u * (r0 + ns) + 1u] = make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(group));
                        ++ns;
                    }
                    if (ns > 0u) {
                        const V3 origin = sf.p + sf.n * kRayEps;
                        P.B.so[p] = make_float4(origin.x, origin.y, origin.z, __uint_as_float(ns));
                        P.B.st[p] = make_float4(ps.T.x, ps.T.y, ps.T.z, __uint_as_float(ns));
                        shadow = true;
                    }
                    const uint32_t insideBefore = ps.bits & kBitsInside;
                    alive = sample_bounce<PLAIN>(P.scene, P.settings, depth, smp, sf, ps);
                    extChanged = (ps.bits & kBitsInside) != insideBefore && (ps.bits & kBitsInside);
                }
            }
            store_state(P.B, p, ps, depth, alive, extChanged);
        }
        // Queue appends: the path's shadow rays, the path to the shadow
        // queue, the path to the next bounce's ray queue.
        const uint32_t c[3] = {ns, shadow ? 1u : 0u, alive && more ? 1u : 0u};
        uint32_t pos[3];
        block_reserve3(c, &P.B.counts[P.B.shadowRayBase + depth], &P.B.counts[P.B.shadowPathBase + depth],
                       &P.B.counts[depth + 1u], pos);
        for (uint32_t j = 0; j < ns; ++j) P.B.qsr[pos[0] + j] = p * P.B.maxShadow + j;
        if (shadow) P.B.qs[pos[1]] = p;
        if (alive && more) out[pos[2]] = p;
    }
}

// Adds the visible light samples of each shaded path, in DirectLighting's
// order, times the throughput at the hit (CpuPathTracer.cpp:1750-1762);
// with the replay log on, records which of them were visible.
__global__ __launch_bounds__(64) void k_wf_resolve(WfParams P, uint32_t depth) {
    const uint32_t n = P.B.counts[P.B.shadowPathBase + depth];
    for (uint32_t i = blockIdx.x * 64u + threadIdx.x; i < n; i += gridDim.x * 64u) {
        const uint32_t p = P.B.qs[i];
        const float4 tn = P.B.st[p];  // throughput at the hit | record count
        const uint32_t ns = __float_as_uint(tn.w);
        const float4* sc = P.B.sr + 2u * static_cast<size_t>(p) * P.B.maxShadow;
        const uint32_t* vis = P.B.svis + static_cast<size_t>(p) * P.B.maxShadow;
        DirectSum ds;
        ds.init(P.settings);
        uint32_t bits = 0;
        for (uint32_t j = 0; j < ns; ++j) {
            const float4 a = sc[2u * j + 1u];
            const bool visible = vis[j] != 0u;
            ds.add(__float_as_uint(a.w), xyz(a), visible);
            bits |= (visible ? 1u : 0u) << (j & 31u);
        }
        ds.flush();
        if (P.visLog) {
            // At most 32 light samples per vertex when the log is on.
            const PathCoord pc = path_coord(P, p);
            P.visLog[replay_index(P.settings.width, P.settings.height, P.batchCount, pc.x, pc.y, P.sliceBegin + pc.s,
                                  depth)] = bits;
        }
        const float4 L = P.B.L[p];
        const V3 Lv = xyz(L) + xyz(tn) * ds.L;
        P.B.L[p] = make_float4(Lv.x, Lv.y, Lv.z, L.w);
    }
}

// The tail of a pass: from bounce depth0 on few paths are alive, and each
// stage launch per bounce is bound by the latency of its slowest ray (c3:
// 0.4-0.6 ms per bounce for depths 4-15, 6.4 ms of a 88 ms frame, for a
// few percent of the rays).  k_wf_tail finishes those paths in one launch:
// persistent one-wave blocks whose lanes take paths from the depth's ray
// queue (one atomic per wave) and run trace_path's bounce body
// (dev_path.h) on each until it ends -- closest hit and shadow rays on the
// LDS ring stacks as in k_render_plus -- then store its radiance; a lane
// whose path ended takes the next.  Every path runs the operations the
// stage kernels would have run, in the same order, so frames are
// unchanged.  Not used with the replay log on (the tail records no hits).
template <bool TEX, bool PLAIN>
__global__ __launch_bounds__(64) void k_wf_tail(WfParams P, uint32_t depth0) {
    const DevSettings& st = P.settings;
    const DevScene& S = P.scene;
    const uint32_t lane = threadIdx.x;
    const uint32_t n = P.B.counts[depth0];
    const uint32_t* in = (depth0 & 1u) ? P.B.qb : P.B.qa;
    uint32_t* counter = &P.B.counts[P.B.tailBase];
    const uint32_t nLight = direct_sample_count(S, st);
    uint32_t ovNode[kStackSize];
    float ovT[kStackSize];
    Counters<false> c;
    bool active = false, exhausted = false;
    uint32_t p = 0, depth = 0;
    PathCoord pc{};
    PathState ps;
    for (;;) {
        const bool want = !active && !exhausted;
        const unsigned long long wm = __ballot(want);
        if (wm) {
            const int leader = __ffsll(static_cast<long long>(wm)) - 1;
            uint32_t first = 0;
            if (static_cast<int>(lane) == leader) first = atomicAdd(counter, static_cast<uint32_t>(__popcll(wm)));
            first = __shfl(first, leader);
            if (want) {
                const uint32_t idx = first + static_cast<uint32_t>(__popcll(wm & ((1ull << lane) - 1ull)));
                if (idx >= n) {
                    exhausted = true;
                } else {
                    p = in[idx];
                    pc = path_coord(P, p);
                    ps = load_state(P.B, p);
                    depth = depth0;
                    active = true;
                }
            }
        }
        if (__all(exhausted)) break;
        if (!active) continue;
        // One bounce (trace_path's loop body).
        const Sampler smp = path_sampler(P, pc);
        bool ended = false;
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
                    if (!direct_sample<false, PLAIN>(S, st, sf, M, smp, depth, j, &ld, &contrib, &group, &maxDist, c))
                        continue;
                    Ray sr;
                    sr.o = origin;
                    sr.d = ld;
                    sr.tmin = kRayEps;
                    sr.tmax = maxDist;
                    ds.add(group, contrib, !intersect_any_ring(S, sr, lane, ovNode));
                }
                ds.flush();
                ps.L = ps.L + ps.T * ds.L;
                if (!sample_bounce<PLAIN>(S, st, depth, smp, sf, ps)) ended = true;
            }
        }
        ++depth;
        if (depth >= st.maxDepth) ended = true;
        if (ended) {
            P.B.L[p] = make_float4(ps.L.x, ps.L.y, ps.L.z, 0.0f);
            active = false;
        }
    }
}

// Per-pixel sums over the pass's samples in sample order, then the batch
// means on the last pass of the pixel's tile (CpuPathTracer.cpp:2098-2137).
// The running sums between passes are kept per pass tile (slot tile * 64 +
// lane): the passes of one tile range follow each other, so only that
// range's pixels need them.
__global__ __launch_bounds__(64) void k_wf_accum(WfParams P) {
    const uint32_t tile = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t tg = P.tileBase + tile;
    const uint32_t x = (tg % P.tilesX) * 8u + (lane & 7u);
    const uint32_t y = (tg / P.tilesX) * 8u + (lane >> 3);
    const DevSettings& st = P.settings;
    if (x >= st.width || y >= st.height) return;
    const size_t pix = static_cast<size_t>(y) * st.width + x;
    const size_t slot = static_cast<size_t>(tile) * 64u + lane;
    float4 beauty, albedo, normal;
    float depth;
    if (P.sliceBegin == 0u) {
        beauty = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        albedo = normal = beauty;
        depth = kInf;
    } else {
        beauty = P.sums.beauty[slot];
        albedo = P.sums.albedo[slot];
        normal = P.sums.normal[slot];
        depth = P.sums.depth[slot];
    }
    // maxDepth 0: no stage of the pass ran, every sample is the empty path
    // (radiance, albedo and normal 0, depth +inf: SampleResult's defaults,
    // CpuPathTracer.cpp:108-113), as the megakernel's init_path leaves it.
    const bool empty = st.maxDepth == 0u;
    for (uint32_t s = 0; s < P.sliceCount; ++s) {
        const uint32_t p = path_index(P, s, tile, lane & 7u, lane >> 3);
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 L = empty ? zero : P.B.L[p];
        const float4 A = empty ? make_float4(0.0f, 0.0f, 0.0f, kInf) : P.B.A[p];
        const float4 N = empty ? zero : P.B.N[p];
        beauty.x = beauty.x + L.x;
        beauty.y = beauty.y + L.y;
        beauty.z = beauty.z + L.z;
        beauty.w = beauty.w + (A.w < kInf ? 1.0f : 0.0f);
        albedo.x = albedo.x + A.x;
        albedo.y = albedo.y + A.y;
        albedo.z = albedo.z + A.z;
        normal.x = normal.x + N.x;
        normal.y = normal.y + N.y;
        normal.z = normal.z + N.z;
        if (A.w < depth) depth = A.w;
    }
    if (P.sliceBegin + P.sliceCount < P.batchCount) {
        P.sums.beauty[slot] = beauty;
        P.sums.albedo[slot] = albedo;
        P.sums.normal[slot] = normal;
        P.sums.depth[slot] = depth;
        return;
    }
    const float nf = static_cast<float>(P.batchCount);
    const float inv = 1.0f / nf;
    float* b = P.frame.beauty + pix * 4u;
    b[0] = beauty.x / nf;
    b[1] = beauty.y / nf;
    b[2] = beauty.z / nf;
    b[3] = beauty.w / nf;
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

// Launches the traversal kernel for the fetch mode and tree width (the
// counting instantiation when P.wfCount is set).
template <bool ANY, bool COUNT>
void launch_trace_as(const WfParams& P, const WfLaunch& L, uint32_t d, uint32_t blocks, uint32_t tblocks,
                     hipStream_t s) {
    const bool w8 = L.bvhWidth == 8u && P.scene.nodes8 != nullptr;
    switch (L.fetchMode * 2u + (w8 ? 1u : 0u)) {
        case 5: hipLaunchKernelGGL((k_wf_trace<ANY, 2, 8, COUNT>), dim3(tblocks), dim3(64), 0, s, P, d); break;
        case 4: hipLaunchKernelGGL((k_wf_trace<ANY, 2, 4, COUNT>), dim3(tblocks), dim3(64), 0, s, P, d); break;
        case 3: hipLaunchKernelGGL((k_wf_trace<ANY, 1, 8, COUNT>), dim3(blocks), dim3(64), 0, s, P, d); break;
        case 2: hipLaunchKernelGGL((k_wf_trace<ANY, 1, 4, COUNT>), dim3(blocks), dim3(64), 0, s, P, d); break;
        case 1: hipLaunchKernelGGL((k_wf_trace<ANY, 0, 8, COUNT>), dim3(blocks), dim3(64), 0, s, P, d); break;
        default: hipLaunchKernelGGL((k_wf_trace<ANY, 0, 4, COUNT>), dim3(blocks), dim3(64), 0, s, P, d); break;
    }
}

template <bool ANY>
void launch_trace(const WfParams& P, const WfLaunch& L, uint32_t d, uint32_t blocks, uint32_t tblocks, hipStream_t s) {
    if (P.wfCount) {
        launch_trace_as<ANY, true>(P, L, d, blocks, tblocks, s);
    } else {
        launch_trace_as<ANY, false>(P, L, d, blocks, tblocks, s);
    }
}

// The shade kernel for the scene's textures and materials and the depth.
void launch_shade(const WfParams& P, uint32_t d, uint32_t blocks, hipStream_t s) {
    const uint32_t v = (P.scene.textured ? 4u : 0u) + (d == 0u ? 2u : 0u) + (P.scene.plainBsdf ? 1u : 0u);
    const dim3 g(blocks), b(kShadeThreads);
    switch (v) {
        case 7: hipLaunchKernelGGL((k_wf_shade<true, true, true>), g, b, 0, s, P, d); break;
        case 6: hipLaunchKernelGGL((k_wf_shade<true, true, false>), g, b, 0, s, P, d); break;
        case 5: hipLaunchKernelGGL((k_wf_shade<true, false, true>), g, b, 0, s, P, d); break;
        case 4: hipLaunchKernelGGL((k_wf_shade<true, false, false>), g, b, 0, s, P, d); break;
        case 3: hipLaunchKernelGGL((k_wf_shade<false, true, true>), g, b, 0, s, P, d); break;
        case 2: hipLaunchKernelGGL((k_wf_shade<false, true, false>), g, b, 0, s, P, d); break;
        case 1: hipLaunchKernelGGL((k_wf_shade<false, false, true>), g, b, 0, s, P, d); break;
        default: hipLaunchKernelGGL((k_wf_shade<false, false, false>), g, b, 0, s, P, d); break;
    }
}

hipError_t launch_wavefront_pass(const WfParams& P, const WfLaunch& L, hipStream_t s, StageMarks* marks) {
    const uint32_t total = P.sliceCount * P.tiles * 64u;
    if (total == 0) return hipSuccess;
    StageMarks none;
    StageMarks& m = marks ? *marks : none;
    const uint32_t fullBlocks = (total + 63u) / 64u;
    const uint32_t blocks = fullBlocks < L.maxBlocks ? fullBlocks : L.maxBlocks;
    hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(256), 0, s, P.B.counts, P.B.countSlots);
    m.mark(SHIRO_HIP_STAGE_RESET, s);
    const uint32_t tblocks = fullBlocks < L.traceBlocks ? fullBlocks : L.traceBlocks;
    // Bounce d's shadow rays and bounce d + 1's closest hits both need only
    // bounce d's shading, and touch disjoint arrays (svis and the radiance
    // against the hit records), so with an aux stream they run side by side:
    // each persistent grid's last waves drain while the other's run, and
    // the memory-bound resolve overlaps the traversal.  Per-stage timing
    // keeps the serial order.
    const bool overlap = L.aux != nullptr && marks == nullptr;
    hipError_t err = hipSuccess;
    bool joinPending = false;
    auto join = [&]() {
        if (joinPending) {
            const hipError_t e = hipStreamWaitEvent(s, L.join, 0);
            if (e != hipSuccess && err == hipSuccess) err = e;
            joinPending = false;
        }
    };
    auto tailAt = [&](uint32_t d) { return d > 0u && d == L.tailDepth && !P.hitLog; };
    bool traced = false;  // the closest hit of bounce d was launched with bounce d - 1's shadow rays
    for (uint32_t d = 0; d < P.settings.maxDepth; ++d) {
        if (tailAt(d)) {
            // The rest of the pass in one persistent kernel.
            join();
            const dim3 g(L.tailBlocks), b(64);
            switch ((P.scene.textured ? 2u : 0u) + (P.scene.plainBsdf ? 1u : 0u)) {
                case 3: hipLaunchKernelGGL((k_wf_tail<true, true>), g, b, 0, s, P, d); break;
                case 2: hipLaunchKernelGGL((k_wf_tail<true, false>), g, b, 0, s, P, d); break;
                case 1: hipLaunchKernelGGL((k_wf_tail<false, true>), g, b, 0, s, P, d); break;
                default: hipLaunchKernelGGL((k_wf_tail<false, false>), g, b, 0, s, P, d); break;
            }
            m.mark(SHIRO_HIP_STAGE_TAIL, s);
            break;
        }
        if (traced) {
            traced = false;
        } else if (d == 0u && L.primaryPackets) {
            const bool w8 = L.bvhWidth == 8u && P.scene.nodes8 != nullptr;
            if (P.wfCount) {
                if (w8) {
                    hipLaunchKernelGGL((k_wf_primary<8, true>), dim3(blocks), dim3(64), 0, s, P);
                } else {
                    hipLaunchKernelGGL((k_wf_primary<4, true>), dim3(blocks), dim3(64), 0, s, P);
                }
            } else if (w8) {
                hipLaunchKernelGGL((k_wf_primary<8, false>), dim3(blocks), dim3(64), 0, s, P);
            } else {
                hipLaunchKernelGGL((k_wf_primary<4, false>), dim3(blocks), dim3(64), 0, s, P);
            }
            m.mark(SHIRO_HIP_STAGE_PRIMARY, s);
        } else {
            launch_trace<false>(P, L, d, blocks, tblocks, s);
            m.mark(SHIRO_HIP_STAGE_EXTEND, s);
        }
        if (d > 0u && P.B.shadeOrder) {
            hipLaunchKernelGGL(k_wf_classify, dim3(kOrderBlocks), dim3(kOrderThreads), 0, s, P, d);
            hipLaunchKernelGGL(k_wf_order_scan, dim3(kShadeClasses), dim3(kOrderThreads), 0, s, P);
            hipLaunchKernelGGL(k_wf_order, dim3(kOrderBlocks), dim3(kOrderThreads), 0, s, P, d);
            m.mark(SHIRO_HIP_STAGE_ORDER, s);
        }
        join();  // bounce d - 1's resolve updates the radiance bounce d's shading reads
        const uint32_t sblocks = (blocks + kShadeWaves - 1u) / kShadeWaves;
        launch_shade(P, d, sblocks, s);
        m.mark(d == 0u ? SHIRO_HIP_STAGE_SHADE_FIRST : SHIRO_HIP_STAGE_SHADE, s);
        if (overlap) {
            hipError_t e = hipEventRecord(L.fork, s);
            if (e == hipSuccess) e = hipStreamWaitEvent(L.aux, L.fork, 0);
            if (e != hipSuccess) return e;
            launch_trace<true>(P, L, d, blocks, tblocks, L.aux);
            hipLaunchKernelGGL(k_wf_resolve, dim3(blocks), dim3(64), 0, L.aux, P, d);
            e = hipEventRecord(L.join, L.aux);
            if (e != hipSuccess) return e;
            joinPending = true;
            const uint32_t next = d + 1u;
            if (next < P.settings.maxDepth && !tailAt(next)) {
                launch_trace<false>(P, L, next, blocks, tblocks, s);
                traced = true;
            }
        } else {
            launch_trace<true>(P, L, d, blocks, tblocks, s);
            m.mark(SHIRO_HIP_STAGE_SHADOW, s);
            hipLaunchKernelGGL(k_wf_resolve, dim3(blocks), dim3(64), 0, s, P, d);
            m.mark(SHIRO_HIP_STAGE_RESOLVE, s);
        }
    }
    join();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_wf_accum, dim3(P.tiles), dim3(64), 0, s, P);
    m.mark(SHIRO_HIP_STAGE_ACCUM, s);
    return hipGetLastError();
}

// Scratch memory of the entry points that allocate nothing
// (shiro_hip_render_batch_device_noalloc): path state, queues, counters and
// running sums of a wavefront pass carved out of one file-scope device
// array.  Every GPU the library runs on holds its own copy until the process
// ends; the host obtains each device's address with hipGetSymbolAddress and
// uses only that address (the array's name is not a device address on the
// host).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

hipError_t scratch_address(int device, void** ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* address[kMaxDevices] = {};
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!address[device]) {
        // Looked up once per device, with that device current.
        int current = 0;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess) return e;
        e = hipSetDevice(device);
        if (e != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        const hipError_t back = hipSetDevice(current);
        if (e != hipSuccess) return e;
        if (back != hipSuccess) return back;
        address[device] = p;
    }
    *ptr = address[device];
    return hipSuccess;
}

}  // namespace shiro_hip
