// C ABI of the backend (include/shiro_hip.h): context, scene upload, batch
// rendering, device accumulation and the single-stage entry points.
//
// The context owns one HIP stream on one device and the device-resident
// scene.  Calls on a context are serialised by a mutex, as the reference
// backends serialise RenderSampleBatch (CpuPathTracer.h:26-31,
// OptixBackend.h:24-26).  Errors are reported as shiro_hip_result codes with
// a message retrievable through shiro_hip_last_error; nothing throws across
// the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/shiro_hip.h"
#include "bvh_build.h"
#include "dev_math.h"
#include "host_scene.h"
#include "kernels.h"

using namespace shiro_hip;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grows to at least `n` bytes (contents not preserved).
    hipError_t reserve(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) return hipSuccess;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        bytes = n;
        return hipSuccess;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

struct EnvBuffers {
    DevBuf texels, row, cond, rowGuide, condGuide;
    DevBuf rowSin, rowW, rgb;   // device updates: row factors, row totals, staging of host texels
    uint32_t w = 0, h = 0;
    bool table = false;
};

// Texture extension: the texels, and a second copy that sampling reads
// while frozen (shiro_hip_update_texture with EVAL_ONLY).
struct TexBuffers {
    DevBuf eval, sampling;
    uint32_t w = 0, h = 0, channels = 0;
    bool frozen = false;
    size_t count() const { return static_cast<size_t>(w) * h * channels; }
};

// The arrays of the wavefront passes of one stream (wavefront_sizes order:
// path state, shadow records, queues, class order, counters, running sums).
constexpr size_t kWfArrays = 25;
struct WfState {
    DevBuf a[kWfArrays];
};

}  // namespace

struct shiro_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Second stream of the wavefront passes (WfLaunch::aux) and its
    // fork / join events, created on first use.
    hipStream_t auxStream = nullptr;
    hipEvent_t auxFork = nullptr, auxJoin = nullptr;
    std::mutex mu;
    std::string err;
    float lastMs = 0.0f;
    // Scene.
    bool hasScene = false;
    bool emissive = false;
    DevBuf nodes, nodes8, tris, hitdata, materials, derived, uvs, lightList, lightCdf, triClass;
    Bvh bvh;                // host copy of the uploaded trees, for refits
    uint64_t topology = 0;  // topology_hash of the uploaded scene
    std::deque<EnvBuffers> env;           // per dome light (index = dome index)
    std::vector<DevEnv> hostEnvs;         // host copies of the device light arrays
    std::vector<DevDome> hostDomes;
    std::vector<DevDirLight> hostDirs;
    DevBuf envsBuf, domesBuf, dirsBuf;
    std::vector<std::unique_ptr<TexBuffers>> tex;
    DevBuf texTable;
    std::vector<shiro_hip_material> mats;   // host copy of the uploaded materials
    DevScene dev{};
    shiro_hip_scene_info info{};
    // Per-batch scratch.
    DevBuf sumB, sumA, sumN, sumD;
    DevBuf frB, frA, frN, frD;
    DevBuf counters;
    // Wavefront state of the context stream.
    WfState wfs;
    // Accumulator.
    DevBuf acB, acA, acN, acD, acC;
    uint32_t acW = 0, acH = 0, acTotal = 0;
    // Wavefront tuning (shiro_hip_set_tuning / shiro_hip_autotune).
    // Defaults: what the autotuner picks on the 4K frames of the large
    // scenes (one pass of 128M paths, lanes refilled from per-wave pools,
    // the tail kernel two bounces after roulette starts), with wave rounds
    // for small passes.
    shiro_hip_tuning tuning{128u << 20, SHIRO_HIP_FETCH_AUTO, 8192u, 16384u, 8u, 1u, 16u, 4u, 4u, 24u, 24u,
                            SHIRO_HIP_TAIL_AUTO, 1u};
    // Fetch counters of the wavefront traversal kernels
    // (shiro_hip_count_wavefront): on only for that call's render.
    DevBuf wfCount;
    bool wfCountOn = false;
    uint32_t computeUnits = 256;
    // Batched rendering: items, work list and the work counter.
    DevBuf manyItems, manyWork, manyCounter;
    DevBuf plusCounter;  // pixel counter of the megakernel with path regeneration
    std::vector<uint32_t> manyDepth;   // BVH depth per item of the current batch (cost estimate)
    std::vector<const shiro_hip_batch_item*> manyItemsHost;  // non-empty items of the current batch
    // Batched wavefront in one graph: per branch a stream and its wavefront
    // state; the instantiated graph of the last batch and the key (items,
    // tuning, buffer addresses) it was captured for.
    std::vector<hipStream_t> branchStreams;
    std::vector<std::unique_ptr<WfState>> branchStates;
    std::vector<hipEvent_t> joinEv;
    hipEvent_t forkEv = nullptr;
    hipGraphExec_t manyExec = nullptr;
    uint64_t manyKey = 0;
    uint32_t manyCaptures = 0, manyUpdates = 0;  // captures made, and how many of them updated the cached graph
    // Stage timing (shiro_hip_set_stage_timing): events recorded after each
    // launch of the last render call.
    bool stageTiming = false;
    std::vector<hipEvent_t> stageEv;
    std::vector<uint8_t> stageKind;
    StageMarks marks;
    shiro_hip_stage_times stageTimes{};
    // Device address of this device's scratch array (scratch_address), null
    // if the lookup failed (scratchError says why).
    void* scratch = nullptr;
    std::string scratchError;
    // Scene content version: changes with every upload, refit, texture or
    // map update (the replay log is valid only for the version it saw).
    uint64_t contentVersion = 0;
    // Replay log (shiro_hip_set_replay_log): hits and light-sample
    // visibility of the last wavefront batch that fitted `replayCap` bytes,
    // and what that batch was.
    uint64_t replayCap = 0;
    DevBuf replayHits, replayVis;
    bool replayValid = false;
    uint64_t replayVersion = 0;
    shiro_hip_camera replayCamera{};
    shiro_hip_settings replaySettings{};
    uint32_t replayStart = 0, replayCount = 0;
    bool lastAdjointReplayed = false;  // the last adjoint read the replay log
};

namespace {

int fail(shiro_hip_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(shiro_hip_ctx* c, hipError_t e, const char* what) {
    const int code = e == hipErrorOutOfMemory ? SHIRO_HIP_OOM : SHIRO_HIP_LAUNCH;
    return fail(c, code, std::string(what) + ": " + hipGetErrorString(e));
}

int accum_reset_locked(shiro_hip_ctx* ctx, uint32_t width, uint32_t height);

#define HIPC(expr, what)                          \
    do {                                          \
        hipError_t e_ = (expr);                   \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
    } while (0)

DevCamera to_dev(const shiro_hip_camera& c) {
    DevCamera d;
    std::memcpy(d.position, c.position, 12);
    std::memcpy(d.forward, c.forward, 12);
    std::memcpy(d.right, c.right, 12);
    std::memcpy(d.up, c.up, 12);
    d.tanHalfFov = std::tan(c.verticalFovDegrees * 0.5f * kPi / 180.0f);
    d.aspect = c.aspectRatio;
    return d;
}

DevSettings to_dev(const shiro_hip_settings& s) {
    DevSettings d;
    d.width = s.width;
    d.height = s.height;
    d.domeLightSamples = s.domeLightSamples;
    d.maxDepth = s.maxDepth;
    d.diffuseDepth = s.diffuseDepth;
    d.specularDepth = s.specularDepth;
    d.backgroundVisible = s.backgroundVisible;
    d.rrDepth = s.russianRouletteDepth;
    return d;
}

template <typename T>
hipError_t upload(DevBuf& b, const T* src, size_t count, hipStream_t s) {
    hipError_t e = b.reserve(std::max<size_t>(count * sizeof(T), 16));
    if (e != hipSuccess) return e;
    if (count == 0) return hipSuccess;
    e = hipMemcpyAsync(b.p, src, count * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

// Samples per launch: bound a launch to roughly 8M paths so that the cancel
// flag is polled at a useful rate.
uint32_t slice_samples(uint64_t pixels, uint32_t count) {
    if (pixels == 0) return count;
    const uint64_t per = std::max<uint64_t>(1, (8ull << 20) / pixels);
    return static_cast<uint32_t>(std::min<uint64_t>(count, per));
}

int check_settings(shiro_hip_ctx* ctx, const shiro_hip_settings* s) {
    if (!s) return fail(ctx, SHIRO_HIP_INVALID, "settings is null");
    if (s->width == 0 || s->height == 0) return fail(ctx, SHIRO_HIP_INVALID, "zero-sized frame");
    if (static_cast<uint64_t>(s->width) * s->height > (1ull << 31))
        return fail(ctx, SHIRO_HIP_INVALID, "frame larger than 2^31 pixels");
    if (s->transport > SHIRO_HIP_TRANSPORT_MEGAKERNEL_PLUS) return fail(ctx, SHIRO_HIP_INVALID, "unknown transport");
    if (s->maxDepth > 255u || s->diffuseDepth > 255u || s->specularDepth > 255u)
        return fail(ctx, SHIRO_HIP_INVALID, "depth limits above 255 are not supported");
    if (s->domeLightSamples > 64u) return fail(ctx, SHIRO_HIP_INVALID, "domeLightSamples above 64 is not supported");
    return SHIRO_HIP_OK;
}

// Wavefront form of one batch: passes of S samples for the pixels of a
// range of T 8x8 tiles, S chosen so that a pass holds about
// tuning.pathsPerPass paths; T is the whole frame unless the path state of
// one pass would exceed the memory budget (kWavefrontStateBytes, or the
// scratch array for the entry points that allocate nothing), in which case
// the frame is covered range by range, all samples of a range before the
// next.  Large passes pay: the last bounces of a pass hold few paths and run
// latency-bound on a mostly idle GPU, so fewer passes means fewer such tails
// (c3: 4M -> 64M paths per pass cuts a frame from 245 to 189 ms; 64M ->
// 128M, one pass for the 16-spp frame instead of two, 122 -> 114 ms).
constexpr uint64_t kWavefrontStateBytes = 32ull << 30;

// Starts the event list of a render call in stage-timing mode: the first
// event marks the start of the first kernel.
void begin_marks(shiro_hip_ctx* ctx) {
    ctx->marks = StageMarks{};
    if (!ctx->stageTiming) return;
    ctx->marks.ev = ctx->stageEv.data();
    ctx->marks.kind = ctx->stageKind.data();
    ctx->marks.cap = static_cast<uint32_t>(ctx->stageEv.size());
    ctx->marks.mark(0xFF, ctx->stream);
}

// Where the wavefront's arrays live: the context's buffers (grown on
// demand) or consecutive 256-byte aligned pieces of the scratch array.
struct WfMemory {
    unsigned char* scratch = nullptr;  // null: the context's buffers
    size_t scratchBytes = 0;
    size_t used = 0;
    // One array of n bytes: `own` grown to n, or the next scratch piece.
    hipError_t get(DevBuf& own, size_t n, void** out) {
        if (!scratch) {
            hipError_t e = own.reserve(std::max<size_t>(n, 16));
            *out = own.p;
            return e;
        }
        const size_t at = (used + 255u) & ~size_t{255};
        if (at + n > scratchBytes) return hipErrorOutOfMemory;
        *out = scratch + at;
        used = at + n;
        return hipSuccess;
    }
    // Bytes the scratch pieces take for the given sizes (with alignment).
    static size_t footprint(const size_t* n, size_t k) {
        size_t t = 0;
        for (size_t i = 0; i < k; ++i) t = ((t + 255u) & ~size_t{255}) + n[i];
        return t;
    }
};

// Sizes of a pass's arrays for cap paths, M light samples per hit, the
// shading-class order on or off, sum slots and counter slots.  Order:
// o d T L E A N H so st | sr | qa qb qs | qsr svis | qkey qsorted blockCount
// blockStart | counters | sums beauty albedo normal depth.
void wavefront_sizes(uint64_t cap, uint32_t M, bool order, uint64_t sumSlots, uint32_t slots, size_t* n) {
    for (int i = 0; i < 10; ++i) n[i] = cap * 16;
    // o, d, T and E: one array of 64-byte path records (WfBuffers).
    n[0] = cap * 16 * kPathRecord;
    n[1] = n[2] = n[4] = 0;
    n[10] = cap * M * 32;
    n[11] = n[12] = n[13] = cap * 4;
    n[14] = n[15] = cap * M * 4;
    n[16] = order ? cap : 0;
    n[17] = order ? cap * 4 : 0;
    n[18] = order ? kShadeClasses * kOrderBlocks * 4 : 0;
    n[19] = order ? (kShadeClasses + 1u) * kOrderBlocks * 4 : 0;
    n[20] = std::max<size_t>(sizeof(DevCounters), slots * 4u);
    n[21] = n[22] = n[23] = sumSlots * 16;
    n[24] = sumSlots * 4;
}

// Pass plan of one batch: S samples per pass over ranges of T tiles, and
// the array sizes.
struct WfPlan {
    uint32_t tilesX = 0, tilesTotal = 0, M = 1, slots = 0, shadowRayBase = 0, shadowPathBase = 0, fetchBase = 0,
             tailBase = 0;
    uint64_t S = 1, T = 1, cap = 0;
    bool order = false;
    size_t n[kWfArrays] = {};
};

int wavefront_plan(shiro_hip_ctx* ctx, const DevScene& scene, const shiro_hip_settings& st, uint32_t count,
                   uint64_t budget, WfPlan* out) {
    WfPlan& p = *out;
    p.tilesX = (st.width + 7u) / 8u;
    const uint32_t tilesY = (st.height + 7u) / 8u;
    p.tilesTotal = p.tilesX * tilesY;
    const uint64_t perSample = static_cast<uint64_t>(p.tilesTotal) * 64u;
    const uint32_t K = st.domeLightSamples > 1u ? st.domeLightSamples : 1u;
    // Light samples per hit: the device's direct_sample_count (dev_path.h).
    const uint32_t lights = scene.dirCount + scene.domeCount * K +
                            (scene.emissiveCount > 0u && scene.emissiveSum > 0.0f ? 1u : 0u);
    p.M = std::max(1u, lights);
    p.order = ctx->tuning.shadeOrder != 0u;
    p.shadowRayBase = st.maxDepth + 1u;
    p.shadowPathBase = p.shadowRayBase + st.maxDepth;
    // The striped work counters start on a 128-byte line.
    p.fetchBase = (p.shadowPathBase + st.maxDepth + kPoolStride - 1u) / kPoolStride * kPoolStride;
    p.slots = p.fetchBase + 2u * st.maxDepth * kPoolStripes * kPoolStride;
    p.tailBase = p.slots;  // the tail kernel's path counter, on a line of its own
    p.slots += kPoolStride;
    const uint64_t maxCap = (1ull << 32) / (2ull * p.M) - 1u;  // queue entries path * M + index stay below 2^32
    auto bytes_for = [&](uint64_t S, uint64_t T) {
        wavefront_sizes(S * T * 64u, p.M, p.order, S < count ? T * 64u : 0u, p.slots, p.n);
        return WfMemory::footprint(p.n, kWfArrays);
    };
    auto fits = [&](uint64_t S, uint64_t T) { return bytes_for(S, T) <= budget && S * T * 64u <= maxCap; };
    // Largest x in [1, hi] with ok(x), or 0 (ok is monotone).
    auto largest = [](uint64_t hi, const auto& ok) {
        uint64_t lo = 0;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo + 1) / 2;
            if (ok(mid)) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    p.S = std::min<uint64_t>(count, std::max<uint64_t>(1, ctx->tuning.pathsPerPass / perSample));
    p.T = p.tilesTotal;
    if (!fits(p.S, p.T)) p.S = std::max<uint64_t>(1, largest(p.S, [&](uint64_t x) { return fits(x, p.T); }));
    if (!fits(p.S, p.T)) {
        // One sample of the whole frame does not fit: ranges of tiles.
        p.T = largest(p.tilesTotal, [&](uint64_t x) { return fits(p.S, x); });
        if (p.T == 0)
            return fail(ctx, SHIRO_HIP_INVALID,
                        "one tile of wavefront path state (" + std::to_string(bytes_for(1, 1)) +
                            " bytes) exceeds the " + std::to_string(budget) + "-byte memory budget");
    }
    p.cap = p.S * p.T * 64u;
    wavefront_sizes(p.cap, p.M, p.order, p.S < count ? p.T * 64u : 0u, p.slots, p.n);
    return SHIRO_HIP_OK;
}

// `scene` is ctx's own scene except for batched renders; `events` records
// ev0 / ev1 around the passes; a non-null `scratch` (kScratchBytes) holds
// every array instead of `state`'s buffers.  Work goes to `stream` (null:
// the context stream) with the arrays of `state` (null: the context's);
// `budget` (0: the default for the memory used) bounds the path state.
int enqueue_wavefront(shiro_hip_ctx* ctx, const DevScene& scene, const shiro_hip_camera& cam,
                      const shiro_hip_settings& st, uint32_t start, uint32_t count, const DevFrame& frame,
                      const volatile uint8_t* cancel, bool events = true, void* scratch = nullptr,
                      hipStream_t stream = nullptr, WfState* state = nullptr, uint64_t budget = 0,
                      float4* hitLog = nullptr, uint32_t* visLog = nullptr) {
    if (!stream) stream = ctx->stream;
    if (!state) state = &ctx->wfs;
    if (budget == 0) budget = scratch ? kScratchBytes : kWavefrontStateBytes;
    WfPlan plan;
    int rc = wavefront_plan(ctx, scene, st, count, budget, &plan);
    if (rc) return rc;
    const size_t* n = plan.n;
    WfMemory mem{static_cast<unsigned char*>(scratch), scratch ? kScratchBytes : 0};
    void* a[kWfArrays];
    for (size_t i = 0; i < kWfArrays; ++i) HIPC(mem.get(state->a[i], n[i], &a[i]), "alloc wavefront state");
    WfParams P{};
    P.scene = scene;
    P.camera = to_dev(cam);
    P.settings = to_dev(st);
    P.sums = DevBatchSums{static_cast<float4*>(a[21]), static_cast<float4*>(a[22]), static_cast<float4*>(a[23]),
                          static_cast<float*>(a[24])};
    P.frame = frame;
    WfBuffers& B = P.B;
    float4** f4[11] = {&B.o, &B.d, &B.T, &B.L, &B.E, &B.A, &B.N, &B.H, &B.so, &B.st, &B.sr};
    for (int i = 0; i < 11; ++i) *f4[i] = static_cast<float4*>(a[i]);
    B.d = B.o + 1;
    B.T = B.o + 2;
    B.E = B.o + 3;
    B.qa = static_cast<uint32_t*>(a[11]);
    B.qb = static_cast<uint32_t*>(a[12]);
    B.qs = static_cast<uint32_t*>(a[13]);
    B.qsr = static_cast<uint32_t*>(a[14]);
    B.svis = static_cast<uint32_t*>(a[15]);
    B.counts = static_cast<uint32_t*>(a[20]);
    B.countSlots = plan.slots;
    B.shadowRayBase = plan.shadowRayBase;
    B.shadowPathBase = plan.shadowPathBase;
    B.fetchBase = plan.fetchBase;
    B.tailBase = plan.tailBase;
    B.maxShadow = plan.M;
    B.cap = static_cast<uint32_t>(plan.cap);
    B.shadeOrder = plan.order ? 1u : 0u;
    B.qkey = plan.order ? static_cast<uint8_t*>(a[16]) : nullptr;
    B.qsorted = plan.order ? static_cast<uint32_t*>(a[17]) : nullptr;
    B.blockCount = plan.order ? static_cast<uint32_t*>(a[18]) : nullptr;
    B.blockStart = plan.order ? static_cast<uint32_t*>(a[19]) : nullptr;
    P.sampleStart = start;
    P.batchCount = count;
    P.tilesX = plan.tilesX;
    P.hitLog = hitLog;
    P.visLog = visLog;
    P.leafWait = ctx->tuning.leafWaitLanes;
    P.shadowLeafWait = ctx->tuning.shadowLeafWaitLanes;
    P.wfCount = ctx->wfCountOn ? ctx->wfCount.as<unsigned long long>() : nullptr;
    WfLaunch L{};
    L.maxBlocks = ctx->tuning.stageBlocks;
    L.traceBlocks = ctx->tuning.traceBlocks;
    // Automatic work distribution: per-wave pools pay off on large passes
    // (c2/c3, 133M paths: 75 -> 63 and 105 -> 83 ms per frame against wave
    // rounds) and cost on small ones (c1, 4.2M paths: 1.53 -> 1.82 ms; c5,
    // 2.1M: 1.98 -> 2.66 ms), where most waves get a chunk or two and then
    // probe every stripe's counter.
    const uint64_t passPaths = static_cast<uint64_t>(plan.S) * plan.T * 64u;
    L.fetchMode = ctx->tuning.fetchMode == SHIRO_HIP_FETCH_AUTO ? (passPaths >= kFetchPoolPaths ? 2u : 0u)
                                                                : ctx->tuning.fetchMode;
    L.primaryPackets = ctx->tuning.primaryPackets;
    L.bvhWidth = ctx->tuning.bvhWidth == 8u && scene.nodes8 ? 8u : 4u;
    L.tailDepth = ctx->tuning.tailDepth == SHIRO_HIP_TAIL_AUTO ? st.russianRouletteDepth + 2u : ctx->tuning.tailDepth;
    L.tailBlocks = ctx->computeUnits * 16u;
    {
        // Shadow rays beside the next bounce's closest hits, except under a
        // caller's stream capture (the graph keeps the serial order) and in
        // the entry point that allocates nothing (no stream is created
        // there).
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPC(hipStreamIsCapturing(stream, &cs), "capture status");
        if (cs == hipStreamCaptureStatusNone && !scratch && ctx->tuning.shadowOverlap) {
            if (!ctx->auxStream) {
                HIPC(hipStreamCreateWithFlags(&ctx->auxStream, hipStreamNonBlocking), "aux stream");
                HIPC(hipEventCreateWithFlags(&ctx->auxFork, hipEventDisableTiming), "aux event");
                HIPC(hipEventCreateWithFlags(&ctx->auxJoin, hipEventDisableTiming), "aux event");
            }
            L.aux = ctx->auxStream;
            L.fork = ctx->auxFork;
            L.join = ctx->auxJoin;
        }
    }
    if (events) {
        HIPC(hipEventRecord(ctx->ev0, stream), "event");
        begin_marks(ctx);
    }
    const uint32_t S = static_cast<uint32_t>(plan.S), T = static_cast<uint32_t>(plan.T);
    for (uint32_t t0 = 0; t0 < plan.tilesTotal; t0 += T) {
        P.tileBase = t0;
        P.tiles = std::min(T, plan.tilesTotal - t0);
        for (uint32_t b = 0; b < count; b += S) {
            if (cancel && *cancel) {
                (void)hipStreamSynchronize(stream);
                return fail(ctx, SHIRO_HIP_CANCELLED, "cancelled");
            }
            P.sliceBegin = b;
            P.sliceCount = std::min(S, count - b);
            // Path groups of G samples as asked, or the largest smaller
            // choice that divides the pass's samples per pixel.
            uint32_t g = L.primaryPackets ? L.primaryPackets : 1u;
            while (g > 1u && P.sliceCount % g != 0u) g /= 4u;
            P.groupLog = g == 64u ? 6u : (g == 16u ? 4u : (g == 4u ? 2u : 0u));
            HIPC(launch_wavefront_pass(P, L, stream, ctx->stageTiming && events ? &ctx->marks : nullptr),
                 "wavefront launch");
            if (cancel) HIPC(hipStreamSynchronize(stream), "wavefront");
        }
    }
    if (events) HIPC(hipEventRecord(ctx->ev1, stream), "event");
    return SHIRO_HIP_OK;
}

// Enqueues the slices of one batch; `count` > 0.  frame holds device
// pointers.  If cancel is non-null the host flag is polled between slices.
int enqueue_batch(shiro_hip_ctx* ctx, const shiro_hip_camera& cam, const shiro_hip_settings& st, uint32_t start,
                  uint32_t count, const DevFrame& frame, bool countFetches, const volatile uint8_t* cancel) {
    if (st.transport == SHIRO_HIP_TRANSPORT_WAVEFRONT && !countFetches) {
        // The replay log: one hit record (16 bytes) and one visibility word
        // per vertex, for batches within the cap whose vertices take at
        // most 32 light samples.
        const uint64_t entries = static_cast<uint64_t>(st.width) * st.height * count * st.maxDepth;
        const DevScene& S = ctx->dev;
        const uint64_t lights = static_cast<uint64_t>(S.dirCount) +
                                static_cast<uint64_t>(S.domeCount) * std::max(1u, st.domeLightSamples) +
                                (S.emissiveCount > 0u && S.emissiveSum > 0.0f ? 1u : 0u);
        const bool record = ctx->replayCap > 0 && entries > 0 && entries * 20u <= ctx->replayCap && lights <= 32u;
        ctx->replayValid = false;
        if (!record) return enqueue_wavefront(ctx, ctx->dev, cam, st, start, count, frame, cancel);
        HIPC(ctx->replayHits.reserve(entries * 16u), "alloc replay log");
        HIPC(ctx->replayVis.reserve(entries * 4u), "alloc replay log");
        const int rc = enqueue_wavefront(ctx, ctx->dev, cam, st, start, count, frame, cancel, true, nullptr, nullptr,
                                         nullptr, 0, ctx->replayHits.as<float4>(), ctx->replayVis.as<uint32_t>());
        if (rc) return rc;
        ctx->replayValid = true;
        ctx->replayVersion = ctx->contentVersion;
        ctx->replayCamera = cam;
        ctx->replaySettings = st;
        ctx->replayStart = start;
        ctx->replayCount = count;
        return SHIRO_HIP_OK;
    }
    const uint64_t np = static_cast<uint64_t>(st.width) * st.height;
    const uint32_t slice = slice_samples(np, count);
    RenderParams P{};
    P.scene = ctx->dev;
    P.camera = to_dev(cam);
    P.settings = to_dev(st);
    if (slice < count) {
        HIPC(ctx->sumB.reserve(np * 16), "alloc sums");
        HIPC(ctx->sumA.reserve(np * 16), "alloc sums");
        HIPC(ctx->sumN.reserve(np * 16), "alloc sums");
        HIPC(ctx->sumD.reserve(np * 4), "alloc sums");
    }
    P.sums = DevBatchSums{ctx->sumB.as<float4>(), ctx->sumA.as<float4>(), ctx->sumN.as<float4>(), ctx->sumD.as<float>()};
    P.frame = frame;
    P.sampleStart = start;
    P.batchCount = count;
    P.tilesX = (st.width + 7u) / 8u;
    P.counters = countFetches ? ctx->counters.as<DevCounters>() : nullptr;
    HIPC(hipEventRecord(ctx->ev0, ctx->stream), "event");
    begin_marks(ctx);
    for (uint32_t b = 0; b < count; b += slice) {
        if (cancel && *cancel) {
            (void)hipStreamSynchronize(ctx->stream);
            return fail(ctx, SHIRO_HIP_CANCELLED, "cancelled");
        }
        P.sliceBegin = b;
        P.sliceCount = std::min(slice, count - b);
        if (st.transport == SHIRO_HIP_TRANSPORT_MEGAKERNEL_PLUS && !countFetches) {
            HIPC(ctx->plusCounter.reserve(256), "alloc counter");
            HIPC(launch_render_plus(P, ctx->plusCounter.as<uint32_t>(), ctx->computeUnits * 16u, ctx->stream),
                 "render launch");
        } else {
            HIPC(launch_render(P, countFetches, ctx->stream), "render launch");
        }
        if (ctx->stageTiming) ctx->marks.mark(SHIRO_HIP_STAGE_MEGAKERNEL, ctx->stream);
        if (cancel) HIPC(hipStreamSynchronize(ctx->stream), "render");
    }
    HIPC(hipEventRecord(ctx->ev1, ctx->stream), "event");
    return SHIRO_HIP_OK;
}

int finish_timing(shiro_hip_ctx* ctx) {
    HIPC(hipEventSynchronize(ctx->ev1), "render");
    float ms = 0.0f;
    HIPC(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "timing");
    ctx->lastMs = ms;
    if (ctx->stageTiming && ctx->marks.ev) {
        shiro_hip_stage_times& t = ctx->stageTimes;
        t = shiro_hip_stage_times{};
        for (uint32_t i = 1; i < ctx->marks.n; ++i) {
            float d = 0.0f;
            HIPC(hipEventElapsedTime(&d, ctx->marks.ev[i - 1], ctx->marks.ev[i]), "stage timing");
            const uint8_t k = ctx->marks.kind[i];
            if (k < SHIRO_HIP_STAGE_COUNT) {
                t.ms[k] += d;
                t.launches[k] += 1u;
            }
        }
        t.complete = ctx->marks.overflow ? 0u : 1u;
        ctx->marks = StageMarks{};
    }
    return SHIRO_HIP_OK;
}

void clear_host_frame(const shiro_hip_settings& st, shiro_hip_frame* out) {
    const size_t np = static_cast<size_t>(st.width) * st.height;
    if (out->beauty) std::fill(out->beauty, out->beauty + np * 4, 0.0f);
    if (out->albedo) std::fill(out->albedo, out->albedo + np * 3, 0.0f);
    if (out->normal) std::fill(out->normal, out->normal + np * 3, 0.0f);
    if (out->depth) std::fill(out->depth, out->depth + np, std::numeric_limits<float>::infinity());
}

// Scratch device array for the single-stage entry points.
struct Tmp {
    DevBuf b;
    template <typename T>
    hipError_t in(const T* src, size_t n, hipStream_t s) {
        hipError_t e = b.reserve(std::max<size_t>(n * sizeof(T), 16));
        if (e != hipSuccess || n == 0) return e;
        return hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    }
    template <typename T>
    hipError_t alloc(size_t n) {
        return b.reserve(std::max<size_t>(n * sizeof(T), 16));
    }
    template <typename T>
    hipError_t out(T* dst, size_t n, hipStream_t s) {
        if (n == 0) return hipSuccess;
        return hipMemcpyAsync(dst, b.p, n * sizeof(T), hipMemcpyDeviceToHost, s);
    }
};

}  // namespace

extern "C" {

void shiro_hip_default_settings(shiro_hip_settings* s) {
    std::memset(s, 0, sizeof(*s));
    s->width = 1280;
    s->height = 720;
    s->samplesPerPixel = 32;
    s->samplesPerUpdate = 1;
    s->domeLightSamples = 1;
    s->maxDepth = 3;
    s->diffuseDepth = 2;
    s->specularDepth = 2;
    s->maxSubdivLevel = 2;
    s->backend = 0;
    s->backgroundVisible = 1;
    s->enableHeadlight = 1;
    s->enableCaustics = 1;
    s->enableSss = 1;
    s->russianRouletteDepth = 2;
    s->transport = SHIRO_HIP_TRANSPORT_MEGAKERNEL;
}

void shiro_hip_default_material(shiro_hip_material* m) { *m = default_material(); }

int shiro_hip_create(int device, shiro_hip_ctx** out) {
    if (!out) return SHIRO_HIP_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SHIRO_HIP_NO_DEVICE;
    if (device < 0 || device >= n) return SHIRO_HIP_INVALID;
    auto* ctx = new shiro_hip_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        ctx->counters.reserve(sizeof(DevCounters)) != hipSuccess) {
        delete ctx;
        return SHIRO_HIP_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) {
        ctx->tuning.traceBlocks = static_cast<uint32_t>(prop.multiProcessorCount) * 32u;
        ctx->computeUnits = static_cast<uint32_t>(prop.multiProcessorCount);
    }
    // This device's copy of the scratch array of the no-allocation entry
    // point; on failure that entry point reports the error, the others work.
    const hipError_t se = scratch_address(device, &ctx->scratch);
    if (se != hipSuccess) {
        ctx->scratch = nullptr;
        ctx->scratchError = std::string("scratch address: ") + hipGetErrorString(se);
    }
    *out = ctx;
    return SHIRO_HIP_OK;
}

void shiro_hip_destroy(shiro_hip_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->stageEv) (void)hipEventDestroy(e);
    if (ctx->manyExec) (void)hipGraphExecDestroy(ctx->manyExec);
    for (hipStream_t b : ctx->branchStreams) {
        (void)hipStreamSynchronize(b);
        (void)hipStreamDestroy(b);
    }
    for (hipEvent_t e : ctx->joinEv) (void)hipEventDestroy(e);
    if (ctx->forkEv) (void)hipEventDestroy(ctx->forkEv);
    if (ctx->auxStream) {
        (void)hipStreamSynchronize(ctx->auxStream);
        (void)hipStreamDestroy(ctx->auxStream);
    }
    if (ctx->auxFork) (void)hipEventDestroy(ctx->auxFork);
    if (ctx->auxJoin) (void)hipEventDestroy(ctx->auxJoin);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* shiro_hip_last_error(const shiro_hip_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int shiro_hip_get_status(const shiro_hip_ctx* ctx, shiro_hip_status* out) {
    if (!out) return SHIRO_HIP_INVALID;
    std::memset(out, 0, sizeof(*out));
    out->kind = 1;
    out->name = "hip";
    out->usesGpu = 1;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SHIRO_HIP_OK;
    const int dev = ctx ? ctx->device : 0;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return SHIRO_HIP_OK;
    out->available = 1;
    std::snprintf(out->deviceName, sizeof(out->deviceName), "%s", p.gcnArchName);
    out->computeUnits = static_cast<uint32_t>(p.multiProcessorCount);
    return SHIRO_HIP_OK;
}

int shiro_hip_supports_scene(const shiro_hip_scene* s) {
    if (!s) return 0;
    if (s->distantLightCount > kMaxSceneDirLights) return 0;
    if (s->domeLightCount > kMaxSceneDomes) return 0;
    return 1;
}

// FNV-1a over what fixes the triangle list of a scene: mesh count, and per
// mesh its vertex count and index array.
uint64_t topology_hash(const shiro_hip_scene& s) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    };
    mix(&s.meshCount, 4);
    for (uint32_t m = 0; m < s.meshCount && s.meshes; ++m) {
        const shiro_hip_mesh& M = s.meshes[m];
        mix(&M.vertexCount, 4);
        mix(&M.indexCount, 4);
        if (M.indices) mix(M.indices, static_cast<size_t>(M.indexCount) * 4u);
    }
    return h;
}

// Scene upload; with `refit` the BVH of the uploaded scene is refitted to
// the new vertex positions instead of rebuilt (same topology required).
int upload_scene(shiro_hip_ctx* ctx, const shiro_hip_scene* s, bool refit) {
    if (!ctx) return SHIRO_HIP_INVALID;
    if (!s) return fail(ctx, SHIRO_HIP_INVALID, "scene is null");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ++ctx->contentVersion;
    HIPC(hipSetDevice(ctx->device), "set device");
    const uint64_t topo = topology_hash(*s);
    if (refit && (!ctx->hasScene || topo != ctx->topology))
        return fail(ctx, SHIRO_HIP_INVALID, "refit needs the uploaded scene's meshes and indices unchanged");
    ctx->hasScene = false;
    const auto t0 = std::chrono::steady_clock::now();
    FlatScene flat;
    std::string err;
    if (!flatten_scene(*s, &flat, &err)) return fail(ctx, SHIRO_HIP_INVALID, err);
    const size_t nt = flat.material.size();
    Bvh bvh;
    if (refit) {
        bvh = std::move(ctx->bvh);
        if (!refit_bvh(&bvh, flat.tri.data(), nt, &err)) return fail(ctx, SHIRO_HIP_INVALID, err);
    } else if (!build_bvh(flat.tri.data(), nt, kMaxBvhDepth, &bvh, &err)) {
        return fail(ctx, SHIRO_HIP_INVALID, err);
    }
    if (!validate_bvh(bvh, flat.tri.data(), nt, kMaxBvhDepth, &err)) return fail(ctx, SHIRO_HIP_INVALID, "BVH check: " + err);
    const auto t1 = std::chrono::steady_clock::now();
    // Emissive-triangle lights and the triangle -> light map.
    std::vector<EmissiveEntry> lights;
    float lightSum = 0.0f;
    build_emissive(*s, &lights, &lightSum);
    std::vector<uint32_t> lightOf(nt, 0u);
    for (size_t i = 0; i < lights.size(); ++i)
        if (lights[i].tri < nt) lightOf[lights[i].tri] = static_cast<uint32_t>(i + 1);
    // Triangles and hit data in leaf order.
    std::vector<float> tris(nt * 12), hit(nt * 12);
    for (size_t slot = 0; slot < nt; ++slot) {
        const uint32_t g = bvh.order[slot];
        float* t = &tris[slot * 12];
        float* h = &hit[slot * 12];
        const float* v = &flat.tri[static_cast<size_t>(g) * 9];
        const float* n = &flat.nrm[static_cast<size_t>(g) * 9];
        const uint32_t endFlag = bvh.leafEnd[slot] ? kLeafEndFlag : 0u;
        const uint32_t flags = flat.hasNormals[g] ? kTriHasNormals : 0u;
        for (int k = 0; k < 3; ++k) {
            std::memcpy(t + 4 * k, v + 3 * k, 12);
            std::memcpy(h + 4 * k, n + 3 * k, 12);
        }
        std::memcpy(t + 3, &g, 4);
        std::memcpy(t + 7, &endFlag, 4);
        t[11] = 0.0f;
        std::memcpy(h + 3, &flat.material[g], 4);
        std::memcpy(h + 7, &flags, 4);
        std::memcpy(h + 11, &lightOf[g], 4);
    }
    // Materials (+ the default material used for out-of-range indices).
    std::vector<shiro_hip_material> mats(s->materials, s->materials + s->materialCount);
    mats.push_back(default_material());
    ctx->mats = mats;
    std::vector<MaterialDerived> der(mats.size());
    for (size_t i = 0; i < mats.size(); ++i) der[i] = derive_material(mats[i]);
    assign_shade_classes(mats.data(), der.data(), der.size());
    std::vector<uint8_t> cls(nt);
    for (size_t j = 0; j < nt; ++j) {
        uint32_t mat;
        std::memcpy(&mat, &hit[j * 12 + 3], 4);
        cls[j] = static_cast<uint8_t>(der[std::min<size_t>(mat, der.size() - 1)].shadeClass);
    }
    if (nt) HIPC(upload(ctx->triClass, cls.data(), cls.size(), ctx->stream), "upload triangle classes");
    HIPC(upload(ctx->nodes, bvh.nodes.data(), bvh.nodes.size(), ctx->stream), "upload nodes");
    if (!bvh.nodes8.empty())
        HIPC(upload(ctx->nodes8, bvh.nodes8.data(), bvh.nodes8.size(), ctx->stream), "upload 8-wide nodes");
    HIPC(upload(ctx->tris, tris.data(), tris.size(), ctx->stream), "upload triangles");
    HIPC(upload(ctx->hitdata, hit.data(), hit.size(), ctx->stream), "upload hit data");
    HIPC(upload(ctx->materials, mats.data(), mats.size(), ctx->stream), "upload materials");
    HIPC(upload(ctx->derived, der.data(), der.size(), ctx->stream), "upload materials");
    if (!lights.empty()) {
        std::vector<float> L(lights.size() * 20), cdf(lights.size());
        for (size_t i = 0; i < lights.size(); ++i) {
            const EmissiveEntry& e = lights[i];
            float* q = &L[i * 20];
            std::memcpy(q, e.p, 12);
            q[3] = e.area;
            std::memcpy(q + 4, e.p + 3, 12);
            q[7] = e.weight;
            std::memcpy(q + 8, e.p + 6, 12);
            std::memcpy(q + 11, &e.tri, 4);
            std::memcpy(q + 12, e.normal, 12);
            q[15] = 0.0f;
            std::memcpy(q + 16, e.radiance, 12);
            q[19] = 0.0f;
            cdf[i] = e.cumulative;
        }
        HIPC(upload(ctx->lightList, L.data(), L.size(), ctx->stream), "upload lights");
        HIPC(upload(ctx->lightCdf, cdf.data(), cdf.size(), ctx->stream), "upload lights");
    }
    if (flat.hasUvs) {
        std::vector<float> uv(nt * 6);
        for (size_t slot = 0; slot < nt; ++slot)
            std::memcpy(&uv[slot * 6], &flat.uv[static_cast<size_t>(bvh.order[slot]) * 6], 24);
        HIPC(upload(ctx->uvs, uv.data(), uv.size(), ctx->stream), "upload uvs");
    }
    ctx->tex.clear();
    for (uint32_t i = 0; i < s->textureCount; ++i) ctx->tex.push_back(std::make_unique<TexBuffers>());
    std::vector<DevTexture> table(s->textureCount);
    for (uint32_t i = 0; i < s->textureCount; ++i) {
        const shiro_hip_texture& t = s->textures[i];
        TexBuffers& b = *ctx->tex[i];
        b.w = t.width;
        b.h = t.height;
        b.channels = t.channels;
        HIPC(upload(b.eval, t.texels, b.count(), ctx->stream), "upload texture");
        table[i] = DevTexture{b.eval.as<float>(), b.eval.as<float>(), b.w, b.h, b.channels, 0u};
    }
    if (!table.empty()) HIPC(upload(ctx->texTable, table.data(), table.size(), ctx->stream), "upload textures");
    DevScene d{};
    d.uvs = flat.hasUvs ? ctx->uvs.as<float2>() : nullptr;
    d.triClass = nt ? ctx->triClass.as<uint8_t>() : nullptr;
    d.emissive = lights.empty() ? nullptr : ctx->lightList.as<float4>();
    d.emissiveCdf = lights.empty() ? nullptr : ctx->lightCdf.as<float>();
    d.emissiveCount = static_cast<uint32_t>(lights.size());
    d.emissiveSum = lightSum;
    d.textures = table.empty() ? nullptr : ctx->texTable.as<DevTexture>();
    d.textureCount = s->textureCount;
    d.textured = flat.textured ? 1u : 0u;
    d.plainBsdf = flat.plainBsdf ? 1u : 0u;
    d.nodes = ctx->nodes.as<float4>();
    d.tris = ctx->tris.as<float4>();
    d.hitdata = ctx->hitdata.as<float4>();
    d.materials = ctx->materials.as<shiro_hip_material>();
    d.derived = ctx->derived.as<MaterialDerived>();
    d.nodeCount = static_cast<uint32_t>(bvh.nodeCount());
    d.nodes8 = bvh.nodes8.empty() ? nullptr : ctx->nodes8.as<float4>();
    d.node8Count = static_cast<uint32_t>(bvh.node8Count());
    d.triCount = static_cast<uint32_t>(nt);
    d.materialCount = static_cast<uint32_t>(mats.size());
    uint64_t bytes = ctx->nodes.bytes + (bvh.nodes8.empty() ? 0 : ctx->nodes8.bytes) + ctx->tris.bytes + ctx->hitdata.bytes + ctx->materials.bytes + ctx->derived.bytes;
    if (flat.hasUvs) bytes += ctx->uvs.bytes;
    if (!lights.empty()) bytes += ctx->lightList.bytes + ctx->lightCdf.bytes;
    for (const auto& b : ctx->tex) bytes += b->eval.bytes;
    // Lights.
    d.dirCount = s->distantLightCount;
    ctx->hostDirs.assign(s->distantLightCount, DevDirLight{});
    for (uint32_t i = 0; i < s->distantLightCount; ++i) {
        std::memcpy(ctx->hostDirs[i].direction, s->distantLights[i].direction, 12);
        std::memcpy(ctx->hostDirs[i].radiance, s->distantLights[i].radiance, 12);
    }
    d.domeCount = s->domeLightCount;
    ctx->hostDomes.assign(s->domeLightCount, DevDome{});
    ctx->hostEnvs.assign(s->domeLightCount, DevEnv{});
    if (ctx->env.size() < s->domeLightCount) ctx->env.resize(s->domeLightCount);
    for (uint32_t i = 0; i < s->domeLightCount; ++i) {
        const shiro_hip_dome_light& L = s->domeLights[i];
        DevDome& D = ctx->hostDomes[i];
        std::memcpy(D.radiance, L.radiance, 12);
        std::memcpy(D.right, L.right, 12);
        std::memcpy(D.up, L.up, 12);
        std::memcpy(D.forward, L.forward, 12);
        D.env = -1;
        if (L.pixels && L.width > 0 && L.height > 0) {
            const uint32_t layout = L.layout == SHIRO_HIP_ENV_ANGULAR ? SHIRO_HIP_ENV_ANGULAR : SHIRO_HIP_ENV_LATLONG;
            const size_t np = static_cast<size_t>(L.width) * L.height;
            std::vector<float> tex(np * 4);
            for (size_t p = 0; p < np; ++p) {
                tex[4 * p] = L.pixels[3 * p];
                tex[4 * p + 1] = L.pixels[3 * p + 1];
                tex[4 * p + 2] = L.pixels[3 * p + 2];
                tex[4 * p + 3] = 0.0f;
            }
            std::vector<float> row, cond;
            const bool table = build_env_tables(L.pixels, L.width, L.height, layout, &row, &cond);
            HIPC(upload(ctx->env[i].texels, tex.data(), tex.size(), ctx->stream), "upload env");
            ctx->env[i].w = L.width;
            ctx->env[i].h = L.height;
            ctx->env[i].table = table;
            if (table) {
                std::vector<float> sines(L.height);
                for (uint32_t y = 0; y < L.height; ++y) sines[y] = env_row_sine(y, L.height);
                HIPC(upload(ctx->env[i].rowSin, sines.data(), sines.size(), ctx->stream), "upload env");
                HIPC(upload(ctx->env[i].row, row.data(), row.size(), ctx->stream), "upload env");
                HIPC(upload(ctx->env[i].cond, cond.data(), cond.size(), ctx->stream), "upload env");
                std::vector<uint32_t> rg(kEnvRowGuide + 1u), cg(static_cast<size_t>(L.height) * (kEnvColGuide + 1u));
                build_cdf_guide(row.data(), 0u, L.height, kEnvRowGuide, rg.data());
                for (uint32_t y = 0; y < L.height; ++y) {
                    const uint32_t off = y * (L.width + 1u);
                    build_cdf_guide(cond.data(), off, off + L.width, kEnvColGuide,
                                    cg.data() + static_cast<size_t>(y) * (kEnvColGuide + 1u));
                }
                HIPC(upload(ctx->env[i].rowGuide, rg.data(), rg.size(), ctx->stream), "upload env");
                HIPC(upload(ctx->env[i].condGuide, cg.data(), cg.size(), ctx->stream), "upload env");
            }
            DevEnv& E = ctx->hostEnvs[i];
            E.texels = ctx->env[i].texels.as<float4>();
            E.rowCdf = table ? ctx->env[i].row.as<float>() : nullptr;
            E.condCdf = table ? ctx->env[i].cond.as<float>() : nullptr;
            E.rowGuide = table ? ctx->env[i].rowGuide.as<uint32_t>() : nullptr;
            E.condGuide = table ? ctx->env[i].condGuide.as<uint32_t>() : nullptr;
            E.w = L.width;
            E.h = L.height;
            E.layout = layout;
            E.hasTable = table ? 1u : 0u;
            D.env = static_cast<int32_t>(i);
            bytes += ctx->env[i].texels.bytes + (table ? ctx->env[i].row.bytes + ctx->env[i].cond.bytes +
                                                             ctx->env[i].rowGuide.bytes + ctx->env[i].condGuide.bytes
                                                       : 0);
        }
    }
    HIPC(upload(ctx->dirsBuf, ctx->hostDirs.data(), ctx->hostDirs.size(), ctx->stream), "upload lights");
    HIPC(upload(ctx->domesBuf, ctx->hostDomes.data(), ctx->hostDomes.size(), ctx->stream), "upload lights");
    HIPC(upload(ctx->envsBuf, ctx->hostEnvs.data(), ctx->hostEnvs.size(), ctx->stream), "upload lights");
    d.dirs = ctx->dirsBuf.as<DevDirLight>();
    d.domes = ctx->domesBuf.as<DevDome>();
    d.envs = ctx->envsBuf.as<DevEnv>();
    for (uint32_t i = 0; i < kInlineDirLights && i < ctx->hostDirs.size(); ++i) d.dirIn[i] = ctx->hostDirs[i];
    for (uint32_t i = 0; i < kInlineDomes && i < ctx->hostDomes.size(); ++i) {
        d.domeIn[i] = ctx->hostDomes[i];
        d.envIn[i] = ctx->hostEnvs[i];
    }
    std::memcpy(d.envTop, s->environmentTop, 12);
    std::memcpy(d.envBottom, s->environmentBottom, 12);
    ctx->dev = d;
    ctx->emissive = flat.emissive;
    const auto t2 = std::chrono::steady_clock::now();
    ctx->info.triangleCount = nt;
    ctx->info.bvhNodeCount = bvh.nodeCount();
    ctx->info.bvhLeafCount = bvh.leafCount;
    ctx->info.bvhDepth = bvh.depth;
    ctx->info.bvh8NodeCount = bvh.node8Count();
    ctx->info.bvh8Depth = bvh.depth8;
    ctx->info.bvhStack = bvh.stack;
    ctx->info.bvh8Stack = bvh.stack8;
    ctx->info.maxLeafSize = bvh.maxLeafSize;
    ctx->info.plainBsdf = flat.plainBsdf ? 1u : 0u;
    ctx->info.deviceBytes = bytes;
    ctx->info.buildSeconds = std::chrono::duration<double>(t1 - t0).count();
    ctx->info.uploadSeconds = std::chrono::duration<double>(t2 - t1).count();
    ctx->bvh = std::move(bvh);
    ctx->topology = topo;
    ctx->hasScene = true;
    return SHIRO_HIP_OK;
}

int shiro_hip_upload_scene(shiro_hip_ctx* ctx, const shiro_hip_scene* s) { return upload_scene(ctx, s, false); }

int shiro_hip_refit_scene(shiro_hip_ctx* ctx, const shiro_hip_scene* s) { return upload_scene(ctx, s, true); }

int shiro_hip_update_texture(shiro_hip_ctx* ctx, uint32_t index, const float* texels, uint32_t flags) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    if (index >= ctx->tex.size() || !texels) return fail(ctx, SHIRO_HIP_INVALID, "bad texture index or null texels");
    ++ctx->contentVersion;
    HIPC(hipSetDevice(ctx->device), "set device");
    TexBuffers& b = *ctx->tex[index];
    const size_t bytes = b.count() * sizeof(float);
    const hipMemcpyKind kind = (flags & SHIRO_HIP_TEXELS_DEVICE) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const bool freeze = (flags & SHIRO_HIP_TEXELS_EVAL_ONLY) != 0;
    if (freeze && !b.frozen) {
        // Keep the current texels for sampling.
        HIPC(b.sampling.reserve(bytes), "alloc texture");
        HIPC(hipMemcpyAsync(b.sampling.p, b.eval.p, bytes, hipMemcpyDeviceToDevice, ctx->stream), "freeze texture");
    }
    HIPC(hipMemcpyAsync(b.eval.p, texels, bytes, kind, ctx->stream), "update texture");
    if (freeze != b.frozen) {
        b.frozen = freeze;
        const DevTexture entry{b.eval.as<float>(), freeze ? b.sampling.as<float>() : b.eval.as<float>(), b.w, b.h,
                               b.channels, 0u};
        HIPC(hipMemcpyAsync(ctx->texTable.as<DevTexture>() + index, &entry, sizeof(entry), hipMemcpyHostToDevice,
                            ctx->stream),
             "update texture");
    }
    HIPC(hipStreamSynchronize(ctx->stream), "update texture");
    return SHIRO_HIP_OK;
}

int shiro_hip_scene_info_get(const shiro_hip_ctx* ctx, shiro_hip_scene_info* out) {
    if (!ctx || !out) return SHIRO_HIP_INVALID;
    if (!ctx->hasScene) return SHIRO_HIP_NO_SCENE;
    *out = ctx->info;
    return SHIRO_HIP_OK;
}

int shiro_hip_render_batch(shiro_hip_ctx* ctx, const shiro_hip_camera* camera, const shiro_hip_settings* settings,
                           uint32_t sampleStart, uint32_t sampleCount, const volatile uint8_t* cancel,
                           shiro_hip_frame* out) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!camera || !out || !out->beauty || !out->albedo || !out->normal || !out->depth)
        return fail(ctx, SHIRO_HIP_INVALID, "null camera or frame buffer");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    clear_host_frame(*settings, out);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    if (sampleCount == 0) return SHIRO_HIP_OK;
    HIPC(hipSetDevice(ctx->device), "set device");
    const uint64_t np = static_cast<uint64_t>(settings->width) * settings->height;
    HIPC(ctx->frB.reserve(np * 16), "alloc frame");
    HIPC(ctx->frA.reserve(np * 12), "alloc frame");
    HIPC(ctx->frN.reserve(np * 12), "alloc frame");
    HIPC(ctx->frD.reserve(np * 4), "alloc frame");
    const DevFrame f{ctx->frB.as<float>(), ctx->frA.as<float>(), ctx->frN.as<float>(), ctx->frD.as<float>()};
    rc = enqueue_batch(ctx, *camera, *settings, sampleStart, sampleCount, f, false, cancel);
    if (rc) {
        clear_host_frame(*settings, out);
        return rc;
    }
    rc = finish_timing(ctx);
    if (rc) return rc;
    HIPC(hipMemcpyAsync(out->beauty, f.beauty, np * 16, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipMemcpyAsync(out->albedo, f.albedo, np * 12, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipMemcpyAsync(out->normal, f.normal, np * 12, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipMemcpyAsync(out->depth, f.depth, np * 4, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "readback");
    return SHIRO_HIP_OK;
}

int shiro_hip_render_batch_device(shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
                                  const shiro_hip_settings* settings, uint32_t sampleStart, uint32_t sampleCount,
                                  shiro_hip_frame* out, int synchronize) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!camera || !out || !out->beauty || !out->albedo || !out->normal || !out->depth)
        return fail(ctx, SHIRO_HIP_INVALID, "null camera or frame buffer");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    if (sampleCount == 0) return SHIRO_HIP_OK;
    HIPC(hipSetDevice(ctx->device), "set device");
    const DevFrame f{out->beauty, out->albedo, out->normal, out->depth};
    rc = enqueue_batch(ctx, *camera, *settings, sampleStart, sampleCount, f, false, nullptr);
    if (rc) return rc;
    if (synchronize) return finish_timing(ctx);
    return SHIRO_HIP_OK;
}

int shiro_hip_render_batch_device_noalloc(shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
                                          const shiro_hip_settings* settings, uint32_t sampleStart,
                                          uint32_t sampleCount, shiro_hip_frame* out, int synchronize) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!camera || !out || !out->beauty || !out->albedo || !out->normal || !out->depth)
        return fail(ctx, SHIRO_HIP_INVALID, "null camera or frame buffer");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    if (!ctx->scratch) return fail(ctx, SHIRO_HIP_LAUNCH, "no scratch memory: " + ctx->scratchError);
    if (sampleCount == 0) return SHIRO_HIP_OK;
    HIPC(hipSetDevice(ctx->device), "set device");
    const DevFrame f{out->beauty, out->albedo, out->normal, out->depth};
    shiro_hip_settings st = *settings;
    st.transport = SHIRO_HIP_TRANSPORT_WAVEFRONT;
    rc = enqueue_wavefront(ctx, ctx->dev, *camera, st, sampleStart, sampleCount, f, nullptr, true, ctx->scratch);
    if (rc) return rc;
    if (synchronize) return finish_timing(ctx);
    return SHIRO_HIP_OK;
}

int shiro_hip_accum_reset(shiro_hip_ctx* ctx, uint32_t width, uint32_t height) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    return accum_reset_locked(ctx, width, height);
}

}  // extern "C"

namespace {

// shiro_hip_accum_reset with ctx->mu held.
int accum_reset_locked(shiro_hip_ctx* ctx, uint32_t width, uint32_t height) {
    HIPC(hipSetDevice(ctx->device), "set device");
    const uint64_t np = static_cast<uint64_t>(width) * height;
    HIPC(ctx->acB.reserve(std::max<uint64_t>(np * 16, 16)), "alloc accumulator");
    HIPC(ctx->acA.reserve(std::max<uint64_t>(np * 12, 16)), "alloc accumulator");
    HIPC(ctx->acN.reserve(std::max<uint64_t>(np * 12, 16)), "alloc accumulator");
    HIPC(ctx->acD.reserve(std::max<uint64_t>(np * 4, 16)), "alloc accumulator");
    HIPC(ctx->acC.reserve(std::max<uint64_t>(np * 4, 16)), "alloc accumulator");
    HIPC(hipMemsetAsync(ctx->acB.p, 0, np * 16, ctx->stream), "clear accumulator");
    HIPC(hipMemsetAsync(ctx->acA.p, 0, np * 12, ctx->stream), "clear accumulator");
    HIPC(hipMemsetAsync(ctx->acN.p, 0, np * 12, ctx->stream), "clear accumulator");
    HIPC(hipMemsetAsync(ctx->acD.p, 0, np * 4, ctx->stream), "clear accumulator");
    HIPC(hipMemsetAsync(ctx->acC.p, 0, np * 4, ctx->stream), "clear accumulator");
    HIPC(hipStreamSynchronize(ctx->stream), "clear accumulator");
    ctx->acW = width;
    ctx->acH = height;
    ctx->acTotal = 0;
    return SHIRO_HIP_OK;
}

}  // namespace

extern "C" {

int shiro_hip_accum_add(shiro_hip_ctx* ctx, const shiro_hip_frame* batch, uint32_t sampleCount) {
    if (!ctx) return SHIRO_HIP_INVALID;
    if (!batch || batch->width == 0 || batch->height == 0 || sampleCount == 0) return SHIRO_HIP_OK;
    if (batch->width != ctx->acW || batch->height != ctx->acH) {
        // FrameAccumulator resets on a size change (Renderer.cpp:65-67).
        const int rc = shiro_hip_accum_reset(ctx, batch->width, batch->height);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    const AccumParams A{ctx->acB.as<float>(), ctx->acA.as<float>(), ctx->acN.as<float>(), ctx->acD.as<float>(),
                        ctx->acC.as<uint32_t>()};
    const DevFrame f{batch->beauty, batch->albedo, batch->normal, batch->depth};
    HIPC(launch_accum_add(A, f, sampleCount, ctx->acW * ctx->acH, ctx->stream), "accumulate");
    ctx->acTotal += sampleCount;
    return SHIRO_HIP_OK;
}

int shiro_hip_accum_resolve(shiro_hip_ctx* ctx, shiro_hip_frame* out) {
    if (!ctx || !out) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (out->width != ctx->acW || out->height != ctx->acH) return fail(ctx, SHIRO_HIP_INVALID, "frame size mismatch");
    const AccumParams A{ctx->acB.as<float>(), ctx->acA.as<float>(), ctx->acN.as<float>(), ctx->acD.as<float>(),
                        ctx->acC.as<uint32_t>()};
    const DevFrame f{out->beauty, out->albedo, out->normal, out->depth};
    HIPC(launch_accum_resolve(A, f, ctx->acTotal, ctx->acW * ctx->acH, ctx->stream), "resolve");
    HIPC(hipStreamSynchronize(ctx->stream), "resolve");
    return SHIRO_HIP_OK;
}

int shiro_hip_render_progressive(shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
                                 const shiro_hip_settings* settings, uint32_t sampleStart, uint32_t sampleCount,
                                 uint32_t flags, const volatile uint8_t* cancel, shiro_hip_frame* resolved,
                                 uint32_t* accumulatedSamples) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!camera) return fail(ctx, SHIRO_HIP_INVALID, "null camera");
    if (resolved && (!resolved->beauty || !resolved->albedo || !resolved->normal || !resolved->depth))
        return fail(ctx, SHIRO_HIP_INVALID, "null frame buffer");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (resolved) clear_host_frame(*settings, resolved);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    HIPC(hipSetDevice(ctx->device), "set device");
    const uint32_t w = settings->width, h = settings->height;
    const uint64_t np = static_cast<uint64_t>(w) * h;
    if ((flags & SHIRO_HIP_PROGRESSIVE_RESET) || w != ctx->acW || h != ctx->acH || !ctx->acB.p) {
        rc = accum_reset_locked(ctx, w, h);
        if (rc) return rc;
    }
    HIPC(ctx->frB.reserve(np * 16), "alloc frame");
    HIPC(ctx->frA.reserve(np * 12), "alloc frame");
    HIPC(ctx->frN.reserve(np * 12), "alloc frame");
    HIPC(ctx->frD.reserve(np * 4), "alloc frame");
    const DevFrame f{ctx->frB.as<float>(), ctx->frA.as<float>(), ctx->frN.as<float>(), ctx->frD.as<float>()};
    const AccumParams A{ctx->acB.as<float>(), ctx->acA.as<float>(), ctx->acN.as<float>(), ctx->acD.as<float>(),
                        ctx->acC.as<uint32_t>()};
    if (sampleCount > 0) {
        rc = enqueue_batch(ctx, *camera, *settings, sampleStart, sampleCount, f, false, cancel);
        if (rc) {
            if (resolved) clear_host_frame(*settings, resolved);
            return rc;
        }
        if (cancel && *cancel) {
            (void)hipStreamSynchronize(ctx->stream);
            return fail(ctx, SHIRO_HIP_CANCELLED, "cancelled");
        }
        HIPC(launch_accum_add(A, f, sampleCount, static_cast<uint32_t>(np), ctx->stream), "accumulate");
        ctx->acTotal += sampleCount;
        rc = finish_timing(ctx);
        if (rc) return rc;
    }
    if (accumulatedSamples) *accumulatedSamples = ctx->acTotal;
    if (!resolved) return SHIRO_HIP_OK;
    // The batch frame is consumed: resolve into it and read it back.
    HIPC(launch_accum_resolve(A, f, ctx->acTotal, static_cast<uint32_t>(np), ctx->stream), "resolve");
    HIPC(hipMemcpyAsync(resolved->beauty, f.beauty, np * 16, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipMemcpyAsync(resolved->albedo, f.albedo, np * 12, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipMemcpyAsync(resolved->normal, f.normal, np * 12, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipMemcpyAsync(resolved->depth, f.depth, np * 4, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "readback");
    return SHIRO_HIP_OK;
}

void* shiro_hip_stream(shiro_hip_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int shiro_hip_synchronize(shiro_hip_ctx* ctx) {
    if (!ctx) return SHIRO_HIP_INVALID;
    HIPC(hipStreamSynchronize(ctx->stream), "synchronize");
    return SHIRO_HIP_OK;
}

int shiro_hip_last_render_ms(const shiro_hip_ctx* ctx, float* ms) {
    if (!ctx || !ms) return SHIRO_HIP_INVALID;
    *ms = ctx->lastMs;
    return SHIRO_HIP_OK;
}

int shiro_hip_qmc_fill(shiro_hip_ctx* ctx, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t sampleIndex,
                       uint32_t dim0, uint32_t dimCount, float* out) {
    if (!ctx || !out) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIPC(hipSetDevice(ctx->device), "set device");
    const size_t n = static_cast<size_t>(w) * h * dimCount;
    Tmp o;
    HIPC(o.alloc<float>(n), "alloc");
    HIPC(launch_qmc_fill(x0, y0, w, h, sampleIndex, dim0, dimCount, o.b.as<float>(), ctx->stream), "qmc launch");
    HIPC(o.out(out, n, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "qmc");
    return SHIRO_HIP_OK;
}

int shiro_hip_intersect(shiro_hip_ctx* ctx, const float* rays, uint32_t n, uint32_t* prim, float* t, float* uv) {
    if (!ctx || !rays || !prim || !t || !uv) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    HIPC(hipSetDevice(ctx->device), "set device");
    Tmp r, p, tt, u;
    HIPC(r.in(rays, static_cast<size_t>(n) * 8, ctx->stream), "upload rays");
    HIPC(p.alloc<uint32_t>(n), "alloc");
    HIPC(tt.alloc<float>(n), "alloc");
    HIPC(u.alloc<float>(2ull * n), "alloc");
    HIPC(launch_intersect(ctx->dev, r.b.as<float>(), n, p.b.as<uint32_t>(), tt.b.as<float>(), u.b.as<float>(),
                          ctx->stream),
         "intersect launch");
    HIPC(p.out(prim, n, ctx->stream), "readback");
    HIPC(tt.out(t, n, ctx->stream), "readback");
    HIPC(u.out(uv, 2ull * n, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "intersect");
    return SHIRO_HIP_OK;
}

int shiro_hip_occluded(shiro_hip_ctx* ctx, const float* rays, uint32_t n, uint8_t* out) {
    if (!ctx || !rays || !out) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    HIPC(hipSetDevice(ctx->device), "set device");
    Tmp r, o;
    HIPC(r.in(rays, static_cast<size_t>(n) * 8, ctx->stream), "upload rays");
    HIPC(o.alloc<uint8_t>(n), "alloc");
    HIPC(launch_occluded(ctx->dev, r.b.as<float>(), n, o.b.as<uint8_t>(), ctx->stream), "occluded launch");
    HIPC(o.out(out, n, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "occluded");
    return SHIRO_HIP_OK;
}

int shiro_hip_trace_primary(shiro_hip_ctx* ctx, const shiro_hip_camera* camera, const shiro_hip_settings* settings,
                            uint32_t sampleIndex, uint32_t* prim, float* t) {
    if (!ctx || !camera || !prim || !t) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    HIPC(hipSetDevice(ctx->device), "set device");
    const size_t n = static_cast<size_t>(settings->width) * settings->height;
    Tmp p, tt;
    HIPC(p.alloc<uint32_t>(n), "alloc");
    HIPC(tt.alloc<float>(n), "alloc");
    HIPC(launch_trace_primary(ctx->dev, to_dev(*camera), to_dev(*settings), sampleIndex, p.b.as<uint32_t>(),
                              tt.b.as<float>(), ctx->stream),
         "primary launch");
    HIPC(p.out(prim, n, ctx->stream), "readback");
    HIPC(tt.out(t, n, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "primary");
    return SHIRO_HIP_OK;
}

int shiro_hip_eval_bsdf(shiro_hip_ctx* ctx, const uint32_t* materialIndex, const float* normal, const float* coatNormal,
                        const float* view, const float* light, uint32_t n, float* value, float* pdf) {
    if (!ctx || !materialIndex || !normal || !coatNormal || !view || !light || !value || !pdf) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    for (uint32_t i = 0; i < n; ++i)
        if (materialIndex[i] >= ctx->dev.materialCount) return fail(ctx, SHIRO_HIP_INVALID, "material index out of range");
    HIPC(hipSetDevice(ctx->device), "set device");
    Tmp mi, a, b, c, d, v, p;
    HIPC(mi.in(materialIndex, n, ctx->stream), "upload");
    HIPC(a.in(normal, 3ull * n, ctx->stream), "upload");
    HIPC(b.in(coatNormal, 3ull * n, ctx->stream), "upload");
    HIPC(c.in(view, 3ull * n, ctx->stream), "upload");
    HIPC(d.in(light, 3ull * n, ctx->stream), "upload");
    HIPC(v.alloc<float>(3ull * n), "alloc");
    HIPC(p.alloc<float>(n), "alloc");
    HIPC(launch_eval_bsdf(ctx->dev, mi.b.as<uint32_t>(), a.b.as<float>(), b.b.as<float>(), c.b.as<float>(),
                          d.b.as<float>(), n, v.b.as<float>(), p.b.as<float>(), ctx->stream),
         "bsdf launch");
    HIPC(v.out(value, 3ull * n, ctx->stream), "readback");
    HIPC(p.out(pdf, n, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "bsdf");
    return SHIRO_HIP_OK;
}

int shiro_hip_env_sample(shiro_hip_ctx* ctx, uint32_t domeIndex, const float* u, uint32_t n, float* direction,
                         float* pdf, float* radiance) {
    if (!ctx || !u || !direction || !pdf || !radiance) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    if (domeIndex >= ctx->dev.domeCount || ctx->hostDomes[domeIndex].env < 0)
        return fail(ctx, SHIRO_HIP_INVALID, "dome has no environment map");
    HIPC(hipSetDevice(ctx->device), "set device");
    Tmp uu, d, p, r;
    HIPC(uu.in(u, 2ull * n, ctx->stream), "upload");
    HIPC(d.alloc<float>(3ull * n), "alloc");
    HIPC(p.alloc<float>(n), "alloc");
    HIPC(r.alloc<float>(3ull * n), "alloc");
    HIPC(launch_env_sample(ctx->hostEnvs[ctx->hostDomes[domeIndex].env], uu.b.as<float>(), n, d.b.as<float>(),
                           p.b.as<float>(), r.b.as<float>(), ctx->stream),
         "env launch");
    HIPC(d.out(direction, 3ull * n, ctx->stream), "readback");
    HIPC(p.out(pdf, n, ctx->stream), "readback");
    HIPC(r.out(radiance, 3ull * n, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "env");
    return SHIRO_HIP_OK;
}

int shiro_hip_env_eval(shiro_hip_ctx* ctx, uint32_t domeIndex, const float* direction, uint32_t n, float* pdf,
                       float* radiance) {
    if (!ctx || !direction || !pdf || !radiance) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    if (domeIndex >= ctx->dev.domeCount || ctx->hostDomes[domeIndex].env < 0)
        return fail(ctx, SHIRO_HIP_INVALID, "dome has no environment map");
    HIPC(hipSetDevice(ctx->device), "set device");
    Tmp d, p, r;
    HIPC(d.in(direction, 3ull * n, ctx->stream), "upload");
    HIPC(p.alloc<float>(n), "alloc");
    HIPC(r.alloc<float>(3ull * n), "alloc");
    HIPC(launch_env_eval(ctx->hostEnvs[ctx->hostDomes[domeIndex].env], d.b.as<float>(), n, p.b.as<float>(),
                         r.b.as<float>(), ctx->stream),
         "env launch");
    HIPC(p.out(pdf, n, ctx->stream), "readback");
    HIPC(r.out(radiance, 3ull * n, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "env");
    return SHIRO_HIP_OK;
}

int shiro_hip_set_tuning(shiro_hip_ctx* ctx, const shiro_hip_tuning* t) {
    if (!ctx || !t) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (t->pathsPerPass == 0 || (t->fetchMode > 2u && t->fetchMode != SHIRO_HIP_FETCH_AUTO) || t->traceBlocks == 0 ||
        t->stageBlocks == 0 ||
        t->batchBlocksPerCU == 0 || t->batchBlocksPerCU > 64u || t->shadeOrder > 1u ||
        (t->primaryPackets > 1u && t->primaryPackets != 4u && t->primaryPackets != 16u && t->primaryPackets != 64u) ||
        (t->bvhWidth != 4u && t->bvhWidth != 8u) || t->manyBranches == 0 || t->manyBranches > 16u ||
        t->leafWaitLanes > 64u || t->shadowLeafWaitLanes > 64u ||
        (t->tailDepth > 1024u && t->tailDepth != SHIRO_HIP_TAIL_AUTO) || t->shadowOverlap > 1u)
        return fail(ctx, SHIRO_HIP_INVALID, "invalid tuning");
    ctx->tuning = *t;
    return SHIRO_HIP_OK;
}

int shiro_hip_get_tuning(const shiro_hip_ctx* ctx, shiro_hip_tuning* t) {
    if (!ctx || !t) return SHIRO_HIP_INVALID;
    *t = ctx->tuning;
    return SHIRO_HIP_OK;
}

int shiro_hip_autotune(shiro_hip_ctx* ctx, const shiro_hip_camera* camera, const shiro_hip_settings* settings,
                       uint32_t trialSamples, shiro_hip_tuning* best, char* report, size_t reportSize) {
    if (!ctx || !camera) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    HIPC(hipSetDevice(ctx->device), "set device");
    shiro_hip_settings st = *settings;
    st.transport = SHIRO_HIP_TRANSPORT_WAVEFRONT;
    const uint32_t spp = std::max(1u, trialSamples);
    const uint64_t np = static_cast<uint64_t>(st.width) * st.height;
    HIPC(ctx->frB.reserve(np * 16), "alloc frame");
    HIPC(ctx->frA.reserve(np * 12), "alloc frame");
    HIPC(ctx->frN.reserve(np * 12), "alloc frame");
    HIPC(ctx->frD.reserve(np * 4), "alloc frame");
    const DevFrame f{ctx->frB.as<float>(), ctx->frA.as<float>(), ctx->frN.as<float>(), ctx->frD.as<float>()};
    const shiro_hip_tuning base = ctx->tuning;
    const uint32_t cus = std::max(1u, base.traceBlocks / 32u);
    std::vector<shiro_hip_tuning> cands;
    for (uint32_t paths : {4u << 20, 16u << 20, 64u << 20, 128u << 20}) {
        for (uint32_t mode : {0u, 1u, 2u}) {
            for (uint32_t tb : {16u, 32u, 64u}) {
                if (mode != 2u && tb != 32u) continue;
                shiro_hip_tuning t = base;
                t.pathsPerPass = paths;
                t.fetchMode = mode;
                t.traceBlocks = cus * tb;
                cands.push_back(t);
            }
        }
    }
    std::string log;
    float bestMs = 0.0f;
    shiro_hip_tuning bestT = base;
    // One untimed warm run, then the timed run; appends the report line.
    auto timeCandidate = [&](const shiro_hip_tuning& t, float* ms) -> int {
        ctx->tuning = t;
        for (int rep = 0; rep < 2; ++rep) {
            int r = enqueue_batch(ctx, *camera, st, 0, spp, f, false, nullptr);
            if (!r) r = finish_timing(ctx);
            if (r) {
                ctx->tuning = base;
                return r;
            }
            *ms = ctx->lastMs;
        }
        const uint32_t width = t.bvhWidth == 8u && ctx->dev.nodes8 ? 8u : 4u;
        char line[192];
        std::snprintf(line, sizeof(line),
                      "pathsPerPass %u fetchMode %u traceBlocks %u primaryPackets %u bvhWidth %u tailDepth %u: "
                      "%.3f ms\n",
                      t.pathsPerPass, t.fetchMode, t.traceBlocks, t.primaryPackets, width, t.tailDepth, *ms);
        log += line;
        return 0;
    };
    for (size_t i = 0; i < cands.size(); ++i) {
        float ms = 0.0f;
        rc = timeCandidate(cands[i], &ms);
        if (rc) return rc;
        if (i == 0 || ms < bestMs) {
            bestMs = ms;
            bestT = cands[i];
        }
    }
    // Second stage: the best candidate with camera-ray packets of 1, 4, 16
    // and 64 samples per packet (when packets are on).
    if (bestT.primaryPackets) {
        const shiro_hip_tuning first = bestT;
        for (uint32_t g : {1u, 4u, 16u, 64u}) {
            if (g == first.primaryPackets) continue;
            shiro_hip_tuning t = first;
            t.primaryPackets = g;
            float ms = 0.0f;
            rc = timeCandidate(t, &ms);
            if (rc) return rc;
            if (ms < bestMs) {
                bestMs = ms;
                bestT = t;
            }
        }
    }
    // Third stage: the best candidate with the other tree width, when the
    // scene has an 8-wide tree.
    if (ctx->dev.nodes8) {
        shiro_hip_tuning t = bestT;
        t.bvhWidth = bestT.bvhWidth == 8u ? 4u : 8u;
        float ms = 0.0f;
        rc = timeCandidate(t, &ms);
        if (rc) return rc;
        if (ms < bestMs) {
            bestMs = ms;
            bestT = t;
        }
    }
    // Fourth stage: the bounce from which one persistent kernel finishes
    // the paths still alive (the latency-bound tail of the pass).
    {
        const shiro_hip_tuning first = bestT;
        for (uint32_t d : {0u, 3u, 4u, 5u, 6u, 8u}) {
            if (d >= settings->maxDepth || d == first.tailDepth) continue;
            shiro_hip_tuning t = first;
            t.tailDepth = d;
            float ms = 0.0f;
            rc = timeCandidate(t, &ms);
            if (rc) return rc;
            if (ms < bestMs) {
                bestMs = ms;
                bestT = t;
            }
        }
    }
    ctx->tuning = bestT;
    if (best) *best = bestT;
    if (report && reportSize > 0) std::snprintf(report, reportSize, "%s", log.c_str());
    return SHIRO_HIP_OK;
}

int shiro_hip_count_fetches(shiro_hip_ctx* ctx, const shiro_hip_camera* camera, const shiro_hip_settings* settings,
                            uint32_t sampleStart, uint32_t sampleCount, shiro_hip_fetch_counts* out) {
    if (!ctx || !camera || !out) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    std::memset(out, 0, sizeof(*out));
    if (sampleCount == 0) return SHIRO_HIP_OK;
    HIPC(hipSetDevice(ctx->device), "set device");
    const uint64_t np = static_cast<uint64_t>(settings->width) * settings->height;
    HIPC(ctx->frB.reserve(np * 16), "alloc frame");
    HIPC(ctx->frA.reserve(np * 12), "alloc frame");
    HIPC(ctx->frN.reserve(np * 12), "alloc frame");
    HIPC(ctx->frD.reserve(np * 4), "alloc frame");
    HIPC(hipMemsetAsync(ctx->counters.p, 0, sizeof(DevCounters), ctx->stream), "clear counters");
    const DevFrame f{ctx->frB.as<float>(), ctx->frA.as<float>(), ctx->frN.as<float>(), ctx->frD.as<float>()};
    rc = enqueue_batch(ctx, *camera, *settings, sampleStart, sampleCount, f, true, nullptr);
    if (rc) return rc;
    DevCounters c;
    HIPC(hipMemcpyAsync(&c, ctx->counters.p, sizeof(c), hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "count");
    out->samples = c.samples;
    out->rays = c.rays;
    out->shadowRays = c.shadowRays;
    out->nodeFetches = c.nodeFetches;
    out->triFetches = c.triFetches;
    out->envTaps = c.envTaps;
    out->cdfSteps = c.cdfSteps;
    out->hitFetches = c.hitFetches;
    out->shadowNodeFetches = c.shadowNodeFetches;
    out->shadowTriFetches = c.shadowTriFetches;
    out->primaryRays = c.primaryRays;
    out->primaryNodeFetches = c.primaryNodeFetches;
    out->primaryTriFetches = c.primaryTriFetches;
    return SHIRO_HIP_OK;
}

int shiro_hip_count_wavefront(shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
                              const shiro_hip_settings* settings, uint32_t sampleStart, uint32_t sampleCount,
                              shiro_hip_wavefront_counts* out) {
    if (!ctx || !camera || !out) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    std::memset(out, 0, sizeof(*out));
    out->bvhWidth = ctx->tuning.bvhWidth == 8u && ctx->dev.nodes8 ? 8u : 4u;
    if (sampleCount == 0) return SHIRO_HIP_OK;
    HIPC(hipSetDevice(ctx->device), "set device");
    const uint64_t np = static_cast<uint64_t>(settings->width) * settings->height;
    HIPC(ctx->frB.reserve(np * 16), "alloc frame");
    HIPC(ctx->frA.reserve(np * 12), "alloc frame");
    HIPC(ctx->frN.reserve(np * 12), "alloc frame");
    HIPC(ctx->frD.reserve(np * 4), "alloc frame");
    constexpr size_t kBytes = sizeof(unsigned long long) * kWfCountKinds * kWfCountFields;
    HIPC(ctx->wfCount.reserve(kBytes), "alloc counters");
    HIPC(hipMemsetAsync(ctx->wfCount.p, 0, kBytes, ctx->stream), "clear counters");
    const DevFrame f{ctx->frB.as<float>(), ctx->frA.as<float>(), ctx->frN.as<float>(), ctx->frD.as<float>()};
    ctx->wfCountOn = true;
    rc = enqueue_wavefront(ctx, ctx->dev, *camera, *settings, sampleStart, sampleCount, f, nullptr);
    ctx->wfCountOn = false;
    if (rc) return rc;
    unsigned long long c[kWfCountKinds * kWfCountFields];
    HIPC(hipMemcpyAsync(c, ctx->wfCount.p, kBytes, hipMemcpyDeviceToHost, ctx->stream), "readback");
    HIPC(hipStreamSynchronize(ctx->stream), "count");
    out->samples = np * sampleCount;
    for (int k = 0; k < kWfCountKinds; ++k) {
        const unsigned long long* v = c + k * kWfCountFields;
        out->rays[k] = v[kWfCountRays];
        out->nodeFetches[k] = v[kWfCountNodes];
        out->triFetches[k] = v[kWfCountTris];
        out->waveNodeFetches[k] = v[kWfCountWaveNodes];
        out->waveTriFetches[k] = v[kWfCountWaveTris];
    }
    out->shadowOccluded = c[2 * kWfCountFields + kWfCountOccluded];
    return SHIRO_HIP_OK;
}

int shiro_hip_valu_probe(shiro_hip_ctx* ctx, uint32_t iterations, float* ms, uint64_t* waveInstructions) {
    if (!ctx || !ms || !waveInstructions || iterations == 0) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIPC(hipSetDevice(ctx->device), "set device");
    const uint32_t blocks = ctx->computeUnits * 8u;  // 8 blocks of 4 waves: 32 waves per CU
    Tmp out;
    HIPC(out.alloc<float>(static_cast<size_t>(blocks) * 256u), "alloc probe");
    HIPC(launch_valu_probe(out.b.as<float>(), blocks, 16u, ctx->stream), "probe");  // warm-up
    HIPC(hipEventRecord(ctx->ev0, ctx->stream), "event");
    HIPC(launch_valu_probe(out.b.as<float>(), blocks, iterations, ctx->stream), "probe");
    HIPC(hipEventRecord(ctx->ev1, ctx->stream), "event");
    HIPC(hipEventSynchronize(ctx->ev1), "probe");
    HIPC(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1), "probe");
    *waveInstructions = static_cast<uint64_t>(blocks) * 4u * iterations * 8u;
    return SHIRO_HIP_OK;
}

int shiro_hip_set_stage_timing(shiro_hip_ctx* ctx, uint32_t enable) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIPC(hipSetDevice(ctx->device), "set device");
    if (enable && ctx->stageEv.empty()) {
        constexpr uint32_t kStageEvents = 8192;
        ctx->stageEv.reserve(kStageEvents);
        for (uint32_t i = 0; i < kStageEvents; ++i) {
            hipEvent_t e = nullptr;
            HIPC(hipEventCreate(&e), "stage events");
            ctx->stageEv.push_back(e);
        }
        ctx->stageKind.assign(kStageEvents, 0);
    }
    ctx->stageTiming = enable != 0;
    ctx->stageTimes = shiro_hip_stage_times{};
    return SHIRO_HIP_OK;
}

int shiro_hip_get_stage_times(shiro_hip_ctx* ctx, shiro_hip_stage_times* out) {
    if (!ctx || !out) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = ctx->stageTimes;
    return SHIRO_HIP_OK;
}

int shiro_hip_update_environment(shiro_hip_ctx* ctx, uint32_t dome, const float* texels, uint32_t flags) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    if (dome >= ctx->dev.domeCount || ctx->hostDomes[dome].env < 0 || !texels)
        return fail(ctx, SHIRO_HIP_INVALID, "bad dome index, dome without a map, or null texels");
    ++ctx->contentVersion;
    HIPC(hipSetDevice(ctx->device), "set device");
    EnvBuffers& E = ctx->env[ctx->hostDomes[dome].env];
    const size_t n = static_cast<size_t>(E.w) * E.h;
    const float* rgb = texels;
    if (!(flags & SHIRO_HIP_TEXELS_DEVICE)) {
        HIPC(E.rgb.reserve(n * 12), "alloc env staging");
        HIPC(hipMemcpyAsync(E.rgb.p, texels, n * 12, hipMemcpyHostToDevice, ctx->stream), "upload env texels");
        rgb = E.rgb.as<float>();
    }
    EnvUpdate U{};
    U.rgb = rgb;
    U.texels = E.texels.as<float4>();
    U.w = E.w;
    U.h = E.h;
    U.rebuild = E.table && !(flags & SHIRO_HIP_TEXELS_EVAL_ONLY) ? 1u : 0u;
    if (U.rebuild) {
        HIPC(E.rowW.reserve(static_cast<size_t>(E.h) * 4), "alloc env scratch");
        U.rowSin = E.rowSin.as<float>();
        U.rowW = E.rowW.as<float>();
        U.rowCdf = E.row.as<float>();
        U.cond = E.cond.as<float>();
        U.rowGuide = E.rowGuide.as<uint32_t>();
        U.condGuide = E.condGuide.as<uint32_t>();
    }
    HIPC(launch_env_update(U, ctx->stream), "env update");
    HIPC(hipStreamSynchronize(ctx->stream), "env update");
    return SHIRO_HIP_OK;
}

int render_adjoint(shiro_hip_ctx* ctx, const shiro_hip_camera* camera, const shiro_hip_settings* settings,
                   uint32_t sampleStart, uint32_t sampleCount, const float* dBeauty, float* const* dTexels,
                   uint32_t dTexelsCount, float* const* dEnv, uint32_t dEnvCount) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!camera || !dBeauty) return fail(ctx, SHIRO_HIP_INVALID, "null camera or adjoint image");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!ctx->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "no scene uploaded");
    if (settings->maxDepth > kAdjointMaxDepth)
        return fail(ctx, SHIRO_HIP_INVALID, "adjoint supports maxDepth <= " + std::to_string(kAdjointMaxDepth));
    if (dTexelsCount > ctx->tex.size() || dTexelsCount > kAdjointMaxTextures || (dTexelsCount > 0 && !dTexels))
        return fail(ctx, SHIRO_HIP_INVALID, "more texel gradients than textures");
    if (ctx->tex.size() > kAdjointMaxTextures)
        return fail(ctx, SHIRO_HIP_INVALID, "adjoint supports at most " + std::to_string(kAdjointMaxTextures) +
                                                " textures");
    if (dEnvCount > ctx->dev.domeCount || (dEnvCount > 0 && !dEnv))
        return fail(ctx, SHIRO_HIP_INVALID, "more environment gradients than dome lights");
    if (dEnvCount > static_cast<uint32_t>(kMaxDomes))
        return fail(ctx, SHIRO_HIP_INVALID, "the adjoint differentiates the maps of at most " +
                                                std::to_string(kMaxDomes) + " dome lights");
    for (uint32_t i = 0; i < dEnvCount; ++i) {
        if (dEnv[i] && ctx->hostDomes[i].env < 0)
            return fail(ctx, SHIRO_HIP_INVALID, "environment gradient for dome " + std::to_string(i) + " without a map");
        if (dEnv[i] && static_cast<uint64_t>(ctx->env[ctx->hostDomes[i].env].w) * ctx->env[ctx->hostDomes[i].env].h >=
                           (1ull << 27))
            return fail(ctx, SHIRO_HIP_INVALID, "environment gradient of a map with 2^27 texels or more");
    }
    for (uint32_t i = 0; i < dTexelsCount; ++i)
        if (dTexels[i] && static_cast<uint64_t>(ctx->tex[i]->w) * ctx->tex[i]->h >= (1ull << 27))
            return fail(ctx, SHIRO_HIP_INVALID, "texel gradient of a texture with 2^27 texels or more");
    for (size_t i = 0; i < ctx->mats.size(); ++i) {
        const shiro_hip_material& m = ctx->mats[i];
        if (m.baseColorTexture < 0 && m.roughnessTexture < 0) continue;
        if (m.specularAnisotropy != 0.0f || m.thinFilmThickness > 0.0f || m.sheen > 0.0f || m.transmission > 0.0f)
            return fail(ctx, SHIRO_HIP_INVALID,
                        "material " + std::to_string(i) +
                            " is textured but outside the differentiable class (anisotropy, thin film, sheen or "
                            "transmission)");
    }
    if (sampleCount == 0) return SHIRO_HIP_OK;
    HIPC(hipSetDevice(ctx->device), "set device");
    AdjointParams P{};
    P.scene = ctx->dev;
    P.camera = to_dev(*camera);
    P.settings = to_dev(*settings);
    P.dBeauty = reinterpret_cast<const float4*>(dBeauty);
    for (uint32_t i = 0; i < dTexelsCount; ++i) P.grad[i] = dTexels[i];
    for (uint32_t i = 0; i < dEnvCount; ++i) {
        P.envGrad[i] = dEnv[i];
        if (dEnv[i]) P.envWanted = 1u;
    }
    P.sampleStart = sampleStart;
    P.sampleCount = sampleCount;
    P.tilesX = (settings->width + 7u) / 8u;
    P.tiles = P.tilesX * ((settings->height + 7u) / 8u);
    // Blocks of 16 samples of a 2x2 pixel square (4 of a 4x4 square, or
    // one sample of an 8x8 tile, when the sample count is not a multiple).
    P.groupLog = sampleCount % 16u == 0u ? 4u : (sampleCount % 4u == 0u ? 2u : 0u);
    // Replay the log of the forward pass if it rendered these samples of
    // this frame from the scene content as it is now.
    if (ctx->replayValid && ctx->replayVersion == ctx->contentVersion && ctx->replayStart == sampleStart &&
        ctx->replayCount == sampleCount && std::memcmp(&ctx->replayCamera, camera, sizeof(*camera)) == 0) {
        shiro_hip_settings a = ctx->replaySettings, b = *settings;
        a.transport = b.transport = 0u;
        if (std::memcmp(&a, &b, sizeof(a)) == 0) {
            P.hitLog = ctx->replayHits.as<float4>();
            P.visLog = ctx->replayVis.as<uint32_t>();
        }
    }
    ctx->lastAdjointReplayed = P.hitLog != nullptr;
    HIPC(hipEventRecord(ctx->ev0, ctx->stream), "event");
    HIPC(launch_adjoint(P, ctx->stream), "adjoint launch");
    HIPC(hipEventRecord(ctx->ev1, ctx->stream), "event");
    return finish_timing(ctx);
}

int shiro_hip_render_adjoint(shiro_hip_ctx* ctx, const shiro_hip_camera* camera, const shiro_hip_settings* settings,
                             uint32_t sampleStart, uint32_t sampleCount, const float* dBeauty, float* const* dTexels,
                             uint32_t dTexelsCount) {
    return render_adjoint(ctx, camera, settings, sampleStart, sampleCount, dBeauty, dTexels, dTexelsCount, nullptr, 0);
}

int shiro_hip_render_adjoint_env(shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
                                 const shiro_hip_settings* settings, uint32_t sampleStart, uint32_t sampleCount,
                                 const float* dBeauty, float* const* dTexels, uint32_t dTexelsCount,
                                 float* const* dEnv, uint32_t dEnvCount) {
    return render_adjoint(ctx, camera, settings, sampleStart, sampleCount, dBeauty, dTexels, dTexelsCount, dEnv,
                          dEnvCount);
}

int shiro_hip_set_replay_log(shiro_hip_ctx* ctx, uint64_t maxBytes) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->replayCap = maxBytes;
    ctx->replayValid = false;
    if (maxBytes == 0) {
        ctx->replayHits.release();
        ctx->replayVis.release();
    }
    return SHIRO_HIP_OK;
}

int shiro_hip_last_adjoint_replayed(const shiro_hip_ctx* ctx, uint32_t* replayed) {
    if (!ctx || !replayed) return SHIRO_HIP_INVALID;
    *replayed = ctx->lastAdjointReplayed ? 1u : 0u;
    return SHIRO_HIP_OK;
}

namespace {

// Validates the items and fills the per-item megakernel parameters.
int many_params(shiro_hip_ctx* ctx, const shiro_hip_batch_item* items, uint32_t count, std::vector<RenderParams>* out,
                bool* tex) {
    out->clear();
    *tex = false;
    for (uint32_t i = 0; i < count; ++i) {
        const shiro_hip_batch_item& it = items[i];
        int rc = check_settings(ctx, &it.settings);
        if (rc) return rc;
        if (!it.scene || !it.scene->hasScene) return fail(ctx, SHIRO_HIP_NO_SCENE, "batch item without a scene");
        if (it.scene->device != ctx->device) return fail(ctx, SHIRO_HIP_INVALID, "batch item on another device");
        if (!it.frame.beauty || !it.frame.albedo || !it.frame.normal || !it.frame.depth)
            return fail(ctx, SHIRO_HIP_INVALID, "batch item with a null frame buffer");
        if (it.sampleCount == 0) continue;
        RenderParams P{};
        P.scene = it.scene->dev;
        P.camera = to_dev(it.camera);
        P.settings = to_dev(it.settings);
        P.frame = DevFrame{it.frame.beauty, it.frame.albedo, it.frame.normal, it.frame.depth};
        P.sampleStart = it.sampleStart;
        P.sliceBegin = 0;
        P.sliceCount = it.sampleCount;
        P.batchCount = it.sampleCount;
        P.tilesX = (it.settings.width + 7u) / 8u;
        P.counters = nullptr;
        if (P.scene.textured) *tex = true;
        out->push_back(P);
    }
    return SHIRO_HIP_OK;
}

// FNV-1a over 8-byte words, for the batch graph's key.
struct KeyHash {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        size_t i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            h = (h ^ w) * 1099511628211ull;
        }
        for (; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
};

// The batch as one graph (SHIRO_HIP_MANY_WAVEFRONT_GRAPH).  Items go
// longest first (cost: samples x (4 + BVH depth)) to the branch with the
// least work so far; each branch is a stream that renders its items one
// after the other through the wavefront with its own state (path state of
// at most kWavefrontStateBytes / branches).  The fork/join of the branches
// is captured once into a hipGraph and instantiated; a call on the same
// items (scenes, cameras, settings, frames, tuning, buffer addresses)
// replays it without capturing, and a call on other items captures anew and
// updates the instantiated graph in place when its topology allows, else
// instantiates it again.  The frames equal single renders bit for bit: every
// item runs the same passes as alone, on its own arrays.
int enqueue_many_graph(shiro_hip_ctx* ctx, const std::vector<RenderParams>& P) {
    const uint32_t nItems = static_cast<uint32_t>(P.size());
    const uint32_t B = std::max(1u, std::min(ctx->tuning.manyBranches, nItems));
    while (ctx->branchStreams.size() < B) {
        hipStream_t st = nullptr;
        HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "branch stream");
        ctx->branchStreams.push_back(st);
        hipEvent_t e = nullptr;
        HIPC(hipEventCreateWithFlags(&e, hipEventDisableTiming), "branch event");
        ctx->joinEv.push_back(e);
        ctx->branchStates.push_back(std::make_unique<WfState>());
    }
    if (!ctx->forkEv) HIPC(hipEventCreateWithFlags(&ctx->forkEv, hipEventDisableTiming), "branch event");
    // Longest first onto the least loaded branch.
    std::vector<uint32_t> order(nItems), branch(nItems);
    for (uint32_t i = 0; i < nItems; ++i) order[i] = i;
    auto cost = [&](uint32_t i) {
        const shiro_hip_batch_item& it = *ctx->manyItemsHost[i];
        return static_cast<double>(it.settings.width) * it.settings.height * it.sampleCount *
               (4.0 + static_cast<double>(ctx->manyDepth[i]));
    };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });
    std::vector<double> load(B, 0.0);
    for (uint32_t i : order) {
        const uint32_t b = static_cast<uint32_t>(std::min_element(load.begin(), load.end()) - load.begin());
        branch[i] = b;
        load[b] += cost(i);
    }
    // Every array sized before the capture (nothing is allocated inside it).
    const uint64_t budget = kWavefrontStateBytes / B;
    for (uint32_t i : order) {
        const shiro_hip_batch_item& it = *ctx->manyItemsHost[i];
        WfPlan plan;
        int rc = wavefront_plan(ctx, P[i].scene, it.settings, it.sampleCount, budget, &plan);
        if (rc) return rc;
        WfState& S = *ctx->branchStates[branch[i]];
        for (size_t k = 0; k < kWfArrays; ++k) HIPC(S.a[k].reserve(std::max<size_t>(plan.n[k], 16)), "alloc state");
    }
    KeyHash key;
    key.add(&B, sizeof(B));
    key.add(&ctx->tuning, sizeof(ctx->tuning));
    for (uint32_t i : order) {
        const shiro_hip_batch_item& it = *ctx->manyItemsHost[i];
        key.add(&branch[i], sizeof(uint32_t));
        key.add(&P[i].scene, sizeof(DevScene));
        key.add(&it.camera, sizeof(it.camera));
        key.add(&it.settings, sizeof(it.settings));
        key.add(&it.sampleStart, sizeof(it.sampleStart));
        key.add(&it.sampleCount, sizeof(it.sampleCount));
        key.add(&P[i].frame, sizeof(DevFrame));
    }
    for (uint32_t b = 0; b < B; ++b)
        for (size_t k = 0; k < kWfArrays; ++k) key.add(&ctx->branchStates[b]->a[k].p, sizeof(void*));
    if (!ctx->manyExec || key.h != ctx->manyKey) {
        hipGraph_t graph = nullptr;
        HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "graph capture");
        int rc = SHIRO_HIP_OK;
        hipError_t e = hipEventRecord(ctx->forkEv, ctx->stream);
        for (uint32_t b = 0; b < B && e == hipSuccess; ++b) e = hipStreamWaitEvent(ctx->branchStreams[b], ctx->forkEv, 0);
        for (uint32_t i : order) {
            if (e != hipSuccess || rc) break;
            const shiro_hip_batch_item& it = *ctx->manyItemsHost[i];
            const uint32_t b = branch[i];
            rc = enqueue_wavefront(ctx, P[i].scene, it.camera, it.settings, it.sampleStart, it.sampleCount,
                                   P[i].frame, nullptr, false, nullptr, ctx->branchStreams[b],
                                   ctx->branchStates[b].get(), budget);
        }
        for (uint32_t b = 0; b < B && e == hipSuccess && !rc; ++b) {
            e = hipEventRecord(ctx->joinEv[b], ctx->branchStreams[b]);
            if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->joinEv[b], 0);
        }
        const hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
        if (rc || e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (rc) return rc;
            return hip_fail(ctx, e != hipSuccess ? e : e2, "graph capture");
        }
        ++ctx->manyCaptures;
        bool updated = false;
        if (ctx->manyExec) {
            hipGraphNode_t errNode = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
            updated = hipGraphExecUpdate(ctx->manyExec, graph, &errNode, &res) == hipSuccess &&
                      res == hipGraphExecUpdateSuccess;
            if (updated) {
                ++ctx->manyUpdates;
            } else {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(ctx->manyExec);
                ctx->manyExec = nullptr;
            }
        }
        if (!updated) {
            e = hipGraphInstantiate(&ctx->manyExec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                ctx->manyExec = nullptr;
                (void)hipGraphDestroy(graph);
                return hip_fail(ctx, e, "graph instantiate");
            }
        }
        (void)hipGraphDestroy(graph);
        ctx->manyKey = key.h;
    }
    HIPC(hipEventRecord(ctx->ev0, ctx->stream), "event");
    HIPC(hipGraphLaunch(ctx->manyExec, ctx->stream), "graph launch");
    HIPC(hipEventRecord(ctx->ev1, ctx->stream), "event");
    return SHIRO_HIP_OK;
}

// Enqueues one batched render in `mode` between ev0 and ev1.
int enqueue_many(shiro_hip_ctx* ctx, const std::vector<RenderParams>& P, bool tex, uint32_t mode) {
    if (mode == SHIRO_HIP_MANY_PERSISTENT) {
        // Work units (item, tile), items in decreasing estimated cost
        // (samples per tile times BVH depth), tiles of an item contiguous.
        std::vector<uint32_t> order(P.size());
        for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
        auto cost = [&](uint32_t i) {
            return static_cast<double>(P[i].batchCount) * (4.0 + static_cast<double>(ctx->manyDepth[i]));
        };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost(a) > cost(b); });
        std::vector<uint2> work;
        for (uint32_t i : order) {
            const uint32_t tiles = P[i].tilesX * ((P[i].settings.height + 7u) / 8u);
            for (uint32_t t = 0; t < tiles; ++t) work.push_back(make_uint2(i, t));
        }
        HIPC(upload(ctx->manyItems, P.data(), P.size(), ctx->stream), "upload batch");
        HIPC(upload(ctx->manyWork, work.data(), work.size(), ctx->stream), "upload batch");
        HIPC(ctx->manyCounter.reserve(256), "alloc batch counter");
        const uint32_t blocks = std::max(1u, ctx->computeUnits * std::max(1u, ctx->tuning.batchBlocksPerCU));
        HIPC(hipEventRecord(ctx->ev0, ctx->stream), "event");
        HIPC(launch_render_many(ctx->manyItems.as<RenderParams>(), ctx->manyWork.as<uint2>(),
                                static_cast<uint32_t>(work.size()), ctx->manyCounter.as<uint32_t>(), blocks, tex,
                                ctx->stream),
             "batched launch");
        HIPC(hipEventRecord(ctx->ev1, ctx->stream), "event");
        return SHIRO_HIP_OK;
    }
    if (mode == SHIRO_HIP_MANY_SEQUENTIAL) {
        HIPC(hipEventRecord(ctx->ev0, ctx->stream), "event");
        for (const RenderParams& q : P) HIPC(launch_render(q, false, ctx->stream), "render launch");
        HIPC(hipEventRecord(ctx->ev1, ctx->stream), "event");
        return SHIRO_HIP_OK;
    }
    if (mode == SHIRO_HIP_MANY_WAVEFRONT_GRAPH) return enqueue_many_graph(ctx, P);
    if (mode == SHIRO_HIP_MANY_WAVEFRONT) {
        HIPC(hipEventRecord(ctx->ev0, ctx->stream), "event");
        for (size_t i = 0; i < P.size(); ++i) {
            const shiro_hip_batch_item& it = *ctx->manyItemsHost[i];
            const int rc = enqueue_wavefront(ctx, P[i].scene, it.camera, it.settings, it.sampleStart, it.sampleCount,
                                             P[i].frame, nullptr, false);
            if (rc) return rc;
        }
        HIPC(hipEventRecord(ctx->ev1, ctx->stream), "event");
        return SHIRO_HIP_OK;
    }
    // Graph: capture the per-item launches, instantiate, replay.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "graph capture");
    hipError_t e = hipSuccess;
    for (const RenderParams& q : P) {
        e = launch_render(q, false, ctx->stream);
        if (e != hipSuccess) break;
    }
    const hipError_t e2 = hipStreamEndCapture(ctx->stream, &graph);
    if (e != hipSuccess) return hip_fail(ctx, e, "graph capture");
    if (e2 != hipSuccess) return hip_fail(ctx, e2, "graph capture");
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_fail(ctx, e, "graph instantiate");
    }
    e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e == hipSuccess) e = hipGraphLaunch(exec, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(ctx, e, "graph launch");
    return SHIRO_HIP_OK;
}

}  // namespace

int shiro_hip_render_many(shiro_hip_ctx* ctx, const shiro_hip_batch_item* items, uint32_t count, uint32_t mode) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (count > 0 && !items) return fail(ctx, SHIRO_HIP_INVALID, "items is null");
    if (mode > SHIRO_HIP_MANY_WAVEFRONT_GRAPH) return fail(ctx, SHIRO_HIP_INVALID, "unknown batch mode");
    HIPC(hipSetDevice(ctx->device), "set device");
    std::vector<RenderParams> P;
    bool tex = false;
    int rc = many_params(ctx, items, count, &P, &tex);
    if (rc) return rc;
    ctx->lastMs = 0.0f;
    if (P.empty()) return SHIRO_HIP_OK;
    ctx->manyDepth.clear();
    ctx->manyItemsHost.clear();
    for (uint32_t i = 0; i < count; ++i)
        if (items[i].sampleCount > 0) {
            ctx->manyDepth.push_back(items[i].scene->info.bvhDepth);
            ctx->manyItemsHost.push_back(items + i);
        }
    rc = enqueue_many(ctx, P, tex, mode);
    if (rc) return rc;
    return finish_timing(ctx);
}

int shiro_hip_many_graph_stats(const shiro_hip_ctx* ctx, uint32_t* captures, uint32_t* updates) {
    if (!ctx) return SHIRO_HIP_INVALID;
    if (captures) *captures = ctx->manyCaptures;
    if (updates) *updates = ctx->manyUpdates;
    return SHIRO_HIP_OK;
}

int shiro_hip_autotune_many(shiro_hip_ctx* ctx, const shiro_hip_batch_item* items, uint32_t count,
                            shiro_hip_tuning* best, char* report, size_t reportSize) {
    if (!ctx) return SHIRO_HIP_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (count > 0 && !items) return fail(ctx, SHIRO_HIP_INVALID, "items is null");
    HIPC(hipSetDevice(ctx->device), "set device");
    std::vector<RenderParams> P;
    bool tex = false;
    int rc = many_params(ctx, items, count, &P, &tex);
    if (rc) return rc;
    ctx->manyDepth.clear();
    ctx->manyItemsHost.clear();
    for (uint32_t i = 0; i < count; ++i)
        if (items[i].sampleCount > 0) {
            ctx->manyDepth.push_back(items[i].scene->info.bvhDepth);
            ctx->manyItemsHost.push_back(items + i);
        }
    const shiro_hip_tuning base = ctx->tuning;
    shiro_hip_tuning bestT = base;
    float bestMs = 0.0f;
    std::string log;
    bool first = true;
    for (uint32_t bpc : {1u, 2u, 4u, 8u, 16u, 32u}) {
        ctx->tuning.batchBlocksPerCU = bpc;
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && !P.empty(); ++rep) {
            rc = enqueue_many(ctx, P, tex, SHIRO_HIP_MANY_PERSISTENT);
            if (!rc) rc = finish_timing(ctx);
            if (rc) {
                ctx->tuning = base;
                return rc;
            }
            ms = ctx->lastMs;
        }
        char line[96];
        std::snprintf(line, sizeof(line), "batchBlocksPerCU %u: %.3f ms\n", bpc, ms);
        log += line;
        if (first || ms < bestMs) {
            bestMs = ms;
            bestT = ctx->tuning;
            first = false;
        }
    }
    // Branches of the graph-captured wavefront batch: the first call of a
    // count captures and instantiates, the second replays and is timed.
    ctx->tuning = bestT;
    float bestGraph = 0.0f;
    first = true;
    for (uint32_t branches : {1u, 2u, 3u, 4u, 6u, 8u}) {
        if (branches > std::max<size_t>(1, P.size())) break;
        ctx->tuning.manyBranches = branches;
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && !P.empty(); ++rep) {
            rc = enqueue_many(ctx, P, tex, SHIRO_HIP_MANY_WAVEFRONT_GRAPH);
            if (!rc) rc = finish_timing(ctx);
            if (rc) {
                ctx->tuning = base;
                return rc;
            }
            ms = ctx->lastMs;
        }
        char line[96];
        std::snprintf(line, sizeof(line), "wavefront graph, manyBranches %u: %.3f ms\n", branches, ms);
        log += line;
        if (first || ms < bestGraph) {
            bestGraph = ms;
            bestT.manyBranches = branches;
            first = false;
        }
    }
    // The tail kernel from a few bounces on, with the best branch count.
    for (uint32_t d : {0u, 3u, 4u, 5u, 6u, SHIRO_HIP_TAIL_AUTO}) {
        if (P.empty() || d == bestT.tailDepth) continue;
        ctx->tuning = bestT;
        ctx->tuning.tailDepth = d;
        float ms = 0.0f;
        for (int rep = 0; rep < 2; ++rep) {
            rc = enqueue_many(ctx, P, tex, SHIRO_HIP_MANY_WAVEFRONT_GRAPH);
            if (!rc) rc = finish_timing(ctx);
            if (rc) {
                ctx->tuning = base;
                return rc;
            }
            ms = ctx->lastMs;
        }
        char line[96];
        std::snprintf(line, sizeof(line), "wavefront graph, tailDepth %u: %.3f ms\n", d, ms);
        log += line;
        if (ms < bestGraph) {
            bestGraph = ms;
            bestT.tailDepth = d;
        }
    }
    ctx->tuning = bestT;
    if (best) *best = bestT;
    if (report && reportSize > 0) std::snprintf(report, reportSize, "%s", log.c_str());
    return SHIRO_HIP_OK;
}

}  // extern "C"
