// Host-side declarations of the kernel launchers (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "layout.h"
#include "../../include/shiro_hip.h"

namespace shiro_hip {

// Device FrameAccumulator state (Renderer.cpp:47-58): AoS sums like the
// reference's vectors.
struct AccumParams {
    float* beauty;          // 4 per pixel
    float* albedo;          // 3
    float* normal;          // 3
    float* depthSum;        // 1
    uint32_t* depthCount;   // 1
};

// Striped work counters of the persistent traversal kernels (WavePool,
// wavefront.hip): per stage kPoolStripes counters, each on its own 128-byte
// line.
constexpr uint32_t kPoolStripes = 16;
constexpr uint32_t kPoolStride = 32;
// Passes of at least this many paths take traversal work from per-wave pools
// (fetch mode 2) under SHIRO_HIP_FETCH_AUTO, smaller ones in wave rounds.
constexpr uint64_t kFetchPoolPaths = 8ull << 20;

// Wavefront path state (`cap` paths) and queues.  Ray origin, direction,
// throughput and extinction of path p are one 64-byte record (o, d, T and
// E point at its four float4 slots; entry p of each is at kPathRecord * p):
// after the first bounce the stages visit paths in queue and shading-class
// order, scattered over the pass, and a record costs one 64-byte memory
// sector per path where separate arrays cost one each.  The radiance L is
// an array of its own: the resolve of the first bounce and the
// accumulation read it in path order, four paths to a sector.
constexpr uint32_t kPathRecord = 4;
struct WfBuffers {
    float4 *o, *d, *T, *L, *E, *A, *N, *H;  // ray origin|prevPdf, dir|bits, throughput (the path
                                            // record), radiance, extinction (the record's fourth
                                            // slot), albedo|first-hit t, normal, hit record
    float4 *so, *st;                        // shadow header: origin|count, throughput at the hit|count
    float4* sr;                             // shadow rays (path * maxShadow + index): 2 x float4 each,
                                            // direction|max distance (read by the any-hit kernel) and
                                            // contribution|group (read by the resolve)
    uint32_t *qa, *qb;                      // ray queues (ping-pong)
    uint32_t* qs;                           // paths with shadow rays
    uint32_t* qsr;                          // shadow rays (path * maxShadow + index)
    uint32_t* svis;                         // per shadow ray: 1 if unoccluded
    uint32_t* counts;                       // queue lengths and work counters, see k*Base
    uint32_t countSlots;
    uint32_t shadowRayBase;                 // counts[shadowRayBase + d]: shadow rays at depth d
    uint32_t shadowPathBase;                // counts[shadowPathBase + d]: paths with shadow rays
    uint32_t fetchBase;                     // work fetched by the persistent traversal kernels: kPoolStripes
                                            // counters kPoolStride apart from counts[fetchBase + (k * maxDepth
                                            // + d) * kPoolStripes * kPoolStride], k = 0 closest, 1 any hit
    uint32_t tailBase;                      // counts[tailBase]: paths taken by the tail kernel
    uint32_t maxShadow;
    uint32_t cap;
    uint8_t* qkey;                          // shading class per queue position
    uint32_t* qsorted;                      // the hit queue in class order
    uint32_t* blockCount;                   // [class][block]: hits of the class in the block's range
    uint32_t* blockStart;                   // [class][block]: first slot within the class; [64][..]: totals
    uint32_t shadeOrder;                    // 1: shade depth > 0 hits in class order
};

// Fetch counters of the traversal kernels' measurement instantiations:
// WfParams::wfCount[kind * kWfCountFields + field], kind 0 camera rays
// (k_wf_primary, or k_wf_trace at depth 0), 1 closest-hit rays of later
// bounces, 2 shadow rays.
constexpr int kWfCountKinds = 3;
enum : int {
    kWfCountRays,       // rays traced
    kWfCountNodes,      // nodes read, summed over rays
    kWfCountTris,       // triangles tested, summed over rays
    kWfCountWaveNodes,  // distinct nodes per wave load instruction, summed
    kWfCountWaveTris,   // distinct triangles per wave load instruction, summed
    kWfCountOccluded,   // shadow rays found occluded
    kWfCountFields
};

struct WfParams {
    DevScene scene;
    DevCamera camera;
    DevSettings settings;
    DevBatchSums sums;
    DevFrame frame;
    WfBuffers B;
    uint32_t sampleStart, sliceBegin, sliceCount, batchCount;
    uint32_t tilesX;    // 8x8 tiles per row of the frame
    uint32_t tiles;     // tiles of this pass: frame tiles [tileBase, tileBase + tiles) in row-major order
    uint32_t tileBase;
    uint32_t groupLog;  // path layout: groups of 2^groupLog samples (path_coord, wavefront.hip)
    // Replay log (null: off): the closest hit of every path vertex and the
    // visibility bits of its light samples, at replay_index, for the adjoint
    // to replay the batch without tracing.
    float4* hitLog;
    uint32_t* visLog;
    // Traversal knobs: lanes that may still look for a leaf when a wave
    // turns to its leaf tests (closest hit, any hit).
    uint32_t leafWait, shadowLeafWait;
    // Measurement runs only (null otherwise): fetch counters, see
    // kWfCountFields; the counting instantiations of the traversal kernels
    // run when it is set.
    unsigned long long* wfCount;
};

// Entry of the replay log for vertex `depth` of sample s (of the batch) of
// pixel (x, y): depth-major planes of width * height * samples entries.
__host__ __device__ inline size_t replay_index(uint32_t width, uint32_t height, uint32_t samples, uint32_t x,
                                               uint32_t y, uint32_t s, uint32_t depth) {
    return ((static_cast<size_t>(depth) * samples + s) * height + y) * width + x;
}

// Device memory of the entry points that allocate nothing (wavefront.hip):
// the size of the library's scratch array, and the address of `device`'s
// copy of it (looked up once per device and kept).
constexpr size_t kScratchBytes = size_t{256} << 20;
hipError_t scratch_address(int device, void** ptr);

struct WfLaunch {
    uint32_t maxBlocks;    // grid of the grid-stride stage kernels
    uint32_t traceBlocks;  // grid of the persistent traversal kernels
    uint32_t fetchMode;    // traversal work distribution, see k_wf_trace
    uint32_t primaryPackets;  // samples per path group (1, 4, 16, 64; path_coord), 0: G = 1 and no
                              // camera-ray packets
    uint32_t bvhWidth;     // 4 or 8: tree of the traversal kernels (8 only if the scene has one)
    uint32_t tailDepth;    // bounce from which k_wf_tail finishes the pass (0: never)
    uint32_t tailBlocks;   // grid of k_wf_tail (one-wave blocks)
    // Non-null: the any-hit traversal and resolve of bounce d run on `aux`
    // (after `fork`, recorded on the pass's stream once bounce d is shaded)
    // while the closest hit of bounce d + 1 runs on the pass's stream, which
    // waits on `join` before the next shading.
    hipStream_t aux;
    hipEvent_t fork, join;
};

// Optional per-kernel timing: an event recorded after each launch, tagged
// with the stage (SHIRO_HIP_STAGE_*) whose kernel it closes.
struct StageMarks {
    hipEvent_t* ev = nullptr;
    uint8_t* kind = nullptr;
    uint32_t n = 0, cap = 0;
    bool overflow = false;
    void mark(uint8_t k, hipStream_t s) {
        if (!ev) return;
        if (n == cap) {
            overflow = true;
            return;
        }
        (void)hipEventRecord(ev[n], s);
        kind[n++] = k;
    }
};

// Path-replay adjoint (adjoint.hip).
constexpr uint32_t kAdjointMaxDepth = 16;
constexpr uint32_t kAdjointMaxTextures = 16;
struct AdjointParams {
    DevScene scene;
    DevCamera camera;
    DevSettings settings;
    const float4* dBeauty;                // W*H adjoint of the beauty mean (rgb; a ignored)
    float* grad[kAdjointMaxTextures];     // texel gradients per texture (null: not wanted)
    float* envGrad[kMaxDomes];            // per dome light: gradient of its map's RGB texels (null: not wanted)
    uint32_t envWanted;                   // 1 if any envGrad is set
    uint32_t sampleStart, sampleCount;
    uint32_t tilesX, tiles;
    uint32_t groupLog;  // blocks of 2^groupLog samples of a pixel square, as the wavefront path groups
    // Replay log of the forward render of these samples (null: trace the
    // paths again): hits and light-sample visibility at replay_index.
    const float4* hitLog;
    const uint32_t* visLog;
};

hipError_t launch_render(const RenderParams& P, bool count, hipStream_t stream);
// VALU issue-rate probe (measurement): `blocks` blocks of four waves, each
// thread issuing 8 * iters FMAs; out holds blocks * 256 floats.
hipError_t launch_valu_probe(float* out, uint32_t blocks, uint32_t iters, hipStream_t s);
// The megakernel with path regeneration: `blocks` persistent one-wave
// blocks; counter: one device word (cleared by the launch).
hipError_t launch_render_plus(const RenderParams& P, uint32_t* counter, uint32_t blocks, hipStream_t stream);
// items / work / counter in device memory; work = (item, tile) pairs.
hipError_t launch_render_many(const RenderParams* items, const uint2* work, uint32_t workCount, uint32_t* counter,
                              uint32_t blocks, bool tex, hipStream_t s);
hipError_t launch_adjoint(const AdjointParams& P, hipStream_t s);

// A dome map's texels replaced on the device: rgb (w*h*3, device) into
// texels (float4); with rebuild, the importance tables rebuilt from them.
struct EnvUpdate {
    const float* rgb;
    float4* texels;
    const float* rowSin;   // env_row_sine per row (host values)
    float* rowW;           // h floats of scratch
    float* rowCdf;         // h + 1
    float* cond;           // h * (w + 1)
    uint32_t* rowGuide;    // kEnvRowGuide + 1
    uint32_t* condGuide;   // h * (kEnvColGuide + 1)
    uint32_t w, h, rebuild;
};
hipError_t launch_env_update(const EnvUpdate& U, hipStream_t s);
hipError_t launch_wavefront_pass(const WfParams& P, const WfLaunch& L, hipStream_t s, StageMarks* marks);
hipError_t launch_accum_add(const AccumParams& A, const DevFrame& f, uint32_t count, uint32_t np, hipStream_t s);
hipError_t launch_accum_resolve(const AccumParams& A, const DevFrame& f, uint32_t total, uint32_t np, hipStream_t s);
hipError_t launch_qmc_fill(uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t s, uint32_t dim0,
                           uint32_t dimCount, float* out, hipStream_t st);
hipError_t launch_intersect(const DevScene& S, const float* rays, uint32_t n, uint32_t* prim, float* t, float* uv,
                            hipStream_t s);
hipError_t launch_occluded(const DevScene& S, const float* rays, uint32_t n, uint8_t* out, hipStream_t s);
hipError_t launch_trace_primary(const DevScene& S, const DevCamera& cam, const DevSettings& st, uint32_t sample,
                                uint32_t* prim, float* t, hipStream_t s);
hipError_t launch_eval_bsdf(const DevScene& S, const uint32_t* mi, const float* n, const float* cn, const float* v,
                            const float* l, uint32_t count, float* value, float* pdf, hipStream_t s);
hipError_t launch_env_sample(const DevEnv& e, const float* u, uint32_t n, float* dir, float* pdf, float* rgb,
                             hipStream_t s);
hipError_t launch_env_eval(const DevEnv& e, const float* dir, uint32_t n, float* pdf, float* rgb, hipStream_t s);

}  // namespace shiro_hip
