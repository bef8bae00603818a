// Device-resident scene layout shared by the host uploader and the kernels.
//
// HBM layout (all arrays 16-byte aligned, one allocation each):
//   nodes    BVH4 with quantised child boxes, 64 B per node = 4 x float4:
//              [0] origin.xyz (fp32, the union of the child boxes' lo
//                  corner), w = bits: byte a (a = x, y, z) = e_a + 128 with
//                  plane scale 2^e_a; byte 3 = valid-child mask
//              [1] child word (uint bits): internal -> node index,
//                  leaf -> kLeafBit | first triangle slot, empty -> kEmpty
//              [2] quantised planes, one uint per plane kind with byte k
//                  for child k: lo.x, hi.x, lo.y, hi.y
//              [3] lo.z, hi.z, 0, 0
//            A child's box is origin + q * 2^e per axis, rounded outwards
//            on the host.  Two sibling nodes share one 128-byte L2 line.
//   nodes8   the same leaves under 8-wide nodes, 128 B (one L2 line) per
//            node = 8 x float4:
//              [0] origin.xyz, w = exponent bytes + valid-child mask (byte 3)
//              [1], [2] child words 0-3, 4-7
//              [3] x planes: lo children 0-3, lo 4-7, hi 0-3, hi 4-7
//              [4] y planes, [5] z planes (same order), [6], [7] zero
//            Absent (no allocation) when the 8-wide tree's worst-case
//            traversal stack exceeds kStackSize.
//   tris     48 B per triangle in leaf order = 3 x float4:
//              {v0.xyz, global prim id bits}, {v1.xyz, end-of-leaf flag},
//              {v2.xyz, 0}
//            read only by the watertight test during traversal.
//   hitdata  48 B per triangle in leaf order = 3 x float4:
//              {n0.xyz, material index bits}, {n1.xyz, flags},
//              {n2.xyz, emissive-list index + 1 bits (0: not a light)}
//            read once per closest hit (shading normal + material).
//   emissive the emissive-triangle light list (80 B per entry) and its
//            cumulative weights, for light sampling and MIS at hits.
//   materials  shiro_hip_material array (276 B each) + a derived table
//              (MaterialDerived) precomputed on the host.
//   uvs      optional, 24 B per triangle in leaf order (3 x float2), read
//            only at hits on textured materials.
//   textures optional float texels (w x h x channels, row 0 at v = 0).
//   env      per dome: texels as float4 (RGB + pad), rowCdf (H+1),
//            condCdf (H x (W+1)), all fp32; search guides of both (u32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/shiro_hip.h"

namespace shiro_hip {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint32_t kNoHit = 0xffffffffu;
constexpr uint32_t kTriHasNormals = 1u;  // hitdata[1].w flag
constexpr uint32_t kLeafEndFlag = 1u;    // tris[3*slot+1].w: last triangle of its leaf

// Traversal stack capacity per thread; the host rejects BVHs whose depth
// could overflow it (3 pushes per level + 1).
constexpr int kStackSize = 96;
constexpr uint32_t kMaxBvhDepth = (kStackSize - 1) / 3;

constexpr size_t kNodeFloats = 16;   // floats per BVH4 node (64 B)
constexpr size_t kNode8Floats = 32;  // floats per 8-wide node (128 B)

// Buckets of the environment CDF search guides (row table, each column
// table): a sample's lower_bound search covers about three buckets' span
// (c2's 2048x1024 map: 1-2 rows, 6 columns; 4 MB of column guides).
constexpr uint32_t kEnvRowGuide = 4096;
constexpr uint32_t kEnvColGuide = 1024;

// Light counts a scene may hold (a bound on the per-path light samples of
// the wavefront's shadow records, not a storage size: the lights live in
// device arrays).  The adjoint differentiates the maps of the first
// kMaxDomes dome lights.
constexpr uint32_t kMaxSceneDirLights = 1024;
constexpr uint32_t kMaxSceneDomes = 256;
constexpr int kMaxDomes = 4;
// The first lights also ride in the kernel arguments (DevScene), where the
// kernels read them with scalar loads; the device arrays hold them all.
constexpr uint32_t kInlineDirLights = 8;
constexpr uint32_t kInlineDomes = 4;

struct DevEnv {
    const float4* texels;   // w*h, row 0 = theta 0
    const float* rowCdf;    // h+1 (null if no table)
    const float* condCdf;   // h*(w+1)
    const uint32_t* rowGuide;   // kEnvRowGuide + 1 search guide entries (build_cdf_guide)
    const uint32_t* condGuide;  // kEnvColGuide + 1 per row
    uint32_t w, h;
    uint32_t layout;        // SHIRO_HIP_ENV_LATLONG or _ANGULAR
    uint32_t hasTable;
};

struct DevDome {
    float radiance[3], right[3], up[3], forward[3];
    int32_t env;            // index into DevScene::envs, -1 = constant dome
    uint32_t pad;
};

struct DevDirLight {
    float direction[3];     // direction of travel (as given)
    float radiance[3];
};

// Per-material constants computed on the host with the reference's own
// expressions (they depend only on the material): lobe weights and colours.
struct MaterialDerived {
    float diffuseColor[3];
    float specularF0[3];
    float coatF0[3];
    float coatTint[3];
    float wDiffuse, wBaseSpecular, wCoat, wTransmission;
    float baseRoughness;
    float transmissionRoughness;
    float sampleRoughnessBase;   // isotropic AdjustedRoughness(baseRoughness,...)
    float sampleRoughnessCoat;
    float sampleRoughnessTrans;
    float emission[3];           // max(emissionColor*strength, 0)
    float mediumExtinction[3];
    float transmissionTint[3];
    float sampleScaleBase;       // 1 - 0.35 * anisotropy * cos(2 pi rotation), base lobe
    float sampleScaleTrans;      // the same for the transmission lobe
    uint32_t flags;              // kMat* bits
    uint32_t shadeClass;         // materials that take the same branches of the
                                 // shading code share a class (< kShadeClasses - 1)
};

// Shading classes for ordering the wavefront's hit queue: materials are
// grouped by the branches the BSDF and bounce code take; misses get the last
// class.
constexpr uint32_t kShadeClasses = 64;
constexpr uint32_t kShadeMissClass = kShadeClasses - 1u;
// The class sort splits a queue into kOrderBlocks contiguous ranges, one per
// 256-thread block.
constexpr uint32_t kOrderBlocks = 512;
constexpr uint32_t kOrderThreads = 256;

constexpr uint32_t kMatAnisotropic = 1u;   // specular or coat anisotropy != 0
constexpr uint32_t kMatThinFilm = 2u;
constexpr uint32_t kMatSheen = 4u;
constexpr uint32_t kMatNormalOverride = 8u;
constexpr uint32_t kMatCoatNormalOverride = 16u;
constexpr uint32_t kMatTangentOverride = 32u;
constexpr uint32_t kMatEmissive = 64u;
constexpr uint32_t kMatOpacity = 128u;     // opacity < 0.999
constexpr uint32_t kMatTextured = 256u;    // baseColorTexture or roughnessTexture >= 0

// A float texture of the texture extension.  Two copies of the texels: the
// one shading evaluates and the one sampling decisions read (lobe weights,
// sampling roughness, pdfs).  They are the same array unless a caller froze
// the sampling copy (shiro_hip_update_texture with EVAL_ONLY).
struct DevTexture {
    const float* eval;
    const float* sampling;
    uint32_t w, h, channels, pad;
};

struct DevScene {
    const float4* nodes;
    const float4* tris;
    const float4* hitdata;
    const shiro_hip_material* materials;
    const MaterialDerived* derived;
    const float2* uvs;            // 3 per triangle in leaf order, null without UVs
    const uint8_t* triClass;      // shading class of each triangle's material, leaf order
    const float4* emissive;       // 5 per emissive triangle: {p0, area} {p1, weight} {p2, tri bits} {normal} {radiance}
    const float* emissiveCdf;     // cumulative weights (emissiveCount)
    uint32_t emissiveCount;
    float emissiveSum;
    const DevTexture* textures;   // null without textures
    uint32_t textureCount;
    uint32_t textured;            // any material with kMatTextured
    uint32_t plainBsdf;           // no material has anisotropy, thin film, sheen or subsurface wrap
    uint32_t nodeCount;
    const float4* nodes8;         // 8-wide tree over the same leaves, null if absent
    uint32_t node8Count;
    uint32_t triCount;
    uint32_t materialCount;
    uint32_t rootIsLeaf;          // single-leaf scenes: root word in rootWord
    uint32_t rootWord;
    uint32_t rootCount;
    const DevEnv* envs;           // per dome with a map (DevDome::env), device array
    const DevDome* domes;         // domeCount entries, device array
    const DevDirLight* dirs;      // dirCount entries, device array
    DevEnv envIn[kInlineDomes];   // copies of the first entries of those arrays
    DevDome domeIn[kInlineDomes];
    DevDirLight dirIn[kInlineDirLights];
    uint32_t domeCount;
    uint32_t dirCount;
    float envTop[3];
    float envBottom[3];
};

struct DevCamera {
    float position[3], forward[3], right[3], up[3];
    // tan(verticalFovDegrees / 2 in radians), computed once on the host
    // with the float tan of the CPU path (CpuPathTracer.cpp:1636) instead
    // of once per camera ray.
    float tanHalfFov, aspect;
};

struct DevSettings {
    uint32_t width, height;
    uint32_t domeLightSamples;
    uint32_t maxDepth, diffuseDepth, specularDepth;
    uint32_t backgroundVisible;
    uint32_t rrDepth;
};

// Running per-pixel sums of one batch across sample-slice launches.
struct DevBatchSums {
    float4* beauty;     // rgb sum + hit count
    float4* albedo;     // rgb sum
    float4* normal;     // rgb sum
    float* depth;       // running min
};

struct DevFrame {
    float* beauty;      // 4 floats per pixel
    float* albedo;      // 3
    float* normal;      // 3
    float* depth;       // 1
};

// Fetch counters (roofline accounting); all 64-bit.
struct DevCounters {
    unsigned long long samples, rays, shadowRays, nodeFetches, triFetches, envTaps, cdfSteps, hitFetches;
    unsigned long long shadowNodeFetches, shadowTriFetches;
    unsigned long long primaryRays, primaryNodeFetches, primaryTriFetches;
};

struct RenderParams {
    DevScene scene;
    DevCamera camera;
    DevSettings settings;
    DevBatchSums sums;
    DevFrame frame;
    uint32_t sampleStart;     // global index of the first sample of the batch
    uint32_t sliceBegin;      // first sample (relative) of this launch
    uint32_t sliceCount;      // samples in this launch
    uint32_t batchCount;      // samples in the whole batch
    uint32_t tilesX;          // 8x8 tiles per row
    DevCounters* counters;    // null unless counting
};

}  // namespace shiro_hip
