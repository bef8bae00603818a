/*
 * shiro_hip.h - C ABI of the MI355X (gfx950) path-tracing backend.
 *
 * This is the drop-in boundary for Shiro's GPU backend slot.  Shiro's
 * renderer talks to a backend through the virtual interface
 *   class RenderBackend { GetStatus(); RenderSampleBatch(const RenderRequest&); }
 * (reference include/shiro/backend/RenderBackend.h:12-34), and its GPU
 * backend additionally answers SupportsScene() (include/shiro/backend/optix/
 * OptixBackend.h:16) which Renderer::Impl::ResolveBackend consults
 * (src/render/Renderer.cpp:16-31).  The C++ adapter HipBackend
 * (include/shiro/backend/hip/HipBackend.h) implements that interface on top
 * of the functions below; every other caller (Python/ctypes, tests, bench)
 * binds these symbols directly.
 *
 * Conventions
 *  - Plain C, POD structs, no exceptions.  Every function that can fail
 *    returns a shiro_hip_result; shiro_hip_last_error() gives the message.
 *  - Vectors are float[3] in the reference's Vec3f order (x, y, z).
 *  - Frame buffers are row major, pixel (x, y) at y*width + x, with y = 0
 *    the bottom row (reference CpuPathTracer.cpp:1633-1636).
 *  - A context owns one HIP stream on one device; calls on one context are
 *    serialised by an internal mutex, like the reference backends
 *    (CpuPathTracer.h:26-31, OptixBackend.h:24-26).
 */
#ifndef SHIRO_HIP_H
#define SHIRO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHIRO_HIP_ABI_VERSION 1u

typedef enum shiro_hip_result {
    SHIRO_HIP_OK = 0,
    SHIRO_HIP_NO_DEVICE = 1,   /* no HIP device / init failed (OptixBackend.cpp:1049-1058) */
    SHIRO_HIP_OOM = 2,         /* device allocation failed */
    SHIRO_HIP_CANCELLED = 3,   /* cancel flag observed between launches */
    SHIRO_HIP_INVALID = 4,     /* invalid argument or scene */
    SHIRO_HIP_NO_SCENE = 5,    /* render requested before a scene upload */
    SHIRO_HIP_LAUNCH = 6       /* a kernel launch or HIP runtime call failed */
} shiro_hip_result;

/* Environment map layout, reference Renderer.h:18-22 (EnvironmentMapLayout). */
typedef enum shiro_hip_env_layout {
    SHIRO_HIP_ENV_AUTOMATIC = 0,
    SHIRO_HIP_ENV_LATLONG = 1,
    SHIRO_HIP_ENV_ANGULAR = 2
} shiro_hip_env_layout;

/* Transport form used by render calls. */
typedef enum shiro_hip_transport {
    SHIRO_HIP_TRANSPORT_MEGAKERNEL = 0, /* one fused TracePath per thread */
    SHIRO_HIP_TRANSPORT_WAVEFRONT = 1,  /* queue-driven stage kernels */
    SHIRO_HIP_TRANSPORT_MEGAKERNEL_PLUS = 2 /* the megakernel with path regeneration: persistent lanes
                                               take the next sample or pixel as soon as a path ends,
                                               traversal stacks in LDS rings; same frames */
} shiro_hip_transport;

/* PbrMaterial, reference Renderer.h:56-102; same fields, same order.
 * Booleans are uint32_t (0/1) to keep the layout compiler independent.
 * The two texture indices are an extension for per-texel parameters
 * (-1 = none); the reference has no textures. */
typedef struct shiro_hip_material {
    float baseColor[3];
    float emissionColor[3];
    float specularColor[3];
    float transmissionColor[3];
    float transmissionScatter[3];
    float coatColor[3];
    float subsurfaceColor[3];
    float subsurfaceRadius[3];
    float sheenColor[3];
    float normalOverride[3];
    float coatNormalOverride[3];
    float tangentOverride[3];
    float baseWeight;
    float emissionStrength;
    float specularWeight;
    float metallic;
    float roughness;
    float diffuseRoughness;
    float opacity;
    float specularAnisotropy;
    float specularRotation;
    float coatWeight;
    float coatRoughness;
    float coatIor;
    float coatAnisotropy;
    float coatRotation;
    float sheen;
    float sheenRoughness;
    float subsurface;
    float subsurfaceScale;
    float subsurfaceAnisotropy;
    float transmission;
    float transmissionDepth;
    float transmissionScatterAnisotropy;
    float transmissionDispersion;
    float transmissionExtraRoughness;
    float ior;
    float coatAffectColor;
    float coatAffectRoughness;
    float thinFilmThickness;
    float thinFilmIor;
    uint32_t thinWalled;
    uint32_t hasNormalOverride;
    uint32_t hasCoatNormalOverride;
    uint32_t hasTangentOverride;
    int32_t baseColorTexture;
    int32_t roughnessTexture;
} shiro_hip_material;

/* TriangleMesh, reference Renderer.h:104-109.  normals are used only when
 * normalCount == vertexCount (CpuPathTracer.cpp:148-151).  uvs (2 floats per
 * vertex) are an extension used only by textured materials.  Triangles are
 * read as consecutive index triples; a trailing partial triple is ignored
 * (CpuPathTracer.cpp:434). */
typedef struct shiro_hip_mesh {
    const float* positions;
    const float* normals;
    const float* uvs;
    const uint32_t* indices;
    uint32_t vertexCount;
    uint32_t normalCount;
    uint32_t indexCount;
    uint32_t materialIndex;
} shiro_hip_mesh;

/* DirectionalLight, reference Renderer.h:111-114: direction of travel. */
typedef struct shiro_hip_directional_light {
    float direction[3];
    float radiance[3];
} shiro_hip_directional_light;

/* DomeLight, reference Renderer.h:116-124.  pixels == NULL means a dome
 * without an environment texture (constant radiance, uniform sampling,
 * CpuPathTracer.cpp:327-384).  pixels holds width*height RGB triples, row 0
 * is the top of the map (theta = 0). */
typedef struct shiro_hip_dome_light {
    float radiance[3];
    float right[3];
    float up[3];
    float forward[3];
    const float* pixels;
    uint32_t width;
    uint32_t height;
    uint32_t layout; /* shiro_hip_env_layout */
    uint32_t reserved;
} shiro_hip_dome_light;

/* Float texture of the texture extension (no counterpart in the reference).
 * texels[(y * width + x) * channels + c], row 0 at v = 0.  A textured
 * material replaces baseColor by channels 0..2 of its baseColorTexture (a
 * one-channel texture gives grey) and roughness by channel 0 of its
 * roughnessTexture, looked up at the hit's interpolated UV with bilinear
 * filtering (texel centres at (i + 0.5) / size) and repeat addressing; every
 * quantity the material derives from them (diffuse colour, F0, lobe weights,
 * roughnesses) follows per hit. */
typedef struct shiro_hip_texture {
    const float* texels;
    uint32_t width;
    uint32_t height;
    uint32_t channels; /* 1 to 4 */
    uint32_t reserved;
} shiro_hip_texture;

/* Scene, reference Renderer.h:126-133. */
typedef struct shiro_hip_scene {
    const shiro_hip_material* materials;
    uint32_t materialCount;
    const shiro_hip_mesh* meshes;
    uint32_t meshCount;
    const shiro_hip_directional_light* distantLights;
    uint32_t distantLightCount;
    const shiro_hip_dome_light* domeLights;
    uint32_t domeLightCount;
    float environmentTop[3];
    float environmentBottom[3];
    const shiro_hip_texture* textures;
    uint32_t textureCount;
} shiro_hip_scene;

/* Camera, reference Renderer.h:47-54. */
typedef struct shiro_hip_camera {
    float position[3];
    float forward[3];
    float right[3];
    float up[3];
    float verticalFovDegrees;
    float aspectRatio;
} shiro_hip_camera;

/* RenderSettings, reference Renderer.h:29-45 (same fields and defaults;
 * see shiro_hip_default_settings).  russianRouletteDepth is an extension:
 * the reference starts roulette at depth 2 (CpuPathTracer.cpp:2006), which is
 * the default here.  transport selects the kernel form. */
typedef struct shiro_hip_settings {
    uint32_t width;
    uint32_t height;
    uint32_t samplesPerPixel;
    uint32_t samplesPerUpdate;
    uint32_t domeLightSamples;
    uint32_t maxDepth;
    uint32_t diffuseDepth;
    uint32_t specularDepth;
    uint32_t threadLimit;
    uint32_t maxSubdivLevel;
    uint32_t backend;
    uint32_t backgroundVisible;
    uint32_t enableHeadlight;
    uint32_t enableCaustics;
    uint32_t enableSss;
    uint32_t russianRouletteDepth;
    uint32_t transport;
    uint32_t reserved;
} shiro_hip_settings;

/* FrameBuffer, reference FrameBuffer.h:10-37: AoS beauty RGBA (float4),
 * albedo RGB, normal RGB (0.5*n + 0.5), depth (min hit distance, +inf when
 * no sample hit).  Pointers are host or device memory depending on the call. */
typedef struct shiro_hip_frame {
    float* beauty;
    float* albedo;
    float* normal;
    float* depth;
    uint32_t width;
    uint32_t height;
} shiro_hip_frame;

/* BackendStatus, reference RenderBackend.h:21-26. */
typedef struct shiro_hip_status {
    uint32_t kind;       /* BackendKind::Gpu = 1 */
    const char* name;    /* "hip" */
    uint32_t available;
    uint32_t usesGpu;
    char deviceName[64];
    uint32_t computeUnits;
} shiro_hip_status;

/* Opaque backend context: device, stream, device-resident scene. */
typedef struct shiro_hip_ctx shiro_hip_ctx;

/* Fill settings with the reference defaults (Renderer.h:29-45). */
void shiro_hip_default_settings(shiro_hip_settings* settings);
/* Fill a material with the reference defaults (Renderer.h:56-102). */
void shiro_hip_default_material(shiro_hip_material* material);

/* Context creation; replaces OptixBackend::EnsureRuntime's device/stream
 * set-up (OptixBackend.cpp:1071-1139). */
int shiro_hip_create(int device, shiro_hip_ctx** out);
void shiro_hip_destroy(shiro_hip_ctx* ctx);
const char* shiro_hip_last_error(const shiro_hip_ctx* ctx);

/* GetStatus, RenderBackend.h:30 / OptixBackend.cpp:1035-1060.  Works with
 * ctx == NULL (probes the device without creating a context). */
int shiro_hip_get_status(const shiro_hip_ctx* ctx, shiro_hip_status* out);

/* SupportsScene, OptixBackend.h:16: 1 if the scene can be rendered. */
int shiro_hip_supports_scene(const shiro_hip_scene* scene);

/* Scene upload: flattens meshes, builds the BVH4 on the host (binned SAH),
 * builds dome importance tables and uploads everything.  Replaces
 * OptixBackend::UploadSceneData + GAS build (OptixBackend.cpp:751-1026) and
 * BuildAccelerationScene (CpuPathTracer.cpp:477-598).  The scene is copied;
 * the caller may free it after the call. */
int shiro_hip_upload_scene(shiro_hip_ctx* ctx, const shiro_hip_scene* scene);

/* Re-uploads a scene whose meshes and index arrays are those of the
 * uploaded one (vertex positions, normals, UVs, materials, lights and domes
 * may change) and refits the BVH to the moved vertices, keeping its tree,
 * instead of building a new one.  The reference rebuilds on any change of
 * its scene signature (OptixBackend.cpp:150-232, CpuPathTracer.cpp:2030-2041);
 * for deforming meshes this skips the SAH build (c3: 1.3 s).  Returns
 * SHIRO_HIP_INVALID if the topology differs.  The tree's quality degrades
 * with large deformations; upload anew to rebuild it. */
int shiro_hip_refit_scene(shiro_hip_ctx* ctx, const shiro_hip_scene* scene);

/* RenderSampleBatch, RenderBackend.h:33 / CpuPathTracer.cpp:2043-2170:
 * renders samples [sampleStart, sampleStart + sampleCount) of every pixel and
 * writes the batch means into host buffers `out` (sized settings->width x
 * settings->height).  cancel (may be NULL) is a byte polled between launches
 * -- the object representation of the reference's std::atomic<bool> cancel
 * flag (RenderBackend.h:18) -- and when it becomes non-zero the call returns
 * SHIRO_HIP_CANCELLED and `out` holds a cleared frame. */
int shiro_hip_render_batch(
    shiro_hip_ctx* ctx,
    const shiro_hip_camera* camera,
    const shiro_hip_settings* settings,
    uint32_t sampleStart,
    uint32_t sampleCount,
    const volatile uint8_t* cancel,
    shiro_hip_frame* out);

/* Same as shiro_hip_render_batch, but `out` holds device pointers and the
 * work is enqueued on the context stream without a host copy.  When
 * `synchronize` is non-zero the call waits for completion. */
int shiro_hip_render_batch_device(
    shiro_hip_ctx* ctx,
    const shiro_hip_camera* camera,
    const shiro_hip_settings* settings,
    uint32_t sampleStart,
    uint32_t sampleCount,
    shiro_hip_frame* out,
    int synchronize);

/* No counterpart in the reference: shiro_hip_render_batch_device for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Renders with the
 * wavefront transport (the same frame, bit for bit, as either transport of
 * shiro_hip_render_batch_device) and calls no hipMalloc or hipFree: the path
 * state, queues, counters and running sums of its passes live in the
 * library's scratch array, 256 MiB of device memory per GPU (held from the
 * first context on that GPU until the process ends), and a frame whose
 * state does not fit is covered in passes over ranges of 8x8 tiles.  The
 * scene must be uploaded (upload allocates) and `out` holds device
 * pointers.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (from two contexts or two threads' streams) race, so
 * serialise them.  The counters are cleared by a kernel at every pass, so a
 * call does not depend on what an earlier call left there.  Returns
 * SHIRO_HIP_INVALID when not even one tile of path state fits the array. */
int shiro_hip_render_batch_device_noalloc(
    shiro_hip_ctx* ctx,
    const shiro_hip_camera* camera,
    const shiro_hip_settings* settings,
    uint32_t sampleStart,
    uint32_t sampleCount,
    shiro_hip_frame* out,
    int synchronize);

/* Device-side FrameAccumulator (Renderer.cpp:47-110).  accumulate adds the
 * batch frame held in device buffers `batch` with weight sampleCount;
 * resolve writes the resolved frame into device buffers `out`. */
int shiro_hip_accum_reset(shiro_hip_ctx* ctx, uint32_t width, uint32_t height);
int shiro_hip_accum_add(shiro_hip_ctx* ctx, const shiro_hip_frame* batch, uint32_t sampleCount);
int shiro_hip_accum_resolve(shiro_hip_ctx* ctx, shiro_hip_frame* out);

/* Progressive batch with the accumulator on the device: replaces the Hydra
 * worker's RenderSampleBatch -> FrameAccumulator::Accumulate -> Resolve
 * sequence (RenderParam.cpp:789-803 over Renderer.cpp:47-110) for a GPU
 * backend.  Renders samples [sampleStart, sampleStart + sampleCount) into
 * the context's device frame, adds them to the device accumulator with
 * weight sampleCount (cleared first with SHIRO_HIP_PROGRESSIVE_RESET, or
 * when the frame size changed, as Accumulate resets on a size change) and,
 * when `resolved` is non-NULL, writes the resolved running frame (the
 * accumulator's means; depth = mean of the batch minima over the batches
 * that hit) into the host buffers `resolved`.  A NULL `resolved` skips the
 * readback (the display is not refreshed for this batch).  A cancelled
 * batch (cancel as in shiro_hip_render_batch) is not accumulated and the
 * call returns SHIRO_HIP_CANCELLED.  accumulatedSamples (may be NULL)
 * receives the accumulator's sample count after the call. */
enum { SHIRO_HIP_PROGRESSIVE_RESET = 1 };
int shiro_hip_render_progressive(
    shiro_hip_ctx* ctx,
    const shiro_hip_camera* camera,
    const shiro_hip_settings* settings,
    uint32_t sampleStart,
    uint32_t sampleCount,
    uint32_t flags,
    const volatile uint8_t* cancel,
    shiro_hip_frame* resolved,
    uint32_t* accumulatedSamples);

/* Stream and timing.  last_render_ms is the HIP-event time of the most
 * recent render call's kernels on the context stream. */
void* shiro_hip_stream(shiro_hip_ctx* ctx);
int shiro_hip_synchronize(shiro_hip_ctx* ctx);
int shiro_hip_last_render_ms(const shiro_hip_ctx* ctx, float* ms);

/* Scene statistics after upload (triangles, BVH nodes, depth, bytes). */
typedef struct shiro_hip_scene_info {
    uint64_t triangleCount;
    uint64_t bvhNodeCount;
    uint64_t bvhLeafCount;
    uint32_t bvhDepth;
    uint32_t maxLeafSize;
    uint64_t deviceBytes;
    double buildSeconds;
    double uploadSeconds;
    uint64_t bvh8NodeCount;  /* 8-wide tree over the same leaves (0: not built, see bvh8Stack) */
    uint32_t bvh8Depth;
    uint32_t bvhStack;       /* most traversal stack entries of the 4-wide tree */
    uint32_t bvh8Stack;      /* of the 8-wide tree; above 95 it is not uploaded */
    uint32_t plainBsdf;      /* 1: no material needs anisotropy, thin film, sheen or subsurface
                                wrap, so the wavefront shades with the BSDF code that leaves them out */
} shiro_hip_scene_info;
/* Replaces the texels of texture `index` of the uploaded scene (same size
 * and channel count) without rebuilding anything.  flags:
 * SHIRO_HIP_TEXELS_DEVICE: `texels` is device memory (e.g. a torch tensor);
 * SHIRO_HIP_TEXELS_EVAL_ONLY: only the copy that shading evaluates changes,
 * sampling decisions (lobe weights, sampling roughness, pdfs) keep reading
 * the previous texels - a frozen-sampling render whose finite differences
 * the adjoint reproduces.  Synchronous.  No counterpart in the reference. */
enum { SHIRO_HIP_TEXELS_DEVICE = 1, SHIRO_HIP_TEXELS_EVAL_ONLY = 2 };
int shiro_hip_update_texture(
    shiro_hip_ctx* ctx, uint32_t index, const float* texels, uint32_t flags);

/* Path-replay adjoint of a render with respect to texture texels (texture
 * extension; no counterpart in the reference).  Replays the paths of
 * samples [sampleStart, sampleStart + sampleCount) of the uploaded scene,
 * camera and settings (maxDepth <= 16) and ADDS to dTexels[i] (device
 * memory, same layout as texture i's texels; NULL entries or i >=
 * dTexelsCount are skipped) the gradient sum_{pixel, c} dBeauty[pixel][c] *
 * d beauty[pixel][c] / d texel, where beauty is the batch mean the render
 * returns and dBeauty (device, W*H*4 floats, alpha ignored) is the caller's
 * adjoint image.  Sampling is detached (lobe choices, directions, pdfs and
 * MIS weights are constants), so this is the exact gradient of the frozen-
 * sampling render (SHIRO_HIP_TEXELS_EVAL_ONLY) at the current texels.
 * Textured materials must have no specular anisotropy, thin film, sheen or
 * transmission; textures with gradients hold fewer than 2^27 texels.
 * Synchronous. */
int shiro_hip_render_adjoint(
    shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
    const shiro_hip_settings* settings, uint32_t sampleStart,
    uint32_t sampleCount, const float* dBeauty, float* const* dTexels,
    uint32_t dTexelsCount);
/* shiro_hip_render_adjoint that also ADDS to dEnv[i] (device, W*H*3 floats
 * in the layout of dome light i's pixels; NULL entries skipped) the gradient
 * with respect to dome i's map texels.  The maps enter the render through
 * the radiance escaped rays and dome light samples read (bilinear lookups);
 * their importance tables are sampling and are detached, so this is the
 * gradient of the render whose tables stay those of the current texels
 * (shiro_hip_update_environment with SHIRO_HIP_TEXELS_EVAL_ONLY).  Maps with
 * gradients hold fewer than 2^27 texels. */
int shiro_hip_render_adjoint_env(
    shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
    const shiro_hip_settings* settings, uint32_t sampleStart,
    uint32_t sampleCount, const float* dBeauty, float* const* dTexels,
    uint32_t dTexelsCount, float* const* dEnv, uint32_t dEnvCount);
/* Replay log for the adjoint (no counterpart in the reference): with
 * maxBytes > 0, every wavefront render batch whose log fits in maxBytes
 * (20 bytes per pixel, sample and depth of maxDepth) and whose vertices
 * take at most 32 light samples records the closest hit of each path vertex
 * and which of its light samples were visible.  An adjoint of the same
 * samples, camera and settings (the transport aside) with the scene content
 * unchanged since (no upload, refit, texture or map update in between)
 * reads them instead of tracing the rays again; its gradients are those of
 * the traced replay.  0 turns the log off and frees it. */
int shiro_hip_set_replay_log(shiro_hip_ctx* ctx, uint64_t maxBytes);
/* *replayed = 1 if the last adjoint call read the replay log, else 0. */
int shiro_hip_last_adjoint_replayed(const shiro_hip_ctx* ctx, uint32_t* replayed);
/* Replaces the texels of dome light i's map (W*H*3 floats in the layout of
 * its pixels; host memory, or device memory with SHIRO_HIP_TEXELS_DEVICE)
 * without re-uploading the scene, the EnvironmentMap of a DomeLight changing
 * between renders (no counterpart in the reference, which reloads the map).
 * The importance tables are rebuilt on the device with the host build's
 * operation order (the same map samples exactly as if uploaded), unless
 * SHIRO_HIP_TEXELS_EVAL_ONLY keeps the current tables (frozen sampling, for
 * finite differences of the adjoint).  Synchronous. */
int shiro_hip_update_environment(shiro_hip_ctx* ctx, uint32_t dome, const float* texels, uint32_t flags);

int shiro_hip_scene_info_get(const shiro_hip_ctx* ctx, shiro_hip_scene_info* out);

/* ---- Entry points exposing single stages of the hot path (device-side
 * counterparts of the reference functions named; inputs/outputs are host
 * arrays, n elements). ---- */

/* OpenQmcSampler fallback LowDiscrepancySample (OpenQmcSampler.cpp:56-71)
 * over pixels [x0,x0+w) x [y0,y0+h), dims [dim0, dim0+dimCount) for one
 * sample index; out[((y-y0)*w + (x-x0))*dimCount + d]. */
int shiro_hip_qmc_fill(
    shiro_hip_ctx* ctx,
    uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
    uint32_t sampleIndex, uint32_t dim0, uint32_t dimCount,
    float* out);

/* Closest hit of arbitrary rays (IntersectScene, CpuPathTracer.cpp:684-695).
 * rays: n x 8 floats (origin xyz, direction xyz, tMin, tMax).  prim receives
 * the global triangle index (meshes concatenated in order, 0xffffffff on a
 * miss), t the hit distance, uv the barycentrics of vertices 1 and 2. */
int shiro_hip_intersect(
    shiro_hip_ctx* ctx, const float* rays, uint32_t n,
    uint32_t* prim, float* t, float* uv);

/* Any-hit test (Occluded, CpuPathTracer.cpp:697-709): out[i] = 1 if hit. */
int shiro_hip_occluded(
    shiro_hip_ctx* ctx, const float* rays, uint32_t n, uint8_t* out);

/* Primary visibility: camera ray for `sampleIndex` at every pixel
 * (GenerateCameraRay, CpuPathTracer.cpp:1633-1648) and its closest hit. */
int shiro_hip_trace_primary(
    shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
    const shiro_hip_settings* settings, uint32_t sampleIndex,
    uint32_t* prim, float* t);

/* EvaluateBsdf / EvaluateBsdfPdf (CpuPathTracer.cpp:1232-1354) for n
 * tuples: material index into the uploaded scene, shading normal, coat
 * normal, view and light directions (3 floats each). */
int shiro_hip_eval_bsdf(
    shiro_hip_ctx* ctx, const uint32_t* materialIndex,
    const float* normal, const float* coatNormal,
    const float* view, const float* light, uint32_t n,
    float* value, float* pdf);

/* Dome light i of the uploaded scene: EnvironmentMap::Sample
 * (EnvironmentMap.cpp:109-163) for n sample pairs u (2 floats each), giving
 * local direction, pdf and radiance; and EnvironmentMap::Pdf / Evaluate
 * (:94-107, :165-183) of local directions. */
int shiro_hip_env_sample(
    shiro_hip_ctx* ctx, uint32_t domeIndex, const float* u, uint32_t n,
    float* direction, float* pdf, float* radiance);
int shiro_hip_env_eval(
    shiro_hip_ctx* ctx, uint32_t domeIndex, const float* direction,
    uint32_t n, float* pdf, float* radiance);

/* Traversal counters for the roofline: runs the render once with fetch
 * counting enabled and returns totals over the frame. */
typedef struct shiro_hip_fetch_counts {
    uint64_t samples;
    uint64_t rays;          /* closest-hit rays */
    uint64_t shadowRays;    /* any-hit rays */
    uint64_t nodeFetches;   /* BVH4 nodes read by closest-hit rays (128 B each) */
    uint64_t triFetches;    /* triangles tested by closest-hit rays (48 B each) */
    uint64_t envTaps;       /* environment texel reads (16 B each) */
    uint64_t cdfSteps;      /* importance table reads (4 B each) */
    uint64_t hitFetches;    /* per-hit vertex/normal/material fetches */
    uint64_t shadowNodeFetches;  /* BVH4 nodes read by any-hit rays */
    uint64_t shadowTriFetches;   /* triangles tested by any-hit rays */
    uint64_t primaryRays;        /* camera rays: the depth-0 part of rays, */
    uint64_t primaryNodeFetches; /* of nodeFetches */
    uint64_t primaryTriFetches;  /* and of triFetches */
} shiro_hip_fetch_counts;
int shiro_hip_count_fetches(
    shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
    const shiro_hip_settings* settings, uint32_t sampleStart,
    uint32_t sampleCount, shiro_hip_fetch_counts* out);

/* Fetch counters of the wavefront transport's traversal kernels (no
 * counterpart in the reference; measurement): renders the samples once
 * through the wavefront with the context's tuning, running the counting
 * instantiations of the traversal kernels (same traversal, same work
 * distribution and tail kernel: the rays the tail kernel traces are not
 * counted), and returns per kernel kind -- [0] camera rays
 * (k_wf_primary's packets, or the depth-0 k_wf_trace), [1] closest-hit rays
 * of the later bounces (k_wf_trace<false>), [2] shadow rays
 * (k_wf_trace<true>) -- the rays, the nodes and triangles read summed over
 * rays, and the distinct nodes and triangles per wave load instruction
 * summed over instructions (a camera packet reads each node once for its
 * 64 rays); for shadow rays also those found occluded. */
typedef struct shiro_hip_wavefront_counts {
    uint64_t samples;
    uint64_t rays[3];
    uint64_t nodeFetches[3];      /* 64 B (4-wide) or 128 B (8-wide) each */
    uint64_t triFetches[3];       /* 48 B each */
    uint64_t waveNodeFetches[3];
    uint64_t waveTriFetches[3];
    uint64_t shadowOccluded;
    uint32_t bvhWidth;            /* tree the kernels walked */
    uint32_t reserved;
} shiro_hip_wavefront_counts;
int shiro_hip_count_wavefront(
    shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
    const shiro_hip_settings* settings, uint32_t sampleStart,
    uint32_t sampleCount, shiro_hip_wavefront_counts* out);

/* ---- Wavefront tuning (no counterpart in the reference: knobs of the
 * wavefront transport, set by the caller or chosen by shiro_hip_autotune). ---- */
typedef struct shiro_hip_tuning {
    uint32_t pathsPerPass;  /* paths in flight per pass (default 128M; the path
                               state takes about (176 + 32 x light samples) B per
                               path, capped at 32 GiB) */
    uint32_t fetchMode;     /* traversal work distribution: 0 wave rounds, 1 lane refill,
                               2 lane refill from per-wave pools of the queue;
                               SHIRO_HIP_FETCH_AUTO (the default): 2 for passes of at least 8M
                               paths, 0 below */
    uint32_t traceBlocks;   /* grid of the fetchMode 2 traversal kernels (default 32 x CUs) */
    uint32_t stageBlocks;   /* grid cap of the grid-stride stage kernels (default 16384) */
    uint32_t batchBlocksPerCU; /* one-wave blocks per CU of the batched persistent kernel (default 8) */
    uint32_t shadeOrder;    /* 1: shade bounces after the first in shading-class order (default 1) */
    uint32_t primaryPackets; /* G: a pass's paths are laid out in groups of 64, each G
                                samples of an 8/sqrt(G)-pixel square (G = 1: 8x8 pixels,
                                4: 4x4, 16: 2x2, 64: one pixel; a smaller G when the
                                pass's samples per pixel are not a multiple), and each
                                group's camera rays are traced as one packet; 0: G = 1
                                and per-lane camera-ray traversal (default 16) */
    uint32_t bvhWidth;      /* 4 or 8: tree the wavefront traversal kernels walk (8 falls
                               back to 4 for a scene without an 8-wide tree) */
    uint32_t manyBranches;  /* SHIRO_HIP_MANY_WAVEFRONT_GRAPH: concurrent branches (streams, each
                               with its own wavefront state) of the batch graph, 1-16 (default 4) */
    uint32_t leafWaitLanes; /* closest-hit traversal: a wave turns to its leaf tests once at most
                               this many of its 64 lanes still search for a leaf (0-64, default 24) */
    uint32_t shadowLeafWaitLanes; /* the same for shadow rays (0-64, default 24) */
    uint32_t tailDepth;     /* from this bounce on (1 or more; 0: never) the paths still alive are
                               finished by one persistent kernel instead of a launch of each
                               stage per bounce (frames unchanged); SHIRO_HIP_TAIL_AUTO (the
                               default): two bounces after roulette starts */
    uint32_t shadowOverlap; /* 1: a bounce's shadow rays and their resolve run on a second stream
                               beside the next bounce's closest hits (frames unchanged; not under
                               the caller's stream capture or in the entry point that allocates
                               nothing); 0: every stage in turn on one stream (default 1) */
} shiro_hip_tuning;
#define SHIRO_HIP_TAIL_AUTO 0xffffffffu
#define SHIRO_HIP_FETCH_AUTO 0xffffffffu
int shiro_hip_set_tuning(shiro_hip_ctx* ctx, const shiro_hip_tuning* tuning);
int shiro_hip_get_tuning(const shiro_hip_ctx* ctx, shiro_hip_tuning* tuning);

/* Times candidate tunings on `trialSamples` samples of the given frame
 * (wavefront transport, uploaded scene): pass sizes, fetch modes and trace
 * grids, then the best of those with other camera-packet sizes, the other
 * tree width, and the tail kernel from bounce 3, 4, 5, 6 or 8.  Keeps the
 * fastest in the context and returns it; `report` (may be NULL) receives
 * one line per candidate.
 * Pass sizes above trialSamples x pixels time like the whole trial in one
 * pass, so trials should span the pass sizes of interest. */
int shiro_hip_autotune(
    shiro_hip_ctx* ctx, const shiro_hip_camera* camera,
    const shiro_hip_settings* settings, uint32_t trialSamples,
    shiro_hip_tuning* best, char* report, size_t reportSize);

/* ---- Batched multi-scene rendering (no counterpart in the reference: one
 * launch for the sample batches of many scenes, e.g. a farm of small
 * frames). ---- */
typedef struct shiro_hip_batch_item {
    const shiro_hip_ctx* scene;   /* context holding the uploaded scene (same device) */
    shiro_hip_camera camera;
    shiro_hip_settings settings;  /* transport is ignored */
    uint32_t sampleStart;
    uint32_t sampleCount;
    shiro_hip_frame frame;        /* device memory, settings.width x height */
} shiro_hip_batch_item;
enum {
    SHIRO_HIP_MANY_PERSISTENT = 0, /* one persistent megakernel over (item, tile) units, longest first */
    SHIRO_HIP_MANY_GRAPH = 1,      /* one megakernel launch per item, captured in a hipGraph and replayed */
    SHIRO_HIP_MANY_SEQUENTIAL = 2, /* one megakernel launch per item, back to back on the stream */
    SHIRO_HIP_MANY_WAVEFRONT = 3,  /* the wavefront form of each item in turn (ctx's tuning and buffers) */
    SHIRO_HIP_MANY_WAVEFRONT_GRAPH = 4 /* the wavefront form of every item in one hipGraph: items
                                          assigned longest first to tuning.manyBranches concurrent
                                          branches, the fork/join captured once, instantiated and
                                          replayed (updated in place when the items change) */
};
/* Renders every item's batch (means into its device frame, like
 * shiro_hip_render_batch_device; the frames are identical in every mode)
 * on ctx's stream and waits.  shiro_hip_last_render_ms times the kernels
 * (for the graph mode the replay, not the capture).  The scene contexts
 * must not be changed while the call runs. */
int shiro_hip_render_many(
    shiro_hip_ctx* ctx, const shiro_hip_batch_item* items, uint32_t count,
    uint32_t mode);
/* Counts of SHIRO_HIP_MANY_WAVEFRONT_GRAPH on this context: graphs captured
 * (a replay of unchanged items captures nothing) and captures that updated
 * the instantiated graph in place instead of instantiating a new one. */
int shiro_hip_many_graph_stats(const shiro_hip_ctx* ctx, uint32_t* captures, uint32_t* updates);
/* Times the persistent mode for blocks-per-CU in {1, 2, 4, 8, 16, 32}, then
 * the graph-captured wavefront batch for manyBranches in {1, 2, 3, 4, 6, 8}
 * and, with the best of those, for tailDepth in {0, 3, 4, 5, 6} (a warm run
 * and a timed run each; the warm run of a setting captures and instantiates
 * its graph, the timed run replays it), keeps the fastest of each in ctx's
 * tuning and returns it; `report` (may be NULL) receives one line per
 * candidate. */
int shiro_hip_autotune_many(
    shiro_hip_ctx* ctx, const shiro_hip_batch_item* items, uint32_t count,
    shiro_hip_tuning* best, char* report, size_t reportSize);

/* ---- Per-kernel timing (no counterpart in the reference: measurement of
 * the render kernels for roofline accounting). ---- */
enum {
    SHIRO_HIP_STAGE_RESET = 0,      /* wavefront queue counters cleared */
    SHIRO_HIP_STAGE_EXTEND = 1,     /* closest-hit traversal of the path queue */
    SHIRO_HIP_STAGE_SHADE = 2,      /* surface/miss terms, light samples, bounce (depth > 0) */
    SHIRO_HIP_STAGE_SHADOW = 3,     /* any-hit traversal of the shadow-ray queue */
    SHIRO_HIP_STAGE_RESOLVE = 4,    /* ordered sum of unoccluded light samples */
    SHIRO_HIP_STAGE_ACCUM = 5,      /* per-pixel sample sums of the pass */
    SHIRO_HIP_STAGE_MEGAKERNEL = 6, /* one megakernel launch */
    SHIRO_HIP_STAGE_ORDER = 7,      /* hit queue sorted by shading class */
    SHIRO_HIP_STAGE_PRIMARY = 8,    /* closest hit of the camera rays traced as packets (depth 0) */
    SHIRO_HIP_STAGE_SHADE_FIRST = 9, /* the first bounce's shading (its own kernel; SHADE: the later bounces) */
    SHIRO_HIP_STAGE_TAIL = 10,      /* the bounces from tuning.tailDepth on, in one persistent kernel */
    SHIRO_HIP_STAGE_COUNT = 11
};
typedef struct shiro_hip_stage_times {
    float ms[SHIRO_HIP_STAGE_COUNT];          /* summed kernel time per stage */
    uint32_t launches[SHIRO_HIP_STAGE_COUNT]; /* launches per stage */
    uint32_t complete;                        /* 0 if the event pool ran out */
    uint32_t reserved;
} shiro_hip_stage_times;
/* While enabled, every render call records an event after each kernel
 * launch on the context's stream; shiro_hip_get_stage_times returns the
 * per-stage sums of the last synchronised render call. */
int shiro_hip_set_stage_timing(shiro_hip_ctx* ctx, uint32_t enable);
/* Calibration of the VALU issue rate (measurement; no counterpart in the
 * reference): a kernel of independent FP32 FMAs at full occupancy, 32 waves
 * per compute unit for `iterations` iterations of 8 FMAs.  Returns its
 * HIP-event time and the VALU wave-instructions it issued (FMAs only), whose
 * ratio is the sustained issue rate the render kernels' VALU counts are
 * compared with. */
int shiro_hip_valu_probe(shiro_hip_ctx* ctx, uint32_t iterations, float* ms, uint64_t* waveInstructions);
int shiro_hip_get_stage_times(shiro_hip_ctx* ctx, shiro_hip_stage_times* out);

/* ---- Mesh ingestion of Shiro's Hydra scene bridge (host code, no device):
 * what HdShiroSceneBridge does to a USD polygon mesh before it reaches a
 * render backend. ---- */
/* Fan triangulation (reference src/hydra/SceneBridge.cpp:1009-1020): a face
 * of c >= 3 vertices v0..v(c-1) gives triangles (v0, vk, vk+1), k = 1..c-2;
 * faces of fewer vertices are skipped (their indices consumed).  *outCount
 * receives the triangle index count; with outIndices NULL only the count is
 * computed.  SHIRO_HIP_INVALID for a negative count, faces needing more
 * than indexCount indices, or outCapacity below the count. */
int shiro_hip_triangulate(const int32_t* faceVertexCounts, uint32_t faceCount, const int32_t* faceVertexIndices,
                          uint32_t indexCount, uint32_t* outIndices, uint64_t outCapacity, uint64_t* outCount);
/* ComputeSmoothNormals (reference SceneBridge.cpp:147-180), used when a
 * mesh has no normal per point: cross(p1 - p0, p2 - p0) of each triangle
 * added to its three vertices (triangles with an out-of-range index or a
 * cross product of length <= 1e-8 skipped), then normalised; a vertex with
 * a zero sum gets (0, 1, 0).  outNormals: 3 x vertexCount floats. */
int shiro_hip_smooth_normals(const float* positions, uint32_t vertexCount, const uint32_t* indices,
                             uint32_t indexCount, float* outNormals);

/* ---- Environment image files: the pixel decoding EnvironmentMap::Load
 * gets from OpenImageIO (reference src/render/EnvironmentMap.cpp:49-92),
 * for hosts built without it.  Formats: PFM ("PF"/"Pf", either byte order),
 * Radiance RGBE (flat or run-length scanlines, any [+-]Y [+-]X orientation)
 * and OpenEXR (scanline or tiled, single- or multi-part: the first part,
 * level 0 of mipmapped or ripmapped tiles; NONE/RLE/ZIPS/ZIP/PIZ/PXR24/
 * B44/B44A; HALF/FLOAT/UINT; R,G,B or Y channels; deep files and DWAA/DWAB
 * are rejected).  `pixels` receives width*height RGB floats, row 0 the top
 * of the image, with Load's channel expansion (g = r and b = g where the
 * file lacks them): the layout shiro_hip_dome_light.pixels takes.  Needs no
 * device. ---- */
enum { SHIRO_HIP_IMAGE_PFM = 1, SHIRO_HIP_IMAGE_RGBE = 2, SHIRO_HIP_IMAGE_EXR = 3 };
typedef struct shiro_hip_image {
    float* pixels;     /* owned by the library: release with shiro_hip_free_image */
    uint32_t width;
    uint32_t height;
    uint32_t channels; /* channels stored in the file */
    uint32_t format;   /* SHIRO_HIP_IMAGE_* */
} shiro_hip_image;
/* Returns SHIRO_HIP_OK or SHIRO_HIP_INVALID (unreadable, malformed or
 * unsupported file; `err` receives the reason) or SHIRO_HIP_OOM. */
int shiro_hip_load_image(const char* path, shiro_hip_image* out, char* err, size_t errSize);
/* The same decoding from a file image held in memory. */
int shiro_hip_load_image_memory(const void* data, size_t size, shiro_hip_image* out, char* err, size_t errSize);
void shiro_hip_free_image(shiro_hip_image* img);

#ifdef __cplusplus
}
#endif

#endif /* SHIRO_HIP_H */
