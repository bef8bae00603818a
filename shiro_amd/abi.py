"""ctypes mirror of include/shiro_hip.h (the C ABI of libshiro_hip.so).

The structures follow the header field for field; the header documents which
reference type each one mirrors (PbrMaterial, TriangleMesh, DomeLight, ...).
"""
import ctypes as C

f3 = C.c_float * 3


class Material(C.Structure):
    _fields_ = [
        ("baseColor", f3), ("emissionColor", f3), ("specularColor", f3),
        ("transmissionColor", f3), ("transmissionScatter", f3), ("coatColor", f3),
        ("subsurfaceColor", f3), ("subsurfaceRadius", f3), ("sheenColor", f3),
        ("normalOverride", f3), ("coatNormalOverride", f3), ("tangentOverride", f3),
        ("baseWeight", C.c_float), ("emissionStrength", C.c_float),
        ("specularWeight", C.c_float), ("metallic", C.c_float),
        ("roughness", C.c_float), ("diffuseRoughness", C.c_float),
        ("opacity", C.c_float), ("specularAnisotropy", C.c_float),
        ("specularRotation", C.c_float), ("coatWeight", C.c_float),
        ("coatRoughness", C.c_float), ("coatIor", C.c_float),
        ("coatAnisotropy", C.c_float), ("coatRotation", C.c_float),
        ("sheen", C.c_float), ("sheenRoughness", C.c_float),
        ("subsurface", C.c_float), ("subsurfaceScale", C.c_float),
        ("subsurfaceAnisotropy", C.c_float), ("transmission", C.c_float),
        ("transmissionDepth", C.c_float), ("transmissionScatterAnisotropy", C.c_float),
        ("transmissionDispersion", C.c_float), ("transmissionExtraRoughness", C.c_float),
        ("ior", C.c_float), ("coatAffectColor", C.c_float),
        ("coatAffectRoughness", C.c_float), ("thinFilmThickness", C.c_float),
        ("thinFilmIor", C.c_float),
        ("thinWalled", C.c_uint32), ("hasNormalOverride", C.c_uint32),
        ("hasCoatNormalOverride", C.c_uint32), ("hasTangentOverride", C.c_uint32),
        ("baseColorTexture", C.c_int32), ("roughnessTexture", C.c_int32),
    ]


class Mesh(C.Structure):
    _fields_ = [
        ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
        ("uvs", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_uint32)),
        ("vertexCount", C.c_uint32), ("normalCount", C.c_uint32),
        ("indexCount", C.c_uint32), ("materialIndex", C.c_uint32),
    ]


class DirectionalLight(C.Structure):
    _fields_ = [("direction", f3), ("radiance", f3)]


class DomeLight(C.Structure):
    _fields_ = [
        ("radiance", f3), ("right", f3), ("up", f3), ("forward", f3),
        ("pixels", C.POINTER(C.c_float)), ("width", C.c_uint32),
        ("height", C.c_uint32), ("layout", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Texture(C.Structure):
    _fields_ = [
        ("texels", C.POINTER(C.c_float)), ("width", C.c_uint32),
        ("height", C.c_uint32), ("channels", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("materials", C.POINTER(Material)), ("materialCount", C.c_uint32),
        ("meshes", C.POINTER(Mesh)), ("meshCount", C.c_uint32),
        ("distantLights", C.POINTER(DirectionalLight)), ("distantLightCount", C.c_uint32),
        ("domeLights", C.POINTER(DomeLight)), ("domeLightCount", C.c_uint32),
        ("environmentTop", f3), ("environmentBottom", f3),
        ("textures", C.POINTER(Texture)), ("textureCount", C.c_uint32),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("position", f3), ("forward", f3), ("right", f3), ("up", f3),
        ("verticalFovDegrees", C.c_float), ("aspectRatio", C.c_float),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32),
        ("samplesPerPixel", C.c_uint32), ("samplesPerUpdate", C.c_uint32),
        ("domeLightSamples", C.c_uint32), ("maxDepth", C.c_uint32),
        ("diffuseDepth", C.c_uint32), ("specularDepth", C.c_uint32),
        ("threadLimit", C.c_uint32), ("maxSubdivLevel", C.c_uint32),
        ("backend", C.c_uint32), ("backgroundVisible", C.c_uint32),
        ("enableHeadlight", C.c_uint32), ("enableCaustics", C.c_uint32),
        ("enableSss", C.c_uint32), ("russianRouletteDepth", C.c_uint32),
        ("transport", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Frame(C.Structure):
    _fields_ = [
        ("beauty", C.POINTER(C.c_float)), ("albedo", C.POINTER(C.c_float)),
        ("normal", C.POINTER(C.c_float)), ("depth", C.POINTER(C.c_float)),
        ("width", C.c_uint32), ("height", C.c_uint32),
    ]


class Status(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32), ("name", C.c_char_p), ("available", C.c_uint32),
        ("usesGpu", C.c_uint32), ("deviceName", C.c_char * 64),
        ("computeUnits", C.c_uint32),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("triangleCount", C.c_uint64), ("bvhNodeCount", C.c_uint64),
        ("bvhLeafCount", C.c_uint64), ("bvhDepth", C.c_uint32),
        ("maxLeafSize", C.c_uint32), ("deviceBytes", C.c_uint64),
        ("buildSeconds", C.c_double), ("uploadSeconds", C.c_double),
        ("bvh8NodeCount", C.c_uint64), ("bvh8Depth", C.c_uint32), ("bvhStack", C.c_uint32),
        ("bvh8Stack", C.c_uint32), ("plainBsdf", C.c_uint32),
    ]


class FetchCounts(C.Structure):
    _fields_ = [
        ("samples", C.c_uint64), ("rays", C.c_uint64), ("shadowRays", C.c_uint64),
        ("nodeFetches", C.c_uint64), ("triFetches", C.c_uint64),
        ("envTaps", C.c_uint64), ("cdfSteps", C.c_uint64), ("hitFetches", C.c_uint64),
        ("shadowNodeFetches", C.c_uint64), ("shadowTriFetches", C.c_uint64),
        ("primaryRays", C.c_uint64), ("primaryNodeFetches", C.c_uint64), ("primaryTriFetches", C.c_uint64),
    ]


class WavefrontCounts(C.Structure):
    _fields_ = [
        ("samples", C.c_uint64), ("rays", C.c_uint64 * 3), ("nodeFetches", C.c_uint64 * 3),
        ("triFetches", C.c_uint64 * 3), ("waveNodeFetches", C.c_uint64 * 3), ("waveTriFetches", C.c_uint64 * 3),
        ("shadowOccluded", C.c_uint64),
        ("bvhWidth", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("pathsPerPass", C.c_uint32), ("fetchMode", C.c_uint32), ("traceBlocks", C.c_uint32),
        ("stageBlocks", C.c_uint32), ("batchBlocksPerCU", C.c_uint32), ("shadeOrder", C.c_uint32),
        ("primaryPackets", C.c_uint32), ("bvhWidth", C.c_uint32), ("manyBranches", C.c_uint32),
        ("leafWaitLanes", C.c_uint32), ("shadowLeafWaitLanes", C.c_uint32), ("tailDepth", C.c_uint32),
        ("shadowOverlap", C.c_uint32),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("scene", C.c_void_p), ("camera", Camera), ("settings", Settings),
        ("sampleStart", C.c_uint32), ("sampleCount", C.c_uint32), ("frame", Frame),
    ]


MANY_PERSISTENT, MANY_GRAPH, MANY_SEQUENTIAL, MANY_WAVEFRONT, MANY_WAVEFRONT_GRAPH = 0, 1, 2, 3, 4


STAGE_NAMES = ["reset", "extend", "shade", "shadow", "resolve", "accum", "megakernel", "order", "primary", "shade_first",
               "tail"]


class StageTimes(C.Structure):
    _fields_ = [
        ("ms", C.c_float * 11), ("launches", C.c_uint32 * 11),
        ("complete", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Image(C.Structure):
    _fields_ = [
        ("pixels", C.POINTER(C.c_float)), ("width", C.c_uint32), ("height", C.c_uint32),
        ("channels", C.c_uint32), ("format", C.c_uint32),
    ]


IMAGE_PFM, IMAGE_RGBE, IMAGE_EXR = 1, 2, 3

RESULT_NAMES = {
    0: "OK", 1: "NO_DEVICE", 2: "OOM", 3: "CANCELLED", 4: "INVALID",
    5: "NO_SCENE", 6: "LAUNCH",
}

ENV_AUTOMATIC, ENV_LATLONG, ENV_ANGULAR = 0, 1, 2
TEXELS_DEVICE, TEXELS_EVAL_ONLY = 1, 2
PROGRESSIVE_RESET = 1
TRANSPORT_MEGAKERNEL, TRANSPORT_WAVEFRONT, TRANSPORT_MEGAKERNEL_PLUS = 0, 1, 2
TAIL_AUTO = 0xFFFFFFFF  # shiro_hip_tuning.tailDepth: two bounces after roulette starts
FETCH_AUTO = 0xFFFFFFFF  # shiro_hip_tuning.fetchMode: per-wave pools for passes of 8M paths or more

# Every symbol include/shiro_hip.h declares (checked by the ABI test).
EXPORTS = [
    "shiro_hip_default_settings", "shiro_hip_default_material",
    "shiro_hip_create", "shiro_hip_destroy", "shiro_hip_last_error",
    "shiro_hip_get_status", "shiro_hip_supports_scene", "shiro_hip_upload_scene",
    "shiro_hip_render_batch", "shiro_hip_render_batch_device", "shiro_hip_render_batch_device_noalloc",
    "shiro_hip_accum_reset", "shiro_hip_accum_add", "shiro_hip_accum_resolve", "shiro_hip_render_progressive",
    "shiro_hip_stream", "shiro_hip_synchronize", "shiro_hip_last_render_ms",
    "shiro_hip_scene_info_get", "shiro_hip_qmc_fill", "shiro_hip_intersect",
    "shiro_hip_occluded", "shiro_hip_trace_primary", "shiro_hip_eval_bsdf",
    "shiro_hip_env_sample", "shiro_hip_env_eval", "shiro_hip_count_fetches", "shiro_hip_count_wavefront",
    "shiro_hip_set_tuning", "shiro_hip_get_tuning", "shiro_hip_autotune",
    "shiro_hip_set_stage_timing", "shiro_hip_get_stage_times", "shiro_hip_valu_probe",
    "shiro_hip_update_texture", "shiro_hip_render_adjoint",
    "shiro_hip_render_many", "shiro_hip_autotune_many", "shiro_hip_many_graph_stats",
    "shiro_hip_load_image", "shiro_hip_load_image_memory", "shiro_hip_free_image",
    "shiro_hip_refit_scene", "shiro_hip_triangulate", "shiro_hip_smooth_normals",
    "shiro_hip_render_adjoint_env", "shiro_hip_update_environment",
    "shiro_hip_set_replay_log", "shiro_hip_last_adjoint_replayed",
]


def fp(a):
    """float32 numpy array -> float*."""
    return a.ctypes.data_as(C.POINTER(C.c_float))


def up(a):
    """uint32 numpy array -> uint32_t*."""
    return a.ctypes.data_as(C.POINTER(C.c_uint32))
