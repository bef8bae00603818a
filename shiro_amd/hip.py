"""ctypes binding of libshiro_hip.so and a Python HipBackend mirroring Shiro's
RenderBackend interface (reference include/shiro/backend/RenderBackend.h:12-34):

    backend = HipBackend()
    backend.GetStatus()                       -> BackendStatus
    backend.SupportsScene(scene)              -> bool   (OptixBackend.h:16)
    backend.RenderSampleBatch(RenderRequest)  -> FrameBuffer

The library is required: importing works without it, but constructing a
context raises if libshiro_hip.so is missing or no HIP device is present --
there is no CPU fallback in this package.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import abi
from .scene import Camera, FrameBuffer, RenderSettings, Scene, marshal

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libshiro_hip.so")

_P = C.POINTER
_u = C.c_uint32
_fp, _up = _P(C.c_float), _P(C.c_uint32)
_lib = None


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.RESULT_NAMES.get(code, code)}: {msg}")
        self.code = code


def load_library(path: str = LIB_PATH):
    """Loads libshiro_hip.so and declares its signatures (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is not built; run __graft_entry__.build() or `make -C shiro_amd/csrc`")
    # One HIP runtime per process: when PyTorch is installed, load it first so
    # that the library binds to the same libamdhip64 as torch's tensors.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)

    def d(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    V = C.c_void_p
    d("shiro_hip_default_settings", None, _P(abi.Settings))
    d("shiro_hip_default_material", None, _P(abi.Material))
    d("shiro_hip_create", C.c_int, C.c_int, _P(V))
    d("shiro_hip_destroy", None, V)
    d("shiro_hip_last_error", C.c_char_p, V)
    d("shiro_hip_get_status", C.c_int, V, _P(abi.Status))
    d("shiro_hip_supports_scene", C.c_int, _P(abi.Scene))
    d("shiro_hip_upload_scene", C.c_int, V, _P(abi.Scene))
    d("shiro_hip_render_batch", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _u, _P(C.c_uint8), _P(abi.Frame))
    d("shiro_hip_render_batch_device", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _u, _P(abi.Frame), C.c_int)
    d("shiro_hip_render_batch_device_noalloc", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _u, _P(abi.Frame),
      C.c_int)
    d("shiro_hip_accum_reset", C.c_int, V, _u, _u)
    d("shiro_hip_accum_add", C.c_int, V, _P(abi.Frame), _u)
    d("shiro_hip_accum_resolve", C.c_int, V, _P(abi.Frame))
    d("shiro_hip_render_progressive", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _u, _u, _P(C.c_uint8),
      _P(abi.Frame), _P(C.c_uint32))
    d("shiro_hip_stream", V, V)
    d("shiro_hip_synchronize", C.c_int, V)
    d("shiro_hip_last_render_ms", C.c_int, V, _P(C.c_float))
    d("shiro_hip_scene_info_get", C.c_int, V, _P(abi.SceneInfo))
    d("shiro_hip_qmc_fill", C.c_int, V, _u, _u, _u, _u, _u, _u, _u, _fp)
    d("shiro_hip_intersect", C.c_int, V, _fp, _u, _up, _fp, _fp)
    d("shiro_hip_occluded", C.c_int, V, _fp, _u, _P(C.c_uint8))
    d("shiro_hip_trace_primary", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _up, _fp)
    d("shiro_hip_eval_bsdf", C.c_int, V, _up, _fp, _fp, _fp, _fp, _u, _fp, _fp)
    d("shiro_hip_env_sample", C.c_int, V, _u, _fp, _u, _fp, _fp, _fp)
    d("shiro_hip_env_eval", C.c_int, V, _u, _fp, _u, _fp, _fp)
    d("shiro_hip_count_fetches", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _u, _P(abi.FetchCounts))
    d("shiro_hip_count_wavefront", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _u, _P(abi.WavefrontCounts))
    d("shiro_hip_set_tuning", C.c_int, V, _P(abi.Tuning))
    d("shiro_hip_get_tuning", C.c_int, V, _P(abi.Tuning))
    d("shiro_hip_autotune", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _P(abi.Tuning), C.c_char_p,
      C.c_size_t)
    d("shiro_hip_set_stage_timing", C.c_int, V, _u)
    d("shiro_hip_get_stage_times", C.c_int, V, _P(abi.StageTimes))
    d("shiro_hip_valu_probe", C.c_int, V, _u, _P(C.c_float), _P(C.c_uint64))
    d("shiro_hip_update_texture", C.c_int, V, _u, _fp, _u)
    d("shiro_hip_render_adjoint", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _u, _fp, _P(_fp), _u)
    d("shiro_hip_render_adjoint_env", C.c_int, V, _P(abi.Camera), _P(abi.Settings), _u, _u, _fp, _P(_fp), _u,
      _P(_fp), _u)
    d("shiro_hip_update_environment", C.c_int, V, _u, _fp, _u)
    d("shiro_hip_set_replay_log", C.c_int, V, C.c_uint64)
    d("shiro_hip_last_adjoint_replayed", C.c_int, V, _P(C.c_uint32))
    d("shiro_hip_render_many", C.c_int, V, _P(abi.BatchItem), _u, _u)
    d("shiro_hip_autotune_many", C.c_int, V, _P(abi.BatchItem), _u, _P(abi.Tuning), C.c_char_p, C.c_size_t)
    d("shiro_hip_many_graph_stats", C.c_int, V, _P(C.c_uint32), _P(C.c_uint32))
    d("shiro_hip_refit_scene", C.c_int, V, _P(abi.Scene))
    d("shiro_hip_triangulate", C.c_int, _P(C.c_int32), _u, _P(C.c_int32), _u, _up, C.c_uint64, _P(C.c_uint64))
    d("shiro_hip_smooth_normals", C.c_int, _fp, _u, _up, _u, _fp)
    d("shiro_hip_load_image", C.c_int, C.c_char_p, _P(abi.Image), C.c_char_p, C.c_size_t)
    d("shiro_hip_load_image_memory", C.c_int, C.c_void_p, C.c_size_t, _P(abi.Image), C.c_char_p, C.c_size_t)
    d("shiro_hip_free_image", None, _P(abi.Image))
    _lib = L
    return L


def load_image(source) -> tuple:
    """Decodes an environment image file (PFM, Radiance RGBE or scanline
    OpenEXR) the way EnvironmentMap::Load reads it through OpenImageIO
    (EnvironmentMap.cpp:49-92): returns ((H, W, 3) float32 RGB with row 0 the
    top of the image, channels stored in the file, abi.IMAGE_* format).
    `source` is a path or the file's bytes.  Needs no device."""
    L = load_library()
    img = abi.Image()
    err = C.create_string_buffer(512)
    if isinstance(source, (bytes, bytearray, memoryview)):
        buf = bytes(source)
        rc = L.shiro_hip_load_image_memory(buf, len(buf), C.byref(img), err, len(err))
    else:
        rc = L.shiro_hip_load_image(os.fsencode(source), C.byref(img), err, len(err))
    if rc:
        raise HipError(rc, err.value.decode(errors="replace"))
    try:
        n = img.width * img.height * 3
        px = np.ctypeslib.as_array(img.pixels, shape=(n,)).copy().reshape(img.height, img.width, 3)
        return px, int(img.channels), int(img.format)
    finally:
        L.shiro_hip_free_image(C.byref(img))


def triangulate(face_vertex_counts, face_vertex_indices) -> np.ndarray:
    """Fan triangulation of polygon faces as Shiro's Hydra scene bridge does
    it (SceneBridge.cpp:1009-1020); returns the triangle indices (u32)."""
    L = load_library()
    counts = np.ascontiguousarray(face_vertex_counts, np.int32)
    idx = np.ascontiguousarray(face_vertex_indices, np.int32)
    n = C.c_uint64(0)
    cp, ip = counts.ctypes.data_as(_P(C.c_int32)), idx.ctypes.data_as(_P(C.c_int32))
    rc = L.shiro_hip_triangulate(cp, len(counts), ip, len(idx), None, 0, C.byref(n))
    if rc:
        raise HipError(rc, "bad polygon face counts or indices")
    out = np.zeros(n.value, np.uint32)
    rc = L.shiro_hip_triangulate(cp, len(counts), ip, len(idx), abi.up(out), len(out), C.byref(n))
    if rc:
        raise HipError(rc, "triangulation failed")
    return out


def smooth_normals(positions, indices) -> np.ndarray:
    """Area-weighted vertex normals of a triangle mesh
    (ComputeSmoothNormals, SceneBridge.cpp:147-180); (V, 3) float32."""
    L = load_library()
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    out = np.zeros_like(pos)
    rc = L.shiro_hip_smooth_normals(abi.fp(pos), len(pos), abi.up(idx), len(idx), abi.fp(out))
    if rc:
        raise HipError(rc, "smooth normals failed")
    return out


def polygon_mesh(points, face_vertex_counts, face_vertex_indices, normals=None, material_index: int = 0):
    """A TriangleMesh from a polygon mesh the way the Hydra scene bridge
    builds one (SceneBridge.cpp:967-1032, points already in world space):
    fan-triangulated faces, the given per-point normals normalised, or smooth
    normals when there is not one per point; None when nothing renderable
    (fewer than 3 points or triangle indices) is left."""
    from .scene import TriangleMesh
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    idx = triangulate(face_vertex_counts, face_vertex_indices)
    if len(pts) < 3 or len(idx) < 3:
        return None
    if normals is not None and len(normals) == len(pts):
        n = np.asarray(normals, np.float32).reshape(-1, 3)
        ln = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
        safe = np.where(ln > 0, ln, np.float32(1))[:, None]
        nrm = np.where((ln > 0)[:, None], n / safe, np.float32(0)).astype(np.float32)
    else:
        nrm = smooth_normals(pts, idx)
    return TriangleMesh(pts, nrm, idx, material_index)


def load_dome_light(path, layout: int = abi.ENV_AUTOMATIC, **kw):
    """A DomeLight whose pixels come from an image file (the reference's
    dome lights hold EnvironmentMap::Load(path, layout))."""
    from .scene import DomeLight
    px, _, _ = load_image(path)
    return DomeLight(pixels=px, layout=layout, **kw)


def scene_signature(scene: Scene) -> str:
    """Content hash of a scene (the reference re-uploads when its FNV-1a scene
    signature changes, OptixBackend.cpp:150-232)."""
    h = hashlib.blake2b(digest_size=16)
    for m in scene.meshes:
        for a in (m.positions, m.normals, m.indices, m.uvs):
            if a is not None:
                h.update(np.ascontiguousarray(a).tobytes())
        h.update(int(m.materialIndex).to_bytes(4, "little"))
    for mat in scene.materials:
        h.update(bytes(mat.to_c()))
    for d in scene.distantLights:
        h.update(np.asarray(d.direction + d.radiance, np.float32).tobytes())
    for d in scene.domeLights:
        h.update(np.asarray(d.radiance + d.right + d.up + d.forward, np.float32).tobytes())
        h.update(int(d.layout).to_bytes(4, "little"))
        if d.pixels is not None:
            h.update(np.ascontiguousarray(d.pixels).tobytes())
    h.update(np.asarray(tuple(scene.environmentTop) + tuple(scene.environmentBottom), np.float32).tobytes())
    for t in scene.textures:
        h.update(np.asarray(np.shape(t.texels), np.int64).tobytes())
        h.update(np.ascontiguousarray(t.texels, np.float32).tobytes())
    return h.hexdigest()


def topology_signature(scene: Scene) -> str:
    """Hash of what fixes a scene's triangle list (mesh count, vertex counts,
    index arrays): scenes that share it can be refitted into each other."""
    h = hashlib.blake2b(digest_size=16)
    h.update(len(scene.meshes).to_bytes(8, "little"))
    for m in scene.meshes:
        h.update(len(m.positions).to_bytes(8, "little"))
        h.update(np.ascontiguousarray(m.indices, np.uint32).tobytes())
    return h.hexdigest()


@dataclass
class BackendStatus:
    """RenderBackend.h:21-26."""
    kind: int = 1
    name: str = "hip"
    available: bool = False
    usesGpu: bool = True
    deviceName: str = ""
    computeUnits: int = 0


@dataclass
class RenderRequest:
    """RenderBackend.h:12-19."""
    scene: Scene
    camera: Camera
    settings: RenderSettings
    sampleStart: int = 0
    sampleCount: int = 1
    cancel: Optional[object] = None  # pointer to a ctypes c_uint8 flag, non-zero = cancel


class Context:
    """One libshiro_hip context (device, stream, device-resident scene)."""

    def __init__(self, device: int = 0):
        self.L = load_library()
        self.h = C.c_void_p()
        rc = self.L.shiro_hip_create(device, C.byref(self.h))
        if rc != 0:
            raise HipError(rc, "cannot create a HIP context (no device?)")
        self._sig = None
        self._keep = None

    def close(self):
        if self.h:
            self.L.shiro_hip_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise HipError(rc, self.L.shiro_hip_last_error(self.h).decode())
        return rc

    def upload(self, scene: Scene, force: bool = False):
        sig = scene_signature(scene)
        if sig == self._sig and not force:
            return
        m = marshal(scene)
        self._check(self.L.shiro_hip_upload_scene(self.h, m.ptr))
        self._sig = sig

    def refit(self, scene: Scene):
        """Uploads a scene whose meshes and indices are the uploaded scene's
        (moved vertices, other materials or lights) and refits the BVH
        instead of rebuilding it (shiro_hip_refit_scene)."""
        m = marshal(scene)
        self._check(self.L.shiro_hip_refit_scene(self.h, m.ptr))
        self._sig = scene_signature(scene)

    def scene_info(self) -> abi.SceneInfo:
        info = abi.SceneInfo()
        self._check(self.L.shiro_hip_scene_info_get(self.h, C.byref(info)))
        return info

    def render(self, camera: Camera, settings: RenderSettings, sample_start=0, sample_count=None,
               cancel=None) -> FrameBuffer:
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        fb = FrameBuffer(settings.width, settings.height)
        f = fb.to_c()
        cam, st = camera.to_c(), settings.to_c()
        rc = self.L.shiro_hip_render_batch(self.h, C.byref(cam), C.byref(st), sample_start, sample_count,
                                           cancel, C.byref(f))
        fb.rc = rc
        if rc not in (0, 3):
            self._check(rc)
        return fb

    def render_device(self, camera: Camera, settings: RenderSettings, frame_ptrs, sample_start=0,
                      sample_count=None, synchronize=True):
        """frame_ptrs: (beauty, albedo, normal, depth) device addresses."""
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        f = abi.Frame()
        f.beauty, f.albedo, f.normal, f.depth = [C.cast(C.c_void_p(p), _fp) for p in frame_ptrs]
        f.width, f.height = settings.width, settings.height
        cam, st = camera.to_c(), settings.to_c()
        self._check(self.L.shiro_hip_render_batch_device(self.h, C.byref(cam), C.byref(st), sample_start,
                                                         sample_count, C.byref(f), int(synchronize)))

    def last_render_ms(self) -> float:
        ms = C.c_float()
        self._check(self.L.shiro_hip_last_render_ms(self.h, C.byref(ms)))
        return ms.value

    def synchronize(self):
        self._check(self.L.shiro_hip_synchronize(self.h))

    def stream(self) -> int:
        return int(self.L.shiro_hip_stream(self.h) or 0)

    def accum_reset(self, w, h):
        self._check(self.L.shiro_hip_accum_reset(self.h, w, h))

    def accum_add(self, frame_ptrs, w, h, count):
        f = abi.Frame()
        f.beauty, f.albedo, f.normal, f.depth = [C.cast(C.c_void_p(p), _fp) for p in frame_ptrs]
        f.width, f.height = w, h
        self._check(self.L.shiro_hip_accum_add(self.h, C.byref(f), count))

    def accum_resolve(self, frame_ptrs, w, h):
        f = abi.Frame()
        f.beauty, f.albedo, f.normal, f.depth = [C.cast(C.c_void_p(p), _fp) for p in frame_ptrs]
        f.width, f.height = w, h
        self._check(self.L.shiro_hip_accum_resolve(self.h, C.byref(f)))

    def render_progressive(self, camera: Camera, settings: RenderSettings, sample_start, sample_count,
                           reset=False, resolve=True, cancel=None):
        """One progressive batch with the accumulator on the device
        (shiro_hip_render_progressive): returns (resolved FrameBuffer or None,
        accumulated sample count)."""
        fb = FrameBuffer(settings.width, settings.height) if resolve else None
        f = fb.to_c() if fb is not None else None
        n = C.c_uint32(0)
        cam, st = camera.to_c(), settings.to_c()
        rc = self.L.shiro_hip_render_progressive(self.h, C.byref(cam), C.byref(st), sample_start, sample_count,
                                                 abi.PROGRESSIVE_RESET if reset else 0, cancel,
                                                 C.byref(f) if f is not None else None, C.byref(n))
        if fb is not None:
            fb.rc = rc
        if rc not in (0, 3):
            self._check(rc)
        return fb, n.value

    # ---- single stages ----
    def qmc_fill(self, x0, y0, w, h, sample, dim0, dim_count) -> np.ndarray:
        out = np.zeros(w * h * dim_count, np.float32)
        self._check(self.L.shiro_hip_qmc_fill(self.h, x0, y0, w, h, sample, dim0, dim_count, abi.fp(out)))
        return out.reshape(h, w, dim_count)

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        uv = np.zeros((n, 2), np.float32)
        self._check(self.L.shiro_hip_intersect(self.h, abi.fp(rays), n, abi.up(prim), abi.fp(t), abi.fp(uv)))
        return prim, t, uv

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(len(rays), np.uint8)
        self._check(self.L.shiro_hip_occluded(self.h, abi.fp(rays), len(rays), out.ctypes.data_as(_P(C.c_uint8))))
        return out

    def trace_primary(self, camera, settings, sample_index=0):
        n = settings.width * settings.height
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        cam, st = camera.to_c(), settings.to_c()
        self._check(self.L.shiro_hip_trace_primary(self.h, C.byref(cam), C.byref(st), sample_index, abi.up(prim),
                                                   abi.fp(t)))
        return prim, t

    def eval_bsdf(self, mat_index, n, cn, v, l):
        mi = np.ascontiguousarray(mat_index, np.uint32)
        arrs = [np.ascontiguousarray(a, np.float32) for a in (n, cn, v, l)]
        cnt = len(mi)
        val = np.zeros((cnt, 3), np.float32)
        pdf = np.zeros(cnt, np.float32)
        self._check(self.L.shiro_hip_eval_bsdf(self.h, abi.up(mi), *[abi.fp(a) for a in arrs], cnt, abi.fp(val),
                                               abi.fp(pdf)))
        return val, pdf

    def env_sample(self, u, dome=0):
        u = np.ascontiguousarray(u, np.float32)
        n = len(u)
        d = np.zeros((n, 3), np.float32)
        p = np.zeros(n, np.float32)
        c = np.zeros((n, 3), np.float32)
        self._check(self.L.shiro_hip_env_sample(self.h, dome, abi.fp(u), n, abi.fp(d), abi.fp(p), abi.fp(c)))
        return d, p, c

    def env_eval(self, dirs, dome=0):
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(dirs)
        p = np.zeros(n, np.float32)
        c = np.zeros((n, 3), np.float32)
        self._check(self.L.shiro_hip_env_eval(self.h, dome, abi.fp(dirs), n, abi.fp(p), abi.fp(c)))
        return p, c

    def get_tuning(self) -> abi.Tuning:
        t = abi.Tuning()
        self._check(self.L.shiro_hip_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.get_tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._check(self.L.shiro_hip_set_tuning(self.h, C.byref(t)))

    def autotune(self, camera, settings, trial_samples=2):
        """Chooses paths per pass, traversal fetch mode and grid for this
        frame, then tries the best with other camera-packet sizes, the other
        tree width and the tail kernel from several bounces; returns
        (tuning, report)."""
        t = abi.Tuning()
        buf = C.create_string_buffer(8192)
        cam, st = camera.to_c(), settings.to_c()
        self._check(self.L.shiro_hip_autotune(self.h, C.byref(cam), C.byref(st), trial_samples, C.byref(t), buf,
                                              len(buf)))
        return t, buf.value.decode()

    def count_fetches(self, camera, settings, sample_start=0, sample_count=None) -> abi.FetchCounts:
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        out = abi.FetchCounts()
        cam, st = camera.to_c(), settings.to_c()
        self._check(self.L.shiro_hip_count_fetches(self.h, C.byref(cam), C.byref(st), sample_start, sample_count,
                                                   C.byref(out)))
        return out

    def count_wavefront(self, camera, settings, sample_start=0, sample_count=None) -> abi.WavefrontCounts:
        """Fetch counters of the wavefront traversal kernels (camera, bounce
        and shadow rays) over one render of the samples with this context's
        tuning."""
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        out = abi.WavefrontCounts()
        cam, st = camera.to_c(), settings.to_c()
        self._check(self.L.shiro_hip_count_wavefront(self.h, C.byref(cam), C.byref(st), sample_start, sample_count,
                                                     C.byref(out)))
        return out

    # ---- batched multi-scene rendering ----
    @staticmethod
    def batch_items(entries):
        """entries: (scene Context, Camera, RenderSettings, frame_ptrs,
        sample_start, sample_count) tuples -> a ctypes BatchItem array."""
        arr = (abi.BatchItem * max(1, len(entries)))()
        for i, (sctx, cam, st, ptrs, start, count) in enumerate(entries):
            it = arr[i]
            it.scene = sctx.h.value
            it.camera = cam.to_c()
            it.settings = st.to_c()
            it.sampleStart = start
            it.sampleCount = st.samplesPerPixel if count is None else count
            it.frame.beauty, it.frame.albedo, it.frame.normal, it.frame.depth = [
                C.cast(C.c_void_p(p), _fp) for p in ptrs]
            it.frame.width, it.frame.height = st.width, st.height
        return arr

    def render_many(self, items, mode=abi.MANY_PERSISTENT):
        """Renders a BatchItem array (see batch_items) on this context's
        stream; returns the kernel milliseconds."""
        self._check(self.L.shiro_hip_render_many(self.h, items, len(items), mode))
        return self.last_render_ms()

    def many_graph_stats(self):
        """(graphs captured, captures that updated the instantiated graph)
        of the graph-captured wavefront batch mode on this context."""
        a, b = C.c_uint32(), C.c_uint32()
        self._check(self.L.shiro_hip_many_graph_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune_many(self, items):
        best = abi.Tuning()
        buf = C.create_string_buffer(2048)
        self._check(self.L.shiro_hip_autotune_many(self.h, items, len(items), C.byref(best), buf, len(buf)))
        return best, buf.value.decode()

    # ---- texture extension ----
    def update_texture(self, index, texels, eval_only=False):
        """Replaces texture `index`'s texels: a float32 numpy array, or a
        device address (int) of the same size.  eval_only freezes the texels
        sampling reads (finite differences of the frozen-sampling render)."""
        flags = abi.TEXELS_EVAL_ONLY if eval_only else 0
        if isinstance(texels, int):
            ptr = C.cast(C.c_void_p(texels), _fp)
            flags |= abi.TEXELS_DEVICE
            keep = None
        else:
            keep = np.ascontiguousarray(texels, np.float32)
            ptr = abi.fp(keep)
        self._check(self.L.shiro_hip_update_texture(self.h, index, ptr, flags))
        del keep
        self._sig = None  # the device scene no longer matches the uploaded one

    def update_environment(self, dome, texels, eval_only=False):
        """Replaces dome light `dome`'s map texels: an (H, W, 3) float32
        numpy array, or a device address (int) of H*W*3 floats.  The
        importance tables are rebuilt on the device unless eval_only."""
        flags = abi.TEXELS_EVAL_ONLY if eval_only else 0
        if isinstance(texels, int):
            ptr = C.cast(C.c_void_p(texels), _fp)
            flags |= abi.TEXELS_DEVICE
            keep = None
        else:
            keep = np.ascontiguousarray(texels, np.float32)
            ptr = abi.fp(keep)
        self._check(self.L.shiro_hip_update_environment(self.h, dome, ptr, flags))
        del keep
        self._sig = None

    def render_adjoint(self, camera, settings, dbeauty_ptr, grad_ptrs, sample_start=0, sample_count=None,
                       env_grad_ptrs=()):
        """Adds the texel gradients of sum(dbeauty * beauty) to the device
        buffers grad_ptrs[i] (0 = skip) for texture i, and those of dome i's
        map to env_grad_ptrs[i]; dbeauty_ptr is a device W*H*4 float image."""
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        arr = (_fp * max(1, len(grad_ptrs)))(*[C.cast(C.c_void_p(p), _fp) if p else None for p in grad_ptrs])
        env = (_fp * max(1, len(env_grad_ptrs)))(*[C.cast(C.c_void_p(p), _fp) if p else None for p in env_grad_ptrs])
        cam, st = camera.to_c(), settings.to_c()
        self._check(self.L.shiro_hip_render_adjoint_env(self.h, C.byref(cam), C.byref(st), sample_start, sample_count,
                                                        C.cast(C.c_void_p(dbeauty_ptr), _fp), arr, len(grad_ptrs),
                                                        env, len(env_grad_ptrs)))

    def set_replay_log(self, max_bytes):
        """Keeps the hits and light-sample visibility of wavefront batches
        needing at most max_bytes (0: off) for the adjoint of the same
        samples to read instead of tracing again."""
        self._check(self.L.shiro_hip_set_replay_log(self.h, int(max_bytes)))

    def last_adjoint_replayed(self) -> bool:
        out = C.c_uint32()
        self._check(self.L.shiro_hip_last_adjoint_replayed(self.h, C.byref(out)))
        return bool(out.value)

    def valu_probe(self, iterations=20000):
        """(milliseconds, VALU wave-instructions) of the FMA issue-rate probe."""
        ms, n = C.c_float(), C.c_uint64()
        self._check(self.L.shiro_hip_valu_probe(self.h, iterations, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_stage_timing(self, enable=True):
        self._check(self.L.shiro_hip_set_stage_timing(self.h, 1 if enable else 0))

    def stage_times(self) -> dict:
        """Per-stage kernel milliseconds and launch counts of the last
        synchronised render call (stage-timing mode)."""
        out = abi.StageTimes()
        self._check(self.L.shiro_hip_get_stage_times(self.h, C.byref(out)))
        return {name: {"ms": out.ms[i], "launches": out.launches[i]}
                for i, name in enumerate(abi.STAGE_NAMES) if out.launches[i]} | {"complete": bool(out.complete)}


class HipBackend:
    """RenderBackend (RenderBackend.h:28-34) over libshiro_hip.  Like the
    reference's GPU backend it reports unavailability instead of raising,
    returns a cleared frame on failure, and caches the device scene keyed by
    its content hash."""

    def __init__(self, device: int = 0):
        self.device = device
        self._ctx: Optional[Context] = None
        self._attempted = False

    # Batches of at least this many samples (pixels x samples) render with
    # the wavefront transport, smaller ones with the megakernel (the C++
    # adapter's rule); both render the same frame bit for bit.
    WAVEFRONT_MIN_SAMPLES = 1 << 18

    @classmethod
    def _settings(cls, st: RenderSettings, count: int) -> RenderSettings:
        import dataclasses
        big = st.width * st.height * count >= cls.WAVEFRONT_MIN_SAMPLES
        return dataclasses.replace(st, transport=abi.TRANSPORT_WAVEFRONT if big else abi.TRANSPORT_MEGAKERNEL)

    def _runtime(self) -> Optional[Context]:
        if self._ctx is None and not self._attempted:
            self._attempted = True
            try:
                self._ctx = Context(self.device)
            except (OSError, HipError):
                self._ctx = None
        return self._ctx

    def GetStatus(self) -> BackendStatus:
        L = load_library()
        st = abi.Status()
        L.shiro_hip_get_status(None, C.byref(st))
        return BackendStatus(kind=st.kind, name=(st.name or b"hip").decode(), available=bool(st.available),
                             usesGpu=bool(st.usesGpu), deviceName=st.deviceName.decode(),
                             computeUnits=st.computeUnits)

    def SupportsScene(self, scene: Scene) -> bool:
        m = marshal(scene)
        return bool(load_library().shiro_hip_supports_scene(m.ptr))

    def RenderSampleBatch(self, request: RenderRequest) -> FrameBuffer:
        st = request.settings
        ctx = self._runtime()
        if ctx is None:
            return FrameBuffer(st.width, st.height)
        try:
            if scene_signature(request.scene) != ctx._sig:
                # Moved vertices on unchanged meshes: refit the BVH; else
                # upload and rebuild.
                topo = topology_signature(request.scene)
                refitted = False
                if ctx._sig is not None and topo == getattr(ctx, "_topo", None):
                    try:
                        ctx.refit(request.scene)
                        refitted = True
                    except HipError:
                        refitted = False
                if not refitted:
                    ctx.upload(request.scene, force=True)
                ctx._topo = topo
            return ctx.render(request.camera, self._settings(st, request.sampleCount), request.sampleStart,
                              request.sampleCount, request.cancel)
        except HipError as e:
            import sys
            print(f"shiro_hip: {e}", file=sys.stderr)
            return FrameBuffer(st.width, st.height)
