"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the CPU oracle
(oracle/liboracle.so) and of the reference's own CPU path built from its
sources (oracle/_ref/libshiro_ref.so, present only where the reference
checkout was available at build time).

Only tests, the smoke check and the benchmark's CPU-baseline leg use this
module; the product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_HERE)
if _ROOT not in sys.path:
    sys.path.insert(0, _ROOT)

from shiro_amd import abi  # noqa: E402
from shiro_amd.scene import FrameBuffer, marshal  # noqa: E402

ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libshiro_ref.so")

_P = C.POINTER
_f, _u = C.c_float, C.c_uint32
_fp, _up = _P(C.c_float), _P(C.c_uint32)


def _bind(lib, name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


class _Lib:
    _cache = {}

    @classmethod
    def load(cls, path):
        if path not in cls._cache:
            if not os.path.exists(path):
                raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
            cls._cache[path] = C.CDLL(path)
        return cls._cache[path]


def oracle_available() -> bool:
    return os.path.exists(ORACLE_SO)


def reference_available() -> bool:
    return os.path.exists(REF_SO)


class _SceneBase:
    prefix = ""

    def __init__(self, lib, scene):
        self.lib = lib
        self.scene = scene
        self.m = marshal(scene)

    def _fn(self, name, res, *args):
        return _bind(self.lib, self.prefix + name, res, *args)

    def close(self):
        if getattr(self, "h", None):
            self._fn("scene_destroy", None, C.c_void_p)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # EnvironmentMap tables / sampling on dome `i`.
    def env_tables(self, i=0):
        d = self.scene.domeLights[i].pixels
        h, w = d.shape[:2]
        row = np.zeros(h + 1, np.float32)
        cond = np.zeros(h * (w + 1), np.float32)
        rc = self._fn("env_tables", C.c_int, C.c_void_p, _u, _fp, _fp)(self.h, i, abi.fp(row), abi.fp(cond))
        if rc:
            raise RuntimeError(f"env_tables failed ({rc})")
        return row, cond.reshape(h, w + 1)

    def env_sample(self, u, i=0):
        u = np.ascontiguousarray(u, np.float32)
        n = len(u)
        d = np.zeros((n, 3), np.float32)
        p = np.zeros(n, np.float32)
        c = np.zeros((n, 3), np.float32)
        rc = self._fn("env_sample", C.c_int, C.c_void_p, _u, _fp, _u, _fp, _fp, _fp)(
            self.h, i, abi.fp(u), n, abi.fp(d), abi.fp(p), abi.fp(c))
        if rc:
            raise RuntimeError(f"env_sample failed ({rc})")
        return d, p, c

    def env_eval(self, dirs, i=0):
        dirs = np.ascontiguousarray(dirs, np.float32)
        n = len(dirs)
        p = np.zeros(n, np.float32)
        c = np.zeros((n, 3), np.float32)
        rc = self._fn("env_eval", C.c_int, C.c_void_p, _u, _fp, _u, _fp, _fp)(
            self.h, i, abi.fp(dirs), n, abi.fp(p), abi.fp(c))
        if rc:
            raise RuntimeError(f"env_eval failed ({rc})")
        return p, c


class OracleScene(_SceneBase):
    """A scene loaded into the CPU oracle.  bvh=False selects brute force."""
    prefix = "oracle_"

    def __init__(self, scene, bvh=True):
        super().__init__(_Lib.load(ORACLE_SO), scene)
        self.h = self._fn("scene_create", C.c_void_p, C.c_void_p, C.c_int)(self.m.ptr, int(bvh))

    def set_bvh(self, on: bool):
        self._fn("scene_set_bvh", None, C.c_void_p, C.c_int)(self.h, int(on))

    def render(self, camera, settings, sample_start=0, sample_count=None, threads=0, cancel=None):
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        fb = FrameBuffer(settings.width, settings.height)
        f = fb.to_c()
        cam, st = camera.to_c(), settings.to_c()
        rc = self._fn("render", C.c_int, C.c_void_p, _P(abi.Camera), _P(abi.Settings), _u, _u, _u,
                      _P(C.c_uint8), _P(abi.Frame))(
            self.h, C.byref(cam), C.byref(st), sample_start, sample_count, threads, cancel, C.byref(f))
        fb.rc = rc
        return fb

    def render_region(self, camera, settings, x0, y0, w, h, sample_start=0, sample_count=None, threads=1):
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        fb = FrameBuffer(settings.width, settings.height)
        f = fb.to_c()
        cam, st = camera.to_c(), settings.to_c()
        self._fn("render_region", C.c_int, C.c_void_p, _P(abi.Camera), _P(abi.Settings), _u, _u, _u, _u, _u,
                 _u, _u, _P(abi.Frame))(
            self.h, C.byref(cam), C.byref(st), sample_start, sample_count, x0, y0, w, h, threads, C.byref(f))
        return fb

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        uv = np.zeros((n, 2), np.float32)
        self._fn("intersect", None, C.c_void_p, _fp, _u, _up, _fp, _fp)(
            self.h, abi.fp(rays), n, abi.up(prim), abi.fp(t), abi.fp(uv))
        return prim, t, uv

    def occluded(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(len(rays), np.uint8)
        self._fn("occluded", None, C.c_void_p, _fp, _u, _P(C.c_uint8))(
            self.h, abi.fp(rays), len(rays), out.ctypes.data_as(_P(C.c_uint8)))
        return out

    def trace_primary(self, camera, settings, sample_index=0):
        n = settings.width * settings.height
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        cam, st = camera.to_c(), settings.to_c()
        self._fn("trace_primary", None, C.c_void_p, _P(abi.Camera), _P(abi.Settings), _u, _up, _fp)(
            self.h, C.byref(cam), C.byref(st), sample_index, abi.up(prim), abi.fp(t))
        return prim, t

    # ---- texture extension ----
    def update_texture(self, index, texels, eval_only=False):
        arr = np.ascontiguousarray(texels, np.float32)
        rc = self._fn("update_texture", C.c_int, C.c_void_p, _u, _fp, C.c_int)(self.h, index, abi.fp(arr),
                                                                                int(eval_only))
        if rc:
            raise IndexError(index)

    def update_environment(self, dome, texels, eval_only=False):
        arr = np.ascontiguousarray(texels, np.float32)
        rc = self._fn("update_environment", C.c_int, C.c_void_p, _u, _fp, C.c_int)(self.h, dome, abi.fp(arr),
                                                                                    int(eval_only))
        if rc:
            raise IndexError(dome)

    def render_adjoint(self, camera, settings, dbeauty, wanted, sample_start=0, sample_count=None, threads=8,
                       env_wanted=()):
        """Texel gradients of sum(dbeauty * beauty); `wanted` lists texture
        indices, `env_wanted` dome indices; returns {index: gradient shaped
        like the texels} and, with env_wanted, also {("env", dome): gradient
        shaped like the dome's pixels}."""
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        db = np.ascontiguousarray(dbeauty, np.float32).reshape(-1)
        ntex = len(self.scene.textures)
        grads = {i: np.zeros(np.shape(self.scene.textures[i].texels), np.float32) for i in wanted}
        ptrs = (_fp * max(1, ntex))(*[abi.fp(grads[i]) if i in grads else None for i in range(ntex)])
        ndome = len(self.scene.domeLights)
        env = {d: np.zeros(np.shape(self.scene.domeLights[d].pixels), np.float32) for d in env_wanted}
        eptrs = (_fp * max(1, ndome))(*[abi.fp(env[d]) if d in env else None for d in range(ndome)])
        cam, st = camera.to_c(), settings.to_c()
        self._fn("render_adjoint_env", C.c_int, C.c_void_p, _P(abi.Camera), _P(abi.Settings), _u, _u, _fp, _P(_fp),
                 _u, _P(_fp), _u, _u)(self.h, C.byref(cam), C.byref(st), sample_start, sample_count, abi.fp(db), ptrs,
                                      ntex, eptrs, ndome, threads)
        for d, g in env.items():
            grads[("env", d)] = g
        return grads


class ReferenceScene(_SceneBase):
    """A scene loaded into the reference's CPU path tracer (brute force)."""
    prefix = "ref_"

    def __init__(self, scene):
        super().__init__(_Lib.load(REF_SO), scene)
        self.h = self._fn("scene_create", C.c_void_p, C.c_void_p)(self.m.ptr)

    def render(self, camera, settings, sample_start=0, sample_count=None):
        sample_count = settings.samplesPerPixel if sample_count is None else sample_count
        fb = FrameBuffer(settings.width, settings.height)
        f = fb.to_c()
        cam, st = camera.to_c(), settings.to_c()
        self._fn("render", C.c_int, C.c_void_p, _P(abi.Camera), _P(abi.Settings), _u, _u, _P(abi.Frame))(
            self.h, C.byref(cam), C.byref(st), sample_start, sample_count, C.byref(f))
        return fb

    def render_accumulated(self, camera, settings, batches):
        fb = FrameBuffer(settings.width, settings.height)
        f = fb.to_c()
        cam, st = camera.to_c(), settings.to_c()
        b = np.ascontiguousarray(batches, np.uint32)
        self._fn("render_accumulated", C.c_int, C.c_void_p, _P(abi.Camera), _P(abi.Settings), _up, _u,
                 _P(abi.Frame))(self.h, C.byref(cam), C.byref(st), abi.up(b), len(b), C.byref(f))
        return fb

    def intersect(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        self._fn("intersect", None, C.c_void_p, _fp, _u, _up, _fp)(self.h, abi.fp(rays), n, abi.up(prim), abi.fp(t))
        return prim, t


def _lib_fns(lib, prefix):
    class F:
        pass

    f = F()
    f.qmc = _bind(lib, prefix + "qmc", _f, _u, _u, _u, _u)
    f.radical_inverse = _bind(lib, prefix + "radical_inverse", _f, _u, _u)
    f.eval_bsdf = _bind(lib, prefix + "eval_bsdf", None, _P(abi.Material), _up, _fp, _fp, _fp, _fp, _u, _fp, _fp)
    f.lobe_weights = _bind(lib, prefix + "lobe_weights", None, _P(abi.Material), _fp, _fp)
    f.ggx_d = _bind(lib, prefix + "ggx_d", _f, _f, _f)
    f.ggx_g2 = _bind(lib, prefix + "ggx_g2", _f, _f, _f, _f)
    f.fresnel_dielectric = _bind(lib, prefix + "fresnel_dielectric", _f, _f, _f, _f)
    f.sample_ggx_half = _bind(lib, prefix + "sample_ggx_half", None, _fp, _f, _f, _f, _fp)
    f.sample_cosine = _bind(lib, prefix + "sample_cosine", None, _fp, _f, _f, _fp)
    return f


def oracle_fns():
    lib = _Lib.load(ORACLE_SO)
    f = _lib_fns(lib, "oracle_")
    f.qmc_fill = _bind(lib, "oracle_qmc_fill", None, _u, _u, _u, _u, _u, _u, _u, _fp)
    f.specular_pdf = _bind(lib, "oracle_specular_pdf", _f, _fp, _fp, _fp, _f)
    f.accum_add = _bind(lib, "oracle_accum_add", None, _u, _fp, _fp, _fp, _fp, _u, _fp, _fp, _fp, _fp, _up)
    f.accum_resolve = _bind(lib, "oracle_accum_resolve", None, _u, _u, _fp, _fp, _fp, _fp, _up, _fp, _fp, _fp,
                            _fp)
    f.bsdf_grad = _bind(lib, "oracle_bsdf_grad", None, _P(abi.Material), _up, _fp, _fp, _fp, _fp, _u, _fp, _fp)
    f.triangulate = _bind(lib, "oracle_triangulate", C.c_int64, _P(C.c_int32), _u, _P(C.c_int32), _u, _up,
                          C.c_uint64)
    f.smooth_normals = _bind(lib, "oracle_smooth_normals", None, _fp, _u, _up, _u, _fp)
    f.plant_mistake = _bind(lib, "oracle_plant_mistake", None, C.c_int)
    return f


# Planted mistakes of the oracle (oracle.cpp, enum Plant), for tests that
# show a fixture reaches the branch a mistake would change.
PLANTS = {
    "extinction_sign": 1, "pane_no_denominator": 2, "opacity_no_rescale": 3, "opacity_no_skip": 4,
    "delta_as_rough": 5, "ignore_normal": 6, "ignore_coat_normal": 7, "no_attenuation": 8,
    "tint_in_medium": 9, "no_override_flip": 10, "mirror_not_delta": 11,
}


class planted:
    """Context manager: the oracle computes with one planted mistake."""

    def __init__(self, name):
        self.code = PLANTS[name]

    def __enter__(self):
        oracle_fns().plant_mistake(self.code)
        return self

    def __exit__(self, *exc):
        oracle_fns().plant_mistake(0)
        return False


def triangulate(face_vertex_counts, face_vertex_indices):
    """The oracle's fan triangulation (SceneBridge.cpp:1009-1020); None for
    malformed input."""
    f = oracle_fns()
    counts = np.ascontiguousarray(face_vertex_counts, np.int32)
    idx = np.ascontiguousarray(face_vertex_indices, np.int32)
    cp, ip = counts.ctypes.data_as(_P(C.c_int32)), idx.ctypes.data_as(_P(C.c_int32))
    n = f.triangulate(cp, len(counts), ip, len(idx), None, 0)
    if n < 0:
        return None
    out = np.zeros(n, np.uint32)
    f.triangulate(cp, len(counts), ip, len(idx), abi.up(out), n)
    return out


def smooth_normals(positions, indices):
    """The oracle's ComputeSmoothNormals (SceneBridge.cpp:147-180)."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    out = np.zeros_like(pos)
    oracle_fns().smooth_normals(abi.fp(pos), len(pos), abi.up(idx), len(idx), abi.fp(out))
    return out


def reference_fns():
    lib = _Lib.load(REF_SO)
    f = _lib_fns(lib, "ref_")
    f.next2d = _bind(lib, "ref_next2d", None, _u, _u, _u, _u, _fp)
    return f


def eval_bsdf(fns, materials, mat_index, n, cn, v, l):
    """Run EvaluateBsdf/EvaluateBsdfPdf over tuples with the given library."""
    mats = (abi.Material * len(materials))(*[m.to_c() for m in materials])
    mi = np.ascontiguousarray(mat_index, np.uint32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (n, cn, v, l)]
    cnt = len(mi)
    val = np.zeros((cnt, 3), np.float32)
    pdf = np.zeros(cnt, np.float32)
    fns.eval_bsdf(C.cast(mats, _P(abi.Material)), abi.up(mi), *[abi.fp(a) for a in arrs], cnt, abi.fp(val),
                  abi.fp(pdf))
    return val, pdf


def bsdf_grad(materials, mat_index, n, cn, v, l):
    """Oracle d f / d baseColor and d f / d roughness over tuples."""
    fns = oracle_fns()
    mats = (abi.Material * len(materials))(*[m.to_c() for m in materials])
    mi = np.ascontiguousarray(mat_index, np.uint32)
    arrs = [np.ascontiguousarray(a, np.float32) for a in (n, cn, v, l)]
    cnt = len(mi)
    db = np.zeros((cnt, 3), np.float32)
    dr = np.zeros((cnt, 3), np.float32)
    fns.bsdf_grad(C.cast(mats, _P(abi.Material)), abi.up(mi), *[abi.fp(a) for a in arrs], cnt, abi.fp(db), abi.fp(dr))
    return db, dr


class Accumulator:
    """FrameAccumulator (Renderer.cpp:47-110) through the oracle."""

    def __init__(self, width, height):
        self.np = width * height
        self.w, self.h = width, height
        self.b = np.zeros((self.np, 4), np.float32)
        self.a = np.zeros((self.np, 3), np.float32)
        self.n = np.zeros((self.np, 3), np.float32)
        self.d = np.zeros(self.np, np.float32)
        self.c = np.zeros(self.np, np.uint32)
        self.total = 0
        self.f = oracle_fns()

    def add(self, fb, count):
        self.f.accum_add(self.np, abi.fp(fb.beauty), abi.fp(fb.albedo), abi.fp(fb.normal), abi.fp(fb.depth),
                         count, abi.fp(self.b), abi.fp(self.a), abi.fp(self.n), abi.fp(self.d), abi.up(self.c))
        self.total += count

    def resolve(self):
        fb = FrameBuffer(self.w, self.h)
        self.f.accum_resolve(self.np, self.total, abi.fp(self.b), abi.fp(self.a), abi.fp(self.n), abi.fp(self.d),
                             abi.up(self.c), abi.fp(fb.beauty), abi.fp(fb.albedo), abi.fp(fb.normal),
                             abi.fp(fb.depth))
        return fb
