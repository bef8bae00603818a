"""Host-side data model mirroring Shiro's runtime types.

Names, fields and defaults follow the reference's
include/shiro/render/Renderer.h:29-133 (RenderSettings, Camera, PbrMaterial,
TriangleMesh, DirectionalLight, DomeLight, Scene) and FrameBuffer.h:10-37,
so that a scene built for the reference reads the same here.  Geometry and
images are numpy arrays; `marshal()` turns a Scene into the POD structs of
include/shiro_hip.h without copying the arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields
from typing import List, Optional

import numpy as np

from . import abi

_V = lambda *v: np.array(v, dtype=np.float32)  # noqa: E731


@dataclass
class RenderSettings:
    """Renderer.h:29-45.  russianRouletteDepth extends the reference, which
    starts roulette at depth 2 (CpuPathTracer.cpp:2006)."""
    width: int = 1280
    height: int = 720
    samplesPerPixel: int = 32
    samplesPerUpdate: int = 1
    domeLightSamples: int = 1
    maxDepth: int = 3
    diffuseDepth: int = 2
    specularDepth: int = 2
    threadLimit: int = 0
    maxSubdivLevel: int = 2
    backend: int = 0
    backgroundVisible: bool = True
    enableHeadlight: bool = True
    enableCaustics: bool = True
    enableSss: bool = True
    russianRouletteDepth: int = 2
    transport: int = abi.TRANSPORT_MEGAKERNEL

    def to_c(self) -> abi.Settings:
        s = abi.Settings()
        for f in fields(self):
            setattr(s, f.name, int(getattr(self, f.name)))
        return s


@dataclass
class Camera:
    """Renderer.h:47-54."""
    position: np.ndarray = field(default_factory=lambda: _V(0, 0, 5))
    forward: np.ndarray = field(default_factory=lambda: _V(0, 0, -1))
    right: np.ndarray = field(default_factory=lambda: _V(1, 0, 0))
    up: np.ndarray = field(default_factory=lambda: _V(0, 1, 0))
    verticalFovDegrees: float = 45.0
    aspectRatio: float = 16.0 / 9.0

    @staticmethod
    def look_at(eye, target, up=(0, 1, 0), fov=45.0, aspect=16 / 9) -> "Camera":
        eye = np.asarray(eye, np.float64)
        f = np.asarray(target, np.float64) - eye
        f /= np.linalg.norm(f)
        r = np.cross(f, np.asarray(up, np.float64))
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        return Camera(eye.astype(np.float32), f.astype(np.float32), r.astype(np.float32),
                      u.astype(np.float32), float(fov), float(aspect))

    def to_c(self) -> abi.Camera:
        c = abi.Camera()
        for name in ("position", "forward", "right", "up"):
            getattr(c, name)[:] = [float(v) for v in getattr(self, name)]
        c.verticalFovDegrees = self.verticalFovDegrees
        c.aspectRatio = self.aspectRatio
        return c


_MATERIAL_VEC = ("baseColor", "emissionColor", "specularColor", "transmissionColor",
                 "transmissionScatter", "coatColor", "subsurfaceColor", "subsurfaceRadius",
                 "sheenColor", "normalOverride", "coatNormalOverride", "tangentOverride")


@dataclass
class PbrMaterial:
    """Renderer.h:56-102 (same fields and defaults).  baseColorTexture and
    roughnessTexture (-1 = none) extend the reference with per-texel
    parameters."""
    baseColor: tuple = (0.8, 0.8, 0.8)
    emissionColor: tuple = (0.0, 0.0, 0.0)
    specularColor: tuple = (1.0, 1.0, 1.0)
    transmissionColor: tuple = (1.0, 1.0, 1.0)
    transmissionScatter: tuple = (0.0, 0.0, 0.0)
    coatColor: tuple = (1.0, 1.0, 1.0)
    subsurfaceColor: tuple = (1.0, 1.0, 1.0)
    subsurfaceRadius: tuple = (1.0, 1.0, 1.0)
    sheenColor: tuple = (1.0, 1.0, 1.0)
    normalOverride: tuple = (0.0, 0.0, 0.0)
    coatNormalOverride: tuple = (0.0, 0.0, 0.0)
    tangentOverride: tuple = (0.0, 0.0, 0.0)
    baseWeight: float = 1.0
    emissionStrength: float = 0.0
    specularWeight: float = 1.0
    metallic: float = 0.0
    roughness: float = 0.5
    diffuseRoughness: float = 0.0
    opacity: float = 1.0
    specularAnisotropy: float = 0.0
    specularRotation: float = 0.0
    coatWeight: float = 0.0
    coatRoughness: float = 0.03
    coatIor: float = 1.5
    coatAnisotropy: float = 0.0
    coatRotation: float = 0.0
    sheen: float = 0.0
    sheenRoughness: float = 0.3
    subsurface: float = 0.0
    subsurfaceScale: float = 1.0
    subsurfaceAnisotropy: float = 0.0
    transmission: float = 0.0
    transmissionDepth: float = 0.0
    transmissionScatterAnisotropy: float = 0.0
    transmissionDispersion: float = 0.0
    transmissionExtraRoughness: float = 0.0
    ior: float = 1.5
    coatAffectColor: float = 0.0
    coatAffectRoughness: float = 0.0
    thinFilmThickness: float = 0.0
    thinFilmIor: float = 1.5
    thinWalled: bool = False
    hasNormalOverride: bool = False
    hasCoatNormalOverride: bool = False
    hasTangentOverride: bool = False
    baseColorTexture: int = -1
    roughnessTexture: int = -1

    def to_c(self) -> abi.Material:
        m = abi.Material()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name in _MATERIAL_VEC:
                getattr(m, f.name)[:] = [float(x) for x in v]
            elif isinstance(v, bool):
                setattr(m, f.name, int(v))
            elif f.name.endswith("Texture"):
                setattr(m, f.name, int(v))
            else:
                setattr(m, f.name, float(v))
        return m


def usd_preview_surface(diffuseColor=(0.18, 0.18, 0.18), roughness=0.5, metallic=0.0,
                        ior=1.5, opacity=1.0, emissiveColor=(0, 0, 0)) -> PbrMaterial:
    """UsdPreviewSurface -> PbrMaterial as the Hydra bridge maps it
    (SceneBridge.cpp:648-662): thinWalled = opacity < 1, unit specular, and
    emission strength 1 when any emissive component is positive."""
    m = PbrMaterial()
    m.baseColor = tuple(float(c) for c in diffuseColor)
    m.roughness = float(roughness)
    m.metallic = float(metallic)
    m.ior = float(ior)
    m.opacity = float(opacity)
    m.thinWalled = m.opacity < 1.0
    m.specularWeight = 1.0
    m.specularColor = (1.0, 1.0, 1.0)
    m.emissionColor = tuple(float(c) for c in emissiveColor)
    m.emissionStrength = 1.0 if max(m.emissionColor) > 0.0 else 0.0
    return m


def standard_surface(**params) -> PbrMaterial:
    """Autodesk standard_surface -> PbrMaterial (SceneBridge.cpp:664-772):
    unspecified parameters keep the PbrMaterial defaults; parameter names are
    the standard_surface ones (base, base_color, metalness, ...)."""
    names = {
        "base": "baseWeight", "base_color": "baseColor", "diffuse_roughness": "diffuseRoughness",
        "specular": "specularWeight", "specular_color": "specularColor",
        "specular_roughness": "roughness", "metalness": "metallic", "specular_IOR": "ior",
        "specular_anisotropy": "specularAnisotropy", "specular_rotation": "specularRotation",
        "transmission": "transmission", "transmission_color": "transmissionColor",
        "transmission_depth": "transmissionDepth", "transmission_scatter": "transmissionScatter",
        "transmission_extra_roughness": "transmissionExtraRoughness", "thin_walled": "thinWalled",
        "coat": "coatWeight", "coat_color": "coatColor", "coat_roughness": "coatRoughness",
        "coat_IOR": "coatIor", "coat_affect_color": "coatAffectColor",
        "coat_affect_roughness": "coatAffectRoughness", "subsurface": "subsurface",
        "subsurface_color": "subsurfaceColor", "subsurface_radius": "subsurfaceRadius",
        "subsurface_scale": "subsurfaceScale", "sheen": "sheen", "sheen_color": "sheenColor",
        "sheen_roughness": "sheenRoughness", "thin_film_thickness": "thinFilmThickness",
        "thin_film_IOR": "thinFilmIor", "emission": "emissionStrength",
        "emission_color": "emissionColor", "opacity": "opacity",
    }
    m = PbrMaterial()
    for k, v in params.items():
        setattr(m, names.get(k, k), tuple(float(x) for x in v) if isinstance(v, (tuple, list, np.ndarray)) else v)
    if m.emissionStrength <= 0.0 and max(m.emissionColor) > 0.0:
        m.emissionStrength = 1.0
    return m


@dataclass
class TriangleMesh:
    """Renderer.h:104-109.  positions (V,3) f32, normals (V,3) or empty,
    indices (3T,) u32, uvs (V,2) optional extension."""
    positions: np.ndarray
    normals: np.ndarray
    indices: np.ndarray
    materialIndex: int = 0
    uvs: Optional[np.ndarray] = None


@dataclass
class DirectionalLight:
    """Renderer.h:111-114: direction of travel and radiance."""
    direction: tuple = (0.0, -1.0, -1.0)
    radiance: tuple = (2.0, 2.0, 2.0)


@dataclass
class DomeLight:
    """Renderer.h:116-124; `pixels` (H,W,3) f32 replaces the file-loaded
    EnvironmentMap (row 0 = top)."""
    radiance: tuple = (1.0, 1.0, 1.0)
    pixels: Optional[np.ndarray] = None
    layout: int = abi.ENV_AUTOMATIC
    right: tuple = (1.0, 0.0, 0.0)
    up: tuple = (0.0, 1.0, 0.0)
    forward: tuple = (0.0, 0.0, 1.0)


@dataclass
class Texture:
    texels: np.ndarray  # (H, W, C) f32


@dataclass
class Scene:
    """Renderer.h:126-133."""
    materials: List[PbrMaterial] = field(default_factory=list)
    meshes: List[TriangleMesh] = field(default_factory=list)
    distantLights: List[DirectionalLight] = field(default_factory=list)
    domeLights: List[DomeLight] = field(default_factory=list)
    environmentTop: tuple = (0.0, 0.0, 0.0)
    environmentBottom: tuple = (0.0, 0.0, 0.0)
    textures: List[Texture] = field(default_factory=list)

    @property
    def triangle_count(self) -> int:
        return int(sum(len(m.indices) // 3 for m in self.meshes))


class Marshalled:
    """A C view of a Scene plus the arrays it points into (kept alive)."""

    def __init__(self, scene: Scene):
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        mats = (abi.Material * max(1, len(scene.materials)))()
        for i, m in enumerate(scene.materials):
            mats[i] = m.to_c()
        meshes = (abi.Mesh * max(1, len(scene.meshes)))()
        for i, m in enumerate(scene.meshes):
            p = arr(m.positions, np.float32).reshape(-1)
            n = arr(m.normals, np.float32).reshape(-1) if m.normals is not None and len(m.normals) else None
            ix = arr(m.indices, np.uint32).reshape(-1)
            uv = arr(m.uvs, np.float32).reshape(-1) if m.uvs is not None else None
            meshes[i].positions = abi.fp(p)
            meshes[i].normals = abi.fp(n) if n is not None else None
            meshes[i].uvs = abi.fp(uv) if uv is not None else None
            meshes[i].indices = abi.up(ix)
            meshes[i].vertexCount = len(p) // 3
            meshes[i].normalCount = len(n) // 3 if n is not None else 0
            meshes[i].indexCount = len(ix)
            meshes[i].materialIndex = int(m.materialIndex)
        dirs = (abi.DirectionalLight * max(1, len(scene.distantLights)))()
        for i, d in enumerate(scene.distantLights):
            dirs[i].direction[:] = [float(v) for v in d.direction]
            dirs[i].radiance[:] = [float(v) for v in d.radiance]
        domes = (abi.DomeLight * max(1, len(scene.domeLights)))()
        for i, d in enumerate(scene.domeLights):
            for name in ("radiance", "right", "up", "forward"):
                getattr(domes[i], name)[:] = [float(v) for v in getattr(d, name)]
            if d.pixels is not None:
                px = arr(d.pixels, np.float32)
                domes[i].pixels = abi.fp(px)
                domes[i].height, domes[i].width = px.shape[0], px.shape[1]
            else:
                domes[i].pixels = None
                domes[i].width = domes[i].height = 0
            domes[i].layout = int(d.layout)
        texs = (abi.Texture * max(1, len(scene.textures)))()
        for i, t in enumerate(scene.textures):
            tx = arr(t.texels, np.float32)
            if tx.ndim == 2:
                tx = tx[:, :, None]
            texs[i].texels = abi.fp(tx)
            texs[i].height, texs[i].width, texs[i].channels = tx.shape
        s = abi.Scene()
        s.materials = C.cast(mats, C.POINTER(abi.Material))
        s.materialCount = len(scene.materials)
        s.meshes = C.cast(meshes, C.POINTER(abi.Mesh))
        s.meshCount = len(scene.meshes)
        s.distantLights = C.cast(dirs, C.POINTER(abi.DirectionalLight))
        s.distantLightCount = len(scene.distantLights)
        s.domeLights = C.cast(domes, C.POINTER(abi.DomeLight))
        s.domeLightCount = len(scene.domeLights)
        s.environmentTop[:] = [float(v) for v in scene.environmentTop]
        s.environmentBottom[:] = [float(v) for v in scene.environmentBottom]
        s.textures = C.cast(texs, C.POINTER(abi.Texture))
        s.textureCount = len(scene.textures)
        self.c = s
        self._keep = (keep, mats, meshes, dirs, domes, texs)

    @property
    def ptr(self):
        return C.byref(self.c)


def marshal(scene: Scene) -> Marshalled:
    return Marshalled(scene)


class FrameBuffer:
    """FrameBuffer.h:10-37 on host numpy arrays (beauty RGBA, albedo,
    normal, depth; row y*W + x, y = 0 at the bottom)."""

    def __init__(self, width: int, height: int):
        self.width, self.height = int(width), int(height)
        n = self.width * self.height
        self.beauty = np.zeros((n, 4), np.float32)
        self.albedo = np.zeros((n, 3), np.float32)
        self.normal = np.zeros((n, 3), np.float32)
        self.depth = np.full((n,), np.inf, np.float32)

    def to_c(self) -> abi.Frame:
        f = abi.Frame()
        f.beauty, f.albedo = abi.fp(self.beauty), abi.fp(self.albedo)
        f.normal, f.depth = abi.fp(self.normal), abi.fp(self.depth)
        f.width, f.height = self.width, self.height
        return f

    def image(self) -> np.ndarray:
        """Beauty RGB as (H, W, 3), row 0 = bottom."""
        return self.beauty[:, :3].reshape(self.height, self.width, 3)
