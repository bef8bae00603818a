"""Seeded synthetic workloads (scenes, cameras, settings).

Every workload is generated from a base seed, so the same call yields the same
arrays on any machine; the HIP backend and the CPU oracle are fed the same
Scene object.  Meshes get smooth normals the way Shiro's Hydra bridge computes
them (area-weighted face normals, SceneBridge.cpp:147-180); materials are
mapped as the bridge maps UsdPreviewSurface / standard_surface
(SceneBridge.cpp:648-772).

Workloads (sizes exact):
  c1  512x512, 16 spp, depth 4: ground quad + icosphere (subdiv 4) = 5,122
      triangles, UsdPreviewSurface diffuse+GGX r=0.4, one directional light.
  c2  1920x1080, 64 spp, depth 8: 250x200-quad value-noise heightfield
      (100,000 tris) + 64 icospheres (subdiv 3, 1,280 tris each) = 181,920
      triangles; standard_surface, metalness {0,1}, roughness U[0.05,1],
      base_color U[0,1]^3; 2048x1024 dome (log-normal background, 3 hot spots
      of peak 1e4).
  c3  3840x2160, 16 spp, depth 16, roulette from depth 3: 250 displaced
      geodesic spheres (12,500 tris) + 250 displaced boxes (7,500 tris) =
      5,000,000 triangles at random transforms in a 100-unit cube; 40% diffuse,
      30% GGX metal, 20% dielectric, 10% coated; c2's dome + one directional.
  c4  32 scenes in one batch (batched multi-scene rendering): square
      frames of side U[256, 1024], spp in {4, 16, 64}, triangles log-uniform
      in [10k, 500k] (a value-noise heightfield plus 1,280-triangle spheres,
      c2's standard_surface mix), each with its own 1024x512 dome; depth 8.
  c5  256x256, 32 spp, depth 4 (texture extension): UV-mapped geodesic
      sphere (frequency 50, 50,000 tris) on a 2-triangle ground; the sphere's
      base colour (128x128x3) and roughness (128x128x1) are textures; a
      256x128 dome.  `textures_truth` holds the seeded target texels,
      `textures_init` the (non-grey, jittered) starting point of an
      optimisation.
Every workload accepts a `scale` that shrinks the image (and spp where
stated) for parity tests at sizes the oracle finishes in seconds.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np

from . import abi
from .scene import (Camera, DirectionalLight, DomeLight, PbrMaterial, RenderSettings, Scene, Texture,
                    TriangleMesh, standard_surface, usd_preview_surface)

BASE_SEED = 0x5348524F


# ---------------------------------------------------------------------------
# Geometry
# ---------------------------------------------------------------------------
def smooth_normals(positions: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Area-weighted vertex normals, zero sums -> (0,1,0) (SceneBridge.cpp:147-180)."""
    p = positions.astype(np.float32)
    tri = indices.reshape(-1, 3)
    fn = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]]).astype(np.float32)
    ok = np.linalg.norm(fn, axis=1) > 1e-8
    acc = np.zeros_like(p)
    for k in range(3):
        np.add.at(acc, tri[ok, k], fn[ok])
    ln = np.linalg.norm(acc, axis=1)
    out = np.empty_like(p)
    good = ln > 0
    out[good] = acc[good] / ln[good, None]
    out[~good] = (0.0, 1.0, 0.0)
    return out.astype(np.float32)


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def _weld(points: np.ndarray, tris: np.ndarray, decimals: int = 6):
    key = np.round(points, decimals)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    # Use the first original point of each class (exact, not the rounded key).
    first = np.full(len(uniq), -1, np.int64)
    order = np.arange(len(points))[::-1]
    first[inv[order]] = order
    return points[first], inv[tris]


def geodesic_sphere(frequency: int) -> Tuple[np.ndarray, np.ndarray]:
    """Unit geodesic sphere: each icosahedron face split into a frequency^2
    triangle grid, projected to the sphere; 20*frequency^2 triangles.
    frequency 2^k is an icosphere subdivided k times (same triangle count)."""
    v, faces = _icosahedron()
    f = int(frequency)
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    m = (ii + jj) <= f
    ii, jj = ii[m], jj[m]
    lid = -np.ones((f + 1, f + 1), np.int64)
    lid[ii, jj] = np.arange(len(ii))
    a_i, a_j = np.meshgrid(np.arange(f), np.arange(f), indexing="ij")
    up = (a_i + a_j) < f
    t1 = np.stack([lid[a_i[up], a_j[up]], lid[a_i[up] + 1, a_j[up]], lid[a_i[up], a_j[up] + 1]], 1)
    dn = (a_i + a_j) < f - 1
    t2 = np.stack([lid[a_i[dn] + 1, a_j[dn]], lid[a_i[dn] + 1, a_j[dn] + 1], lid[a_i[dn], a_j[dn] + 1]], 1)
    local = np.concatenate([t1, t2])
    pts, tris = [], []
    n = len(ii)
    for k, (a, b, c) in enumerate(faces):
        p = v[a][None] + (v[b] - v[a])[None] * (ii / f)[:, None] + (v[c] - v[a])[None] * (jj / f)[:, None]
        pts.append(p / np.linalg.norm(p, axis=1, keepdims=True))
        tris.append(local + k * n)
    pts = np.concatenate(pts)
    tris = np.concatenate(tris)
    pts, tris = _weld(pts, tris)
    return pts, tris.reshape(-1)


def box_mesh(n: int) -> Tuple[np.ndarray, np.ndarray]:
    """Unit cube [-1,1]^3, each face an n x n quad grid (12 n^2 triangles);
    faces keep their own vertices (sharp edges)."""
    g = np.linspace(-1.0, 1.0, n + 1)
    uu, vv = np.meshgrid(g, g, indexing="ij")
    uu, vv = uu.reshape(-1), vv.reshape(-1)
    qi, qj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    qi, qj = qi.reshape(-1), qj.reshape(-1)
    a = qi * (n + 1) + qj
    b = (qi + 1) * (n + 1) + qj
    c = (qi + 1) * (n + 1) + qj + 1
    d = qi * (n + 1) + qj + 1
    quad = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
    pts, tris = [], []
    # (axis, sign): u, v span the other two axes with outward winding.
    for k, (ax, s) in enumerate([(0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)]):
        o1, o2 = [x for x in range(3) if x != ax]
        p = np.zeros((len(uu), 3))
        p[:, ax] = s
        if s > 0:
            p[:, o1], p[:, o2] = uu, vv
        else:
            p[:, o1], p[:, o2] = vv, uu
        pts.append(p)
        tris.append(quad + k * len(uu))
    return np.concatenate(pts), np.concatenate(tris).reshape(-1)


class ValueNoise3:
    """Seeded trilinear value noise on an integer lattice (hash table)."""

    def __init__(self, seed: int, size: int = 256):
        rng = np.random.default_rng(seed)
        self.size = size
        self.table = rng.uniform(-1.0, 1.0, size).astype(np.float64)
        self.perm = rng.permutation(size)

    def _h(self, i, j, k):
        s = self.size
        return self.table[self.perm[(self.perm[(self.perm[i % s] + j) % s] + k) % s]]

    def __call__(self, p: np.ndarray) -> np.ndarray:
        f = np.floor(p)
        t = p - f
        t = t * t * (3 - 2 * t)
        i = f.astype(np.int64)
        out = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (t[:, 0] if dx else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1]) * \
                        (t[:, 2] if dz else 1 - t[:, 2])
                    out = out + w * self._h(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz)
        return out

    def fbm(self, p: np.ndarray, octaves: int = 4) -> np.ndarray:
        out, amp, freq = 0.0, 1.0, 1.0
        for _ in range(octaves):
            out = out + amp * self(p * freq)
            amp *= 0.5
            freq *= 2.0
        return out


def displaced_sphere(frequency: int, noise: ValueNoise3, amplitude: float, feature: float):
    p, ix = geodesic_sphere(frequency)
    r = 1.0 + amplitude * noise.fbm(p * feature, 3)
    return p * r[:, None], ix


def displaced_box(n: int, noise: ValueNoise3, amplitude: float, feature: float):
    p, ix = box_mesh(n)
    # Push along the dominant axis (the face normal) so faces stay closed at
    # shared edges: the offset is a function of position only.
    d = amplitude * noise.fbm(p * feature, 3)
    return p * (1.0 + d)[:, None], ix


def random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def make_mesh(p, ix, material: int) -> TriangleMesh:
    p = np.ascontiguousarray(p, np.float32)
    ix = np.ascontiguousarray(ix, np.uint32)
    return TriangleMesh(p, smooth_normals(p, ix), ix, material)


# ---------------------------------------------------------------------------
# Environment map
# ---------------------------------------------------------------------------
def _upsample(img: np.ndarray, height: int, width: int) -> np.ndarray:
    """Bilinear upsampling of an (h, w, c) grid of texel centres to
    (height, width, c); u wraps, v clamps."""
    h, w = img.shape[:2]
    v = np.clip((np.arange(height) + 0.5) / height * h - 0.5, 0, h - 1)
    u = (np.arange(width) + 0.5) / width * w - 0.5
    v0 = np.floor(v).astype(np.int64)
    v1 = np.minimum(v0 + 1, h - 1)
    fv = (v - v0)[:, None, None]
    u0 = np.floor(u).astype(np.int64)
    fu = (u - u0)[None, :, None]
    u1 = (u0 + 1) % w
    u0 = u0 % w
    top = img[v0][:, u0] * (1 - fu) + img[v0][:, u1] * fu
    bot = img[v1][:, u0] * (1 - fu) + img[v1][:, u1] * fu
    return top * (1 - fv) + bot * fv


def procedural_dome(width: int, height: int, seed: int, hot_spots: int = 3, peak: float = 1.0e4,
                    spot_width_deg: float = 1.5) -> np.ndarray:
    """(H, W, 3) fp32 lat-long map: log-normal background (log radiance is a
    smooth field with mean 0 and standard deviation 1 per channel, slightly
    decorrelated between channels; evaluated on a grid 4x coarser and
    upsampled bilinearly) plus Gaussian hot spots of the given peak in the
    upper hemisphere.  Row 0 is theta = 0 (+Y)."""
    rng = np.random.default_rng(seed)

    def dirs(h, w):
        theta = (np.arange(h) + 0.5) / h * np.pi
        phi = (np.arange(w) + 0.5) / w * 2 * np.pi
        st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
        return np.stack(np.broadcast_arrays(np.sin(phi)[None] * st, ct, np.cos(phi)[None] * st), -1).reshape(-1, 3)

    ch, cw = max(2, height // 4), max(4, width // 4)
    coarse = dirs(ch, cw)
    base = ValueNoise3(seed).fbm(coarse * 3.0 + 11.0, 5)
    log_rad = np.empty((ch * cw, 3))
    for c in range(3):
        n = 0.85 * base + 0.15 * ValueNoise3(seed + 101 + c).fbm(coarse * 3.0 + 7.0, 5)
        log_rad[:, c] = (n - n.mean()) / (n.std() + 1e-12)
    img = np.exp(_upsample(log_rad.reshape(ch, cw, 3), height, width)).reshape(-1, 3)
    flat = dirs(height, width)
    sw = np.deg2rad(spot_width_deg)
    for _ in range(hot_spots):
        th = rng.uniform(0.15, 0.45) * np.pi
        ph = rng.uniform(0, 2 * np.pi)
        c = np.array([np.sin(ph) * np.sin(th), np.cos(th), np.cos(ph) * np.sin(th)])
        ang = np.arccos(np.clip(flat @ c, -1, 1))
        img += peak * np.exp(-0.5 * (ang / sw) ** 2)[:, None]
    return img.reshape(height, width, 3).astype(np.float32)


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
@dataclass
class Workload:
    name: str
    scene: Scene
    camera: Camera
    settings: RenderSettings

    @property
    def samples(self) -> int:
        s = self.settings
        return s.width * s.height * s.samplesPerPixel


def _settings(w, h, spp, depth, rr=2) -> RenderSettings:
    return RenderSettings(width=w, height=h, samplesPerPixel=spp, samplesPerUpdate=spp,
                          domeLightSamples=1, maxDepth=depth, diffuseDepth=depth,
                          specularDepth=depth, russianRouletteDepth=rr)


def c1(scale: float = 1.0, spp: int = 16, seed: int = BASE_SEED) -> Workload:
    rng = np.random.default_rng(seed + 1)
    w = h = max(8, int(round(512 * scale)))
    mats = [usd_preview_surface(tuple(rng.uniform(0.2, 0.8, 3)), roughness=0.4) for _ in range(2)]
    quad = np.array([[-6, 0, -6], [6, 0, -6], [6, 0, 6], [-6, 0, 6]], np.float32)
    qi = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    sp, si = geodesic_sphere(16)
    sp = sp * 1.0 + np.array([0.0, 1.0, 0.0])
    scene = Scene(materials=mats, meshes=[make_mesh(quad, qi, 0), make_mesh(sp, si, 1)])
    # Direction towards the light is uniform on the upper hemisphere; the
    # light travels the opposite way.
    z = rng.uniform(0.0, 1.0)
    ph = rng.uniform(0, 2 * np.pi)
    r = np.sqrt(1 - z * z)
    to_light = np.array([r * np.cos(ph), z, r * np.sin(ph)])
    scene.distantLights.append(DirectionalLight(tuple(-to_light), (3.0, 3.0, 3.0)))
    cam = Camera.look_at((0.0, 2.2, 5.5), (0.0, 0.8, 0.0), fov=45.0, aspect=w / h)
    return Workload("c1", scene, cam, _settings(w, h, spp, 4))


def c2(scale: float = 1.0, spp: int = 64, seed: int = BASE_SEED, env_size=(2048, 1024)) -> Workload:
    rng = np.random.default_rng(seed + 2)
    noise = ValueNoise3(seed + 20)
    nx, nz = 250, 200
    xs = np.linspace(-50, 50, nx + 1)
    zs = np.linspace(-40, 40, nz + 1)
    gx, gz = np.meshgrid(xs, zs, indexing="ij")
    pts = np.stack([gx.reshape(-1), np.zeros(gx.size), gz.reshape(-1)], 1)
    pts[:, 1] = 3.0 * noise.fbm(pts * 0.06, 5)
    qi, qj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    qi, qj = qi.reshape(-1), qj.reshape(-1)
    a = qi * (nz + 1) + qj
    b = a + 1
    c = (qi + 1) * (nz + 1) + qj + 1
    d = (qi + 1) * (nz + 1) + qj
    hf_idx = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]).reshape(-1)

    def mat():
        return standard_surface(base_color=tuple(rng.uniform(0, 1, 3)),
                                metalness=float(rng.integers(0, 2)),
                                specular_roughness=float(rng.uniform(0.05, 1.0)))

    mats = [mat()]
    meshes = [make_mesh(pts, hf_idx, 0)]
    sp, si = geodesic_sphere(8)
    for k in range(64):
        rad = rng.uniform(1.0, 3.0)
        cx, cz = rng.uniform(-40, 40), rng.uniform(-30, 30)
        cy = 3.0 * noise.fbm(np.array([[cx, 0.0, cz]]) * 0.06, 5)[0] + rad * 0.8
        mats.append(mat())
        meshes.append(make_mesh(sp * rad + np.array([cx, cy, cz]), si, k + 1))
    scene = Scene(materials=mats, meshes=meshes)
    ew, eh = env_size
    scene.domeLights.append(DomeLight(pixels=procedural_dome(ew, eh, seed + 3)))
    w = max(8, int(round(1920 * scale)))
    h = max(8, int(round(1080 * scale)))
    cam = Camera.look_at((0.0, 22.0, 55.0), (0.0, 0.0, 0.0), fov=45.0, aspect=w / h)
    return Workload("c2", scene, cam, _settings(w, h, spp, 8))


def c3(scale: float = 1.0, spp: int = 16, seed: int = BASE_SEED, objects: int = 250,
       sphere_freq: int = 25, box_n: int = 25, env_size=(2048, 1024)) -> Workload:
    rng = np.random.default_rng(seed + 4)
    noise = ValueNoise3(seed + 40)
    sp0, si0 = geodesic_sphere(sphere_freq)
    bp0, bi0 = box_mesh(box_n)
    mats, meshes = [], []

    def material(kind: str) -> PbrMaterial:
        col = tuple(rng.uniform(0.05, 0.95, 3))
        if kind == "diffuse":
            return standard_surface(base_color=col, metalness=0.0, specular=0.0,
                                    specular_roughness=float(rng.uniform(0.3, 1.0)))
        if kind == "metal":
            return standard_surface(base_color=col, metalness=1.0,
                                    specular_roughness=float(rng.uniform(0.05, 0.6)))
        if kind == "glass":
            return standard_surface(base=0.0, transmission=1.0, transmission_color=tuple(rng.uniform(0.6, 1.0, 3)),
                                    specular_IOR=float(rng.uniform(1.33, 1.7)),
                                    specular_roughness=float(rng.uniform(0.0, 0.3)), thin_walled=False)
        return standard_surface(base_color=col, metalness=0.0, coat=1.0,
                                coat_roughness=float(rng.uniform(0.0, 0.2)),
                                specular_roughness=float(rng.uniform(0.2, 0.8)))

    kinds = rng.choice(["diffuse", "metal", "glass", "coated"], size=2 * objects, p=[0.4, 0.3, 0.2, 0.1])
    for k in range(2 * objects):
        is_sphere = k < objects
        local_noise = ValueNoise3(seed + 1000 + k, 64)
        if is_sphere:
            p = sp0 * (1.0 + 0.15 * local_noise.fbm(sp0 * 2.5, 3))[:, None]
            ix = si0
        else:
            p = bp0 * (1.0 + 0.08 * local_noise.fbm(bp0 * 2.5, 3))[:, None]
            ix = bi0
        R = random_rotation(rng)
        s = rng.uniform(1.0, 3.5)
        t = rng.uniform(-50, 50, 3)
        mats.append(material(str(kinds[k])))
        meshes.append(make_mesh((p @ R.T) * s + t, ix, k))
    scene = Scene(materials=mats, meshes=meshes)
    ew, eh = env_size
    scene.domeLights.append(DomeLight(pixels=procedural_dome(ew, eh, seed + 3)))
    d = np.array([-0.4, -1.0, -0.3])
    scene.distantLights.append(DirectionalLight(tuple(d / np.linalg.norm(d)), (3.0, 3.0, 3.0)))
    w = max(8, int(round(3840 * scale)))
    h = max(8, int(round(2160 * scale)))
    cam = Camera.look_at((0.0, 40.0, 140.0), (0.0, 0.0, 0.0), fov=45.0, aspect=w / h)
    return Workload("c3", scene, cam, _settings(w, h, spp, 16, rr=3))


def emissive(scale: float = 1.0, spp: int = 16, seed: int = BASE_SEED) -> Workload:
    """c1 with area lights: an emissive quad above the sphere (2 triangles,
    warm white, strength 4) and a small emissive icosphere (80 triangles,
    blue) beside it, next to c1's directional light.  Exercises the
    emissive-triangle light list, its sampling and the MIS of emission at
    hits."""
    wl = c1(scale=scale, spp=spp, seed=seed)
    lamp = usd_preview_surface((0.0, 0.0, 0.0), roughness=0.5, emissiveColor=(1.0, 0.9, 0.7))
    lamp.emissionStrength = 4.0
    blue = usd_preview_surface((0.1, 0.1, 0.1), roughness=0.3, emissiveColor=(0.3, 0.5, 1.0))
    blue.emissionStrength = 2.0
    scene = wl.scene
    scene.materials = list(scene.materials) + [lamp, blue]
    q = np.array([[-1.2, 3.2, -1.0], [1.2, 3.2, -1.0], [1.2, 3.4, 1.0], [-1.2, 3.4, 1.0]], np.float32)
    qi = np.array([0, 1, 2, 0, 2, 3], np.uint32)  # facing down
    sp, si = geodesic_sphere(2)
    scene.meshes = list(scene.meshes) + [make_mesh(q, qi, len(scene.materials) - 2),
                                         make_mesh(sp * 0.35 + np.array([1.8, 0.35, 0.6]), si,
                                                   len(scene.materials) - 1)]
    wl.name = "emissive"
    return wl


def c4_scene(index: int, seed: int = BASE_SEED, env_size=(1024, 512), scale: float = 1.0,
             spp_scale: float = 1.0, tri_scale: float = 1.0) -> Workload:
    """Scene `index` of the c4 batch (parameters drawn from a per-index
    seed); the scales shrink frames, spp and triangle counts for tests."""
    rng = np.random.default_rng(seed + 4000 + index)
    side = int(rng.integers(256, 1025))
    spp = int(rng.choice([4, 16, 64]))
    tris = int(np.exp(rng.uniform(np.log(10_000), np.log(500_000))) * tri_scale)
    noise = ValueNoise3(seed + 4100 + index)
    sphere_tris = 1280
    nsph = int(min(64, max(1, 0.3 * tris // sphere_tris)))
    quads = max(8, (tris - nsph * sphere_tris) // 2)
    nx = max(2, int(np.sqrt(quads * 1.25)))
    nz = max(2, quads // nx)
    xs = np.linspace(-50, 50, nx + 1)
    zs = np.linspace(-40, 40, nz + 1)
    gx, gz = np.meshgrid(xs, zs, indexing="ij")
    pts = np.stack([gx.reshape(-1), np.zeros(gx.size), gz.reshape(-1)], 1)
    pts[:, 1] = 3.0 * noise.fbm(pts * 0.06, 4)
    qi, qj = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    qi, qj = qi.reshape(-1), qj.reshape(-1)
    a = qi * (nz + 1) + qj
    c = (qi + 1) * (nz + 1) + qj + 1
    hf_idx = np.concatenate([np.stack([a, a + 1, c], 1), np.stack([a, c, c - 1], 1)]).reshape(-1)

    def mat():
        return standard_surface(base_color=tuple(rng.uniform(0, 1, 3)), metalness=float(rng.integers(0, 2)),
                                specular_roughness=float(rng.uniform(0.05, 1.0)))

    mats = [mat()]
    meshes = [make_mesh(pts, hf_idx, 0)]
    sp, si = geodesic_sphere(8)
    for k in range(nsph):
        rad = rng.uniform(1.0, 3.0)
        cx, cz = rng.uniform(-40, 40), rng.uniform(-30, 30)
        cy = 3.0 * noise.fbm(np.array([[cx, 0.0, cz]]) * 0.06, 4)[0] + rad * 0.8
        mats.append(mat())
        meshes.append(make_mesh(sp * rad + np.array([cx, cy, cz]), si, k + 1))
    scene = Scene(materials=mats, meshes=meshes)
    ew, eh = env_size
    scene.domeLights.append(DomeLight(pixels=procedural_dome(ew, eh, seed + 4200 + index)))
    w = h = max(8, int(round(side * scale)))
    spp = max(1, int(round(spp * spp_scale)))
    az = rng.uniform(0, 2 * np.pi)
    eye = (60.0 * np.sin(az), 25.0, 60.0 * np.cos(az))
    cam = Camera.look_at(eye, (0.0, 0.0, 0.0), fov=45.0, aspect=1.0)
    return Workload(f"c4[{index}]", scene, cam, _settings(w, h, spp, 8))


def c4(count: int = 32, seed: int = BASE_SEED, **kw):
    """The c4 batch: `count` scenes (see c4_scene)."""
    return [c4_scene(i, seed, **kw) for i in range(count)]


def sphere_uvs(p: np.ndarray) -> np.ndarray:
    """Lat-long UVs of points on a sphere around the origin: u = longitude,
    v = polar angle / pi (0 at +Y)."""
    d = p / np.linalg.norm(p, axis=1, keepdims=True)
    u = np.arctan2(d[:, 0], d[:, 2]) / (2 * np.pi) + 0.5
    v = np.arccos(np.clip(d[:, 1], -1.0, 1.0)) / np.pi
    return np.stack([u, v], 1).astype(np.float32)


def c5_textures(seed: int, size: int = 128):
    """Seeded target texels: smooth base colour in [0.1, 0.9]^3 and
    roughness in [0.15, 0.85] (value-noise fields on the texture plane)."""
    yy, xx = np.meshgrid((np.arange(size) + 0.5) / size, (np.arange(size) + 0.5) / size, indexing="ij")
    pts = np.stack([xx.ravel() * 6.0, yy.ravel() * 6.0, np.zeros(size * size)], 1)
    cols = []
    for c in range(3):
        n = ValueNoise3(seed + 500 + c).fbm(pts + c * 17.0, 4)
        cols.append(0.5 + 0.4 * np.tanh(1.5 * n))
    albedo = np.stack(cols, 1).reshape(size, size, 3).astype(np.float32)
    r = ValueNoise3(seed + 600).fbm(pts * 0.7 + 3.0, 4)
    rough = (0.5 + 0.35 * np.tanh(1.5 * r)).reshape(size, size, 1).astype(np.float32)
    return albedo, rough


def c5(scale: float = 1.0, spp: int = 32, seed: int = BASE_SEED, frequency: int = 50, tex_size: int = 128,
       env_size=(256, 128)) -> Workload:
    p, ix = geodesic_sphere(frequency)
    p = p.astype(np.float32)
    sphere = make_mesh(p + np.array([0.0, 1.0, 0.0], np.float32), ix, 0)
    sphere.uvs = sphere_uvs(p)
    quad = np.array([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.float32)
    qi = np.array([0, 2, 1, 0, 3, 2], np.uint32)
    mat = usd_preview_surface((0.5, 0.5, 0.5), roughness=0.5)
    mat.baseColorTexture, mat.roughnessTexture = 0, 1
    ground = usd_preview_surface((0.35, 0.35, 0.35), roughness=0.7)
    albedo, rough = c5_textures(seed + 5, tex_size)
    # Starting texels: a non-grey base colour with a small seeded jitter.
    # With grey texels every throughput channel ties, and the max in the
    # reference's clamp_max_component has no derivative at a tie.
    jr = np.random.default_rng(seed + 7)
    init_albedo = (np.array([0.55, 0.5, 0.45], np.float32) +
                   jr.uniform(-0.02, 0.02, albedo.shape)).astype(np.float32)
    init_rough = (0.5 + jr.uniform(-0.02, 0.02, rough.shape)).astype(np.float32)
    scene = Scene(materials=[mat, ground], meshes=[sphere, make_mesh(quad, qi, 1)],
                  textures=[Texture(albedo.copy()), Texture(rough.copy())])
    ew, eh = env_size
    env = procedural_dome(ew, eh, seed + 6, hot_spots=2, peak=40.0, spot_width_deg=6.0)
    scene.domeLights.append(DomeLight(pixels=env.copy()))
    # Starting dome map: the target's contrast reduced towards its mean
    # colour (hot spots dimmed, still bright where the target is).
    init_env = (0.6 * env + 0.4 * env.mean(axis=(0, 1), keepdims=True)).astype(np.float32)
    w = h = max(8, int(round(256 * scale)))
    cam = Camera.look_at((0.0, 1.6, 3.6), (0.0, 0.9, 0.0), fov=40.0, aspect=1.0)
    wl = Workload("c5", scene, cam, _settings(w, h, spp, 4, rr=4))
    # 65,536 pixels: the wavefront packs all 32 samples into one pass, the
    # megakernel (one thread per pixel) would leave most of the GPU idle.
    wl.settings.transport = abi.TRANSPORT_WAVEFRONT
    wl.textures_truth = (albedo, rough)
    wl.textures_init = (init_albedo, init_rough)
    wl.env_truth = env
    wl.env_init = init_env
    return wl


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3, "c5": c5, "emissive": emissive}
