"""Shared helpers for the tests: metrics and the small workloads the oracle
finishes in seconds."""
import numpy as np

from shiro_amd import workloads


def relmse(x, y):
    """Relative MSE over beauty RGB: mean((x-y)^2 / (y^2 + 1e-2))."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return float(np.mean((x - y) ** 2 / (y * y + 1e-2)))


def c1_mini(size=64, spp=8):
    return workloads.c1(scale=size / 512, spp=spp)


def emissive_mini(size=64, spp=8):
    """c1 with an emissive quad and an emissive sphere (area lights)."""
    return workloads.emissive(scale=size / 512, spp=spp)


def c2_mini(width=96, spp=8, env=(256, 128)):
    return workloads.c2(scale=width / 1920, spp=spp, env_size=env)


def angular_mini(width=96, spp=8):
    """c2's geometry under a rotated angular (light-probe) dome."""
    from golden import golden_cases
    return golden_cases.angular_workload(scale=width / 1920, spp=spp)


def c3_mini(width=96, spp=8, objects=12, env=(128, 64)):
    """c3's material mix and lighting on a few low-resolution objects."""
    return workloads.c3(scale=width / 3840, spp=spp, objects=objects, sphere_freq=8, box_n=6, env_size=env)


def random_rays(scene, n, seed=0):
    """Rays from points inside the scene bounds in uniform random directions
    (n x 8: origin, direction, tmin, tmax)."""
    rng = np.random.default_rng(seed)
    pts = np.concatenate([m.positions for m in scene.meshes])
    lo, hi = pts.min(0), pts.max(0)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3:6] = d
    r[:, 6] = 1e-4
    r[:, 7] = np.float32(3.402823466e38)
    return r


def c5_mini(size=32, spp=4, frequency=12, tex=16, env=(64, 32)):
    """c5's textured sphere at a size the oracle renders in seconds."""
    return workloads.c5(scale=size / 256, spp=spp, frequency=frequency, tex_size=tex, env_size=env)


def c5_scene_with(wl, albedo, rough):
    """The c5 scene with its two textures replaced."""
    import copy
    from shiro_amd.scene import Texture
    scene = copy.copy(wl.scene)
    scene.textures = [Texture(np.ascontiguousarray(albedo, np.float32)), Texture(np.ascontiguousarray(rough, np.float32))]
    return scene
