"""Inputs of the golden fixtures, shared by make_golden.py (which runs them
through the reference build) and the tests (which run them through the CPU
oracle and the HIP backend)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from shiro_amd import workloads  # noqa: E402
from shiro_amd.scene import PbrMaterial  # noqa: E402


def qmc_grid():
    px = np.array([0, 1, 2, 7, 100, 1919, 4095], np.uint32)
    py = np.array([0, 1, 3, 255, 1079, 2159], np.uint32)
    s = np.array([0, 1, 2, 15, 63, 64, 1000, 65535, 1 << 20], np.uint32)
    dims = np.concatenate([np.arange(0, 64), [256, 257, 512, 513, 768, 1024, 1792, 2049, 4097, 16384, 16385]]).astype(
        np.uint32)
    return px, py, s, dims


def radical_inverse_cases():
    bases = np.array([2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89],
                     np.uint32)
    idx = np.concatenate([np.arange(0, 300), [1023, 1024, 1025, 65535, 65536, 1 << 20, (1 << 24) + 1, 0xFFFFFFFF]]
                         ).astype(np.uint32)
    return bases, idx


def bsdf_materials():
    rng = np.random.default_rng(11)
    mats = []
    for i in range(40):
        m = PbrMaterial()
        m.baseColor = tuple(rng.uniform(0, 1, 3))
        m.baseWeight = float(rng.uniform(0.5, 1.0))
        m.metallic = float(rng.choice([0.0, 1.0, rng.uniform()]))
        m.roughness = float(rng.choice([0.0, 0.01, rng.uniform(), 1.0]))
        m.specularWeight = float(rng.uniform(0.0, 1.0))
        m.specularColor = tuple(rng.uniform(0.5, 1.0, 3))
        m.ior = float(rng.uniform(1.0, 2.5))
        m.diffuseRoughness = float(rng.uniform()) if i % 3 == 0 else 0.0
        m.coatWeight = float(rng.uniform()) if i % 4 == 1 else 0.0
        m.coatRoughness = float(rng.uniform(0, 0.6))
        m.coatIor = float(rng.uniform(1.2, 2.0))
        m.coatColor = tuple(rng.uniform(0.3, 1.0, 3))
        m.coatAffectColor = float(rng.uniform()) if i % 5 == 1 else 0.0
        m.coatAffectRoughness = float(rng.uniform()) if i % 5 == 2 else 0.0
        m.specularAnisotropy = float(rng.uniform(-1, 1)) if i % 5 == 3 else 0.0
        m.specularRotation = float(rng.uniform())
        m.coatAnisotropy = float(rng.uniform(-1, 1)) if i % 7 == 3 else 0.0
        m.coatRotation = float(rng.uniform())
        m.sheen = float(rng.uniform()) if i % 6 == 3 else 0.0
        m.sheenColor = tuple(rng.uniform(0, 1, 3))
        m.sheenRoughness = float(rng.uniform())
        m.thinFilmThickness = float(rng.uniform(100, 1500)) if i % 7 == 4 else 0.0
        m.thinFilmIor = float(rng.uniform(1.0, 2.0))
        m.subsurface = float(rng.uniform()) if i % 8 == 5 else 0.0
        m.subsurfaceRadius = tuple(rng.uniform(0, 4, 3))
        m.subsurfaceAnisotropy = float(rng.uniform(-1, 1))
        m.transmission = float(rng.uniform()) if i % 3 == 2 else 0.0
        m.transmissionColor = tuple(rng.uniform(0, 1, 3))
        if i % 9 == 6:
            m.hasTangentOverride = True
            m.tangentOverride = tuple(rng.normal(size=3))
        mats.append(m)
    return mats


def _unit(rng, k):
    v = rng.normal(size=(k, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def bsdf_tuples(nmat, n=4000):
    rng = np.random.default_rng(12)
    N = _unit(rng, n)
    CN = N.copy()
    tilt = rng.random(n) < 0.25
    CN[tilt] = _unit(rng, int(tilt.sum()))
    CN[tilt] = np.where((np.sum(CN[tilt] * N[tilt], 1) < 0)[:, None], -CN[tilt], CN[tilt])
    V = _unit(rng, n)
    V = np.where((np.sum(V * N, 1) < 0)[:, None], -V, V).astype(np.float32)
    L = _unit(rng, n)
    mi = rng.integers(0, nmat, n).astype(np.uint32)
    return mi, N, CN, V, L


def scalar_cases(n=500):
    rng = np.random.default_rng(13)
    nh = rng.uniform(0, 1, n).astype(np.float32)
    rough = rng.uniform(0, 1, n).astype(np.float32)
    cosi = rng.uniform(-1, 1, n).astype(np.float32)
    etai = rng.choice([1.0, 1.33, 1.5, 2.4], n).astype(np.float32)
    etat = rng.choice([1.0, 1.33, 1.5, 2.4], n).astype(np.float32)
    return nh, rough, cosi, etai, etat


def env_queries(n=4000):
    rng = np.random.default_rng(14)
    u = rng.random((n, 2)).astype(np.float32)
    u[:8] = [[0, 0], [1, 1], [0.5, 0.5], [0.99999994, 0.99999994], [0, 1], [1, 0], [1e-9, 1e-9], [0.25, 0.75]]
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[:6] = [[0, 1, 0], [0, -1, 0], [1, 0, 0], [0, 0, 1], [0, 0, -1], [-1, 0, 0]]
    return u, d


def angular_dome():
    """A 48x48 angular (light-probe) dome whose basis is turned 0.6 rad about
    +y (WorldToLight/LightToWorld, CpuPathTracer.cpp:315-325)."""
    from shiro_amd import abi
    from shiro_amd.scene import DomeLight
    from shiro_amd.workloads import procedural_dome
    c, s = float(np.cos(0.6)), float(np.sin(0.6))
    px = procedural_dome(48, 48, 777, hot_spots=2, peak=200.0, spot_width_deg=10.0)
    return DomeLight(radiance=(1.0, 0.9, 0.8), pixels=px, layout=abi.ENV_ANGULAR, right=(c, 0.0, -s),
                     up=(0.0, 1.0, 0.0), forward=(s, 0.0, c))


def angular_workload(scale=40 / 1920, spp=2):
    """c2's geometry and materials under the angular dome."""
    wl = workloads.c2(scale=scale, spp=spp, env_size=(16, 8))
    wl.scene.domeLights[0] = angular_dome()
    wl.name = "angular"
    return wl


def image_workloads():
    """Small c1/c2/c3 scenes.  Roulette starts at depth 2 in all of them: the
    reference hardcodes that depth (CpuPathTracer.cpp:2006)."""
    c3 = workloads.c3(scale=64 / 3840, spp=8, objects=6, sphere_freq=6, box_n=4, env_size=(64, 32))
    c3.settings.russianRouletteDepth = 2
    return {
        "c1": workloads.c1(scale=48 / 512, spp=4),
        "c2": workloads.c2(scale=40 / 1920, spp=2, env_size=(64, 32)),
        "c3": c3,
        "emissive": workloads.emissive(scale=48 / 512, spp=4),
        "angular": angular_workload(),
    }


def accumulated_workload():
    wl = workloads.c1(scale=32 / 512, spp=4)
    wl.settings.threadLimit = 3
    return "c1_accumulated", wl, [2, 2]
