"""The HIP backend on the TracePath branches of the feature scenes
(tests/golden/golden_cases.py:feature_workloads): opacity pass-through,
smooth glass and mirrors, thin-walled panes, absorbing media and normal
overrides (CpuPathTracer.cpp:970-986, 1405-1441, 1695-1720, 1768-1877).

Bars: beauty relMSE <= 1e-3 against the reference's own frame (the fixture)
and against the oracle at equal spp and seeds; alpha (hit fraction) equal on
>= 99.9% of pixels; both transports, both tree widths and the shading-class
order render the same frame bit for bit.
"""
import os

import numpy as np
import pytest

from golden import golden_cases as cases
from helpers import relmse
from oracle.bindings import OracleScene
from shiro_amd import abi
from shiro_amd.scene import RenderSettings

pytestmark = pytest.mark.gpu

FEATURES = ["opacity", "delta", "thin", "medium", "overrides"]
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _log(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", FEATURES)
def test_feature_frame_vs_reference_fixture(hip_ctx, name):
    wl = cases.feature_workloads()[name]
    z = np.load(os.path.join(G, f"img_{name}.npz"), allow_pickle=False)
    _log(f"feature {name} {wl.settings.width}x{wl.settings.height} {wl.settings.samplesPerPixel}spp")
    hip_ctx.upload(wl.scene)
    for transport in (abi.TRANSPORT_MEGAKERNEL, abi.TRANSPORT_WAVEFRONT):
        st = RenderSettings(**{**wl.settings.__dict__, "transport": transport})
        g = hip_ctx.render(wl.camera, st)
        assert np.all(np.isfinite(g.beauty))
        e = relmse(g.beauty[:, :3], z["beauty"][:, :3])
        _log(f"  transport {transport}: relMSE vs reference = {e:.3e}")
        assert e <= 1e-3, (transport, e)
        assert np.mean(g.beauty[:, 3] == z["beauty"][:, 3]) >= 0.999
        assert np.mean(np.all(np.isclose(g.albedo, z["albedo"], atol=1e-6), axis=1)) >= 0.999
        assert np.mean(np.all(np.isclose(g.normal, z["normal"], atol=1e-4), axis=1)) >= 0.999


@pytest.mark.parametrize("name", FEATURES)
def test_feature_frame_vs_oracle_larger(hip_ctx, name):
    """The same scenes at 96x72 and 16 spp against the oracle."""
    wl = cases.feature_workloads()[name]
    st = RenderSettings(**{**wl.settings.__dict__, "width": 96, "height": 72, "samplesPerPixel": 16})
    _log(f"feature {name} 96x72 16spp vs oracle")
    hip_ctx.upload(wl.scene)
    g = hip_ctx.render(wl.camera, st)
    o = OracleScene(wl.scene, bvh=True).render(wl.camera, st, threads=8)
    e = relmse(g.beauty[:, :3], o.beauty[:, :3])
    _log(f"  relMSE vs oracle = {e:.3e}")
    assert e <= 1e-3, e
    assert np.mean(g.beauty[:, 3] == o.beauty[:, 3]) >= 0.999
    fin = np.isfinite(o.depth) & np.isfinite(g.depth)
    assert np.mean(np.isfinite(g.depth) == np.isfinite(o.depth)) >= 0.999
    assert np.allclose(g.depth[fin], o.depth[fin], rtol=1e-4)


@pytest.mark.parametrize("name", FEATURES)
def test_feature_transports_identical(hip_ctx, name):
    wl = cases.feature_workloads()[name]
    st = RenderSettings(**{**wl.settings.__dict__, "width": 96, "height": 72, "samplesPerPixel": 16})
    _log(f"feature {name}: megakernel vs wavefront variants")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, st)
    mp = hip_ctx.render(wl.camera, RenderSettings(**{**st.__dict__, "transport": abi.TRANSPORT_MEGAKERNEL_PLUS}))
    for k in ("beauty", "albedo", "normal", "depth"):
        assert np.array_equal(getattr(mp, k), getattr(mk, k)), ("megakernel+", k)
    wst = RenderSettings(**{**st.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    saved = hip_ctx.get_tuning()
    try:
        for width in (4, 8):
            for order, tail in ((0, 0), (1, 0), (1, 1), (1, 2)):
                hip_ctx.set_tuning(bvhWidth=width, shadeOrder=order, pathsPerPass=1 << 16, tailDepth=tail)
                wf = hip_ctx.render(wl.camera, wst)
                for k in ("beauty", "albedo", "normal", "depth"):
                    a, b = getattr(wf, k), getattr(mk, k)
                    assert np.array_equal(a, b), (width, order, tail, k, np.max(np.abs(a - b)))
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})


@pytest.mark.parametrize("depth", [0, 1, 2])
def test_short_depth_limits(hip_ctx, depth):
    """maxDepth 0 renders the empty path (SampleResult's defaults: radiance,
    albedo and normal 0, depth +inf, so alpha 0: CpuPathTracer.cpp:108-113,
    2119-2130) in both transports; depths 1 and 2 agree with the oracle and
    between the transports bit for bit."""
    wl = cases.feature_workloads()["delta"]
    st = RenderSettings(**{**wl.settings.__dict__, "maxDepth": depth})
    _log(f"maxDepth {depth}")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, st)
    wf = hip_ctx.render(wl.camera, RenderSettings(**{**st.__dict__, "transport": abi.TRANSPORT_WAVEFRONT}))
    for k in ("beauty", "albedo", "normal", "depth"):
        assert np.array_equal(getattr(wf, k), getattr(mk, k)), k
    o = OracleScene(wl.scene, bvh=True).render(wl.camera, st)
    if depth == 0:
        assert not mk.beauty.any() and not mk.albedo.any() and not mk.normal.any()
        assert np.all(np.isinf(mk.depth))
        assert not o.beauty.any() and np.all(np.isinf(o.depth))
    else:
        assert relmse(mk.beauty[:, :3], o.beauty[:, :3]) <= 1e-3
        assert np.mean(mk.beauty[:, 3] == o.beauty[:, 3]) >= 0.999


def test_many_lights_vs_oracle(hip_ctx):
    """Twelve directional lights and six dome lights (three with maps, one
    angular; more than the first build held in kernel arguments) against
    the oracle, both transports and megakernel+ identical."""
    from shiro_amd.scene import DirectionalLight, DomeLight
    from shiro_amd.workloads import procedural_dome
    wl = cases.feature_workloads()["overrides"]
    sc = wl.scene
    rng = np.random.default_rng(17)
    sc.distantLights = []
    for k in range(12):
        d = rng.normal(size=3)
        d[1] = -abs(d[1]) - 0.2
        sc.distantLights.append(DirectionalLight(tuple(d / np.linalg.norm(d)), tuple(rng.uniform(0.1, 0.5, 3))))
    sc.domeLights = [DomeLight(radiance=(0.3, 0.3, 0.3), pixels=procedural_dome(32, 16, 60 + k, hot_spots=1,
                                                                                 peak=50.0, spot_width_deg=10.0))
                     for k in range(2)]
    sc.domeLights += [DomeLight(radiance=(0.05, 0.05, 0.08)) for _ in range(3)]
    sc.domeLights.append(cases.angular_dome())
    st = RenderSettings(**{**wl.settings.__dict__, "domeLightSamples": 2})
    _log(f"{len(sc.distantLights)} directional + {len(sc.domeLights)} dome lights")
    hip_ctx.upload(sc)
    mk = hip_ctx.render(wl.camera, st)
    o = OracleScene(sc, bvh=True).render(wl.camera, st)
    e = relmse(mk.beauty[:, :3], o.beauty[:, :3])
    _log(f"  relMSE vs oracle {e:.3e}")
    assert e <= 1e-3
    for t in (abi.TRANSPORT_WAVEFRONT, abi.TRANSPORT_MEGAKERNEL_PLUS):
        g = hip_ctx.render(wl.camera, RenderSettings(**{**st.__dict__, "transport": t}))
        assert np.array_equal(g.beauty, mk.beauty), t


def test_plain_and_general_bsdf_kernels(hip_ctx):
    """The wavefront shades scenes whose materials need no anisotropy, thin
    film, sheen or subsurface wrap with BSDF code that leaves those terms out
    (scene_info().plainBsdf = 1), other scenes with the general code.  Each
    optional term on its own makes a scene general; both forms render what
    the megakernel (always the general code) renders, bit for bit, and the
    general scenes agree with the oracle at relMSE <= 1e-3."""
    import copy
    from shiro_amd.scene import standard_surface
    base = [standard_surface(base_color=(0.7, 0.3, 0.2), specular_roughness=0.4),
            standard_surface(base_color=(0.8, 0.8, 0.8), metalness=1.0, specular_roughness=0.25),
            standard_surface(base_color=(0.2, 0.4, 0.7), coat=1.0, coat_roughness=0.1, specular_roughness=0.5)]
    variants = {"plain": {}, "anisotropy": {"specularAnisotropy": 0.6, "specularRotation": 0.2},
                "coat_anisotropy": {"coatAnisotropy": -0.5}, "thin_film": {"thinFilmThickness": 400.0},
                "sheen": {"sheen": 0.8, "sheenRoughness": 0.3}, "subsurface": {"subsurface": 0.5}}
    for name, change in variants.items():
        mats = [copy.copy(m) for m in base]
        for k, v in change.items():
            setattr(mats[1], k, v)
        wl = cases._feature_scene(mats, 23)
        hip_ctx.upload(wl.scene)
        plain = hip_ctx.scene_info().plainBsdf
        _log(f"{name}: plainBsdf {plain}")
        assert plain == (1 if name == "plain" else 0), name
        st = RenderSettings(**{**wl.settings.__dict__, "width": 64, "height": 48, "samplesPerPixel": 8})
        mk = hip_ctx.render(wl.camera, st)
        wf = hip_ctx.render(wl.camera, RenderSettings(**{**st.__dict__, "transport": abi.TRANSPORT_WAVEFRONT}))
        for k in ("beauty", "albedo", "normal", "depth"):
            assert np.array_equal(getattr(wf, k), getattr(mk, k)), (name, k)
        if name != "plain":
            o = OracleScene(wl.scene, bvh=True).render(wl.camera, st, threads=8)
            e = relmse(mk.beauty[:, :3], o.beauty[:, :3])
            _log(f"  relMSE vs oracle {e:.3e}")
            assert e <= 1e-3, (name, e)
