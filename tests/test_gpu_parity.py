"""Parity of the HIP backend (through the C ABI) with the CPU oracle.

Bars (tolerances written in each test):
  - QMC samples: bit-exact.
  - Closest hit / primary visibility: triangle ids agree on >= 99.99% of rays
    (the rest are ties at shared edges, where the watertight test and the
    oracle's Moller-Trumbore test may pick different neighbours).
  - BSDF / environment: fp32, relative 2e-5 (the device's sin/cos/atan2/acos
    differ from glibc by ulps).
  - Images: beauty relMSE <= 1e-3 against the oracle at equal spp and seeds.
"""
import ctypes as C
import os

import numpy as np
import pytest

from helpers import angular_mini, c1_mini, c2_mini, c3_mini, c5_mini, emissive_mini, random_rays, relmse
from oracle.bindings import Accumulator, OracleScene, oracle_fns
from shiro_amd.scene import Camera, DirectionalLight, PbrMaterial, RenderSettings, Scene, TriangleMesh

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def test_qmc_bit_exact(hip_ctx):
    f = oracle_fns()
    for s in (0, 1, 7, 1000, 65535, 1 << 20):
        _log(f"qmc sample {s}")
        g = hip_ctx.qmc_fill(3, 5, 32, 16, s, 0, 96)
        ref = np.zeros(32 * 16 * 96, np.float32)
        from shiro_amd import abi
        f.qmc_fill(3, 5, 32, 16, s, 0, 96, abi.fp(ref))
        assert np.array_equal(g.reshape(-1).view(np.uint32), ref.view(np.uint32)), f"sample {s}"


@pytest.fixture(scope="module")
def c2m():
    wl = c2_mini()
    return wl, OracleScene(wl.scene, bvh=True)


def test_intersect_vs_oracle(hip_ctx, c2m):
    wl, orc = c2m
    _log("intersect c2-mini")
    hip_ctx.upload(wl.scene)
    rays = random_rays(wl.scene, 100000, seed=1)
    gp, gt, _ = hip_ctx.intersect(rays)
    op, ot, _ = orc.intersect(rays)
    agree = np.mean(gp == op)
    assert agree >= 0.9999, agree
    hit = (gp == op) & (op != 0xFFFFFFFF)
    assert np.allclose(gt[hit], ot[hit], rtol=1e-4, atol=1e-5)
    occ_rays = rays.copy()
    occ_rays[:, 7] = np.random.default_rng(2).uniform(0.1, 60.0, len(rays))
    go = hip_ctx.occluded(occ_rays)
    oo = orc.occluded(occ_rays)
    assert np.mean(go == oo) >= 0.9999


def test_primary_visibility(hip_ctx, c2m):
    wl, orc = c2m
    _log("primary visibility c2 480x270")
    hip_ctx.upload(wl.scene)
    st = RenderSettings(**{**wl.settings.__dict__, "width": 480, "height": 270})
    cam = Camera(**{**wl.camera.__dict__, "aspectRatio": 480 / 270})
    for s in (0, 3):
        gp, gt = hip_ctx.trace_primary(cam, st, s)
        op, ot = orc.trace_primary(cam, st, s)
        assert np.mean(gp == op) >= 0.9999


def _bsdf_materials():
    rng = np.random.default_rng(5)
    mats = []
    for i in range(24):
        m = PbrMaterial()
        m.baseColor = tuple(rng.uniform(0, 1, 3))
        m.metallic = float(rng.choice([0.0, 1.0, rng.uniform()]))
        m.roughness = float(rng.uniform(0.0, 1.0))
        m.diffuseRoughness = float(rng.uniform(0, 1)) if i % 3 == 0 else 0.0
        m.coatWeight = float(rng.uniform()) if i % 4 == 1 else 0.0
        m.coatRoughness = float(rng.uniform(0, 0.5))
        m.specularAnisotropy = float(rng.uniform(-1, 1)) if i % 5 == 2 else 0.0
        m.specularRotation = float(rng.uniform())
        m.sheen = float(rng.uniform()) if i % 6 == 3 else 0.0
        m.thinFilmThickness = float(rng.uniform(100, 900)) if i % 7 == 4 else 0.0
        m.subsurface = float(rng.uniform()) if i % 8 == 5 else 0.0
        m.transmission = float(rng.uniform()) if i % 3 == 2 else 0.0
        mats.append(m)
    return mats


def test_bsdf_vs_oracle(hip_ctx):
    from oracle.bindings import eval_bsdf
    mats = _bsdf_materials()
    quad = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    sc = Scene(materials=mats, meshes=[TriangleMesh(quad, np.zeros((0, 3), np.float32), np.array([0, 1, 2], np.uint32))])
    hip_ctx.upload(sc)
    rng = np.random.default_rng(6)
    n = 20000

    def unit(k):
        v = rng.normal(size=(k, 3))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    N = unit(n)
    CN = N.copy()
    V = unit(n)
    V = np.where((np.sum(V * N, 1) < 0)[:, None], -V, V).astype(np.float32)
    L = unit(n)
    mi = rng.integers(0, len(mats), n).astype(np.uint32)
    gv, gp = hip_ctx.eval_bsdf(mi, N, CN, V, L)
    ov, op = eval_bsdf(oracle_fns(), mats, mi, N, CN, V, L)
    assert np.allclose(gv, ov, rtol=2e-5, atol=1e-6), np.max(np.abs(gv - ov))
    assert np.allclose(gp, op, rtol=2e-5, atol=1e-6), np.max(np.abs(gp - op))


def test_env_vs_oracle(hip_ctx, c2m):
    wl, orc = c2m
    hip_ctx.upload(wl.scene)
    rng = np.random.default_rng(7)
    u = rng.random((50000, 2)).astype(np.float32)
    gd, gp, gc = hip_ctx.env_sample(u)
    od, op, oc = orc.env_sample(u)
    assert np.allclose(gd, od, rtol=2e-5, atol=2e-6)
    assert np.allclose(gp, op, rtol=2e-5, atol=1e-7)
    assert np.allclose(gc, oc, rtol=2e-4, atol=1e-5)
    d = rng.normal(size=(50000, 3)).astype(np.float32)
    gp2, gc2 = hip_ctx.env_eval(d)
    op2, oc2 = orc.env_eval(d)
    # A direction within an ulp of a texel boundary may fall on either side.
    assert np.mean(np.isclose(gp2, op2, rtol=2e-5, atol=1e-7)) >= 0.999
    assert np.mean(np.all(np.isclose(gc2, oc2, rtol=2e-4, atol=1e-5), axis=1)) >= 0.999


MINI = [c1_mini, c2_mini, c3_mini, emissive_mini, angular_mini]
MINI_IDS = ["c1", "c2", "c3", "emissive", "angular"]


def test_env_angular_vs_oracle(hip_ctx):
    """Angular dome: evaluation through the light-probe mapping, uniform
    sphere sampling with pdf 1/(4 pi) (no importance table)."""
    wl = angular_mini()
    _log("angular dome sample/eval")
    hip_ctx.upload(wl.scene)
    orc = OracleScene(wl.scene)
    rng = np.random.default_rng(8)
    u = rng.random((50000, 2)).astype(np.float32)
    gd, gp, gc = hip_ctx.env_sample(u)
    od, op, oc = orc.env_sample(u)
    assert np.allclose(gd, od, rtol=2e-5, atol=2e-6)
    assert np.array_equal(gp, op)
    assert np.mean(np.all(np.isclose(gc, oc, rtol=2e-4, atol=1e-5), axis=1)) >= 0.999
    d = rng.normal(size=(50000, 3)).astype(np.float32)
    gp2, gc2 = hip_ctx.env_eval(d)
    op2, oc2 = orc.env_eval(d)
    assert np.array_equal(gp2, op2)
    assert np.mean(np.all(np.isclose(gc2, oc2, rtol=2e-4, atol=1e-5), axis=1)) >= 0.999


@pytest.mark.parametrize("make", MINI, ids=MINI_IDS)
def test_render_relmse(hip_ctx, make):
    wl = make()
    _log(f"render {wl.name} {wl.settings.width}x{wl.settings.height} {wl.settings.samplesPerPixel}spp")
    hip_ctx.upload(wl.scene)
    g = hip_ctx.render(wl.camera, wl.settings)
    o = OracleScene(wl.scene, bvh=True).render(wl.camera, wl.settings)
    assert np.all(np.isfinite(g.beauty))
    e = relmse(g.beauty[:, :3], o.beauty[:, :3])
    _log(f"relMSE {wl.name} = {e:.3e}")
    assert e <= 1e-3, e
    assert np.mean(g.beauty[:, 3] == o.beauty[:, 3]) >= 0.999
    assert np.mean(np.all(np.isclose(g.albedo, o.albedo, atol=1e-6), axis=1)) >= 0.999
    assert np.mean(np.all(np.isclose(g.normal, o.normal, atol=1e-4), axis=1)) >= 0.999
    fin = np.isfinite(o.depth) & np.isfinite(g.depth)
    assert np.mean(np.isfinite(g.depth) == np.isfinite(o.depth)) >= 0.999
    assert np.allclose(g.depth[fin], o.depth[fin], rtol=1e-4)


def test_render_sliced_batches(hip_ctx):
    """A batch larger than one launch slice carries its running sums across
    launches; the result equals the oracle's single batch (compared on a
    band of rows the oracle renders in seconds)."""
    wl = c1_mini(size=512, spp=40)  # 512^2 * 40 samples span two launches
    _log("render sliced c1 512x512 40spp")
    hip_ctx.upload(wl.scene)
    g = hip_ctx.render(wl.camera, wl.settings)
    y0, rows = 200, 24
    o = OracleScene(wl.scene, bvh=True).render_region(wl.camera, wl.settings, 0, y0, 512, rows, threads=8)
    band = slice(y0 * 512, (y0 + rows) * 512)
    assert relmse(g.beauty[band, :3], o.beauty[band, :3]) <= 1e-3


def test_device_accumulator(hip_ctx):
    import torch
    wl = c1_mini(size=48, spp=4)
    hip_ctx.upload(wl.scene)
    w, h = wl.settings.width, wl.settings.height
    dev = torch.device("cuda:0")
    bufs = [torch.zeros(w * h * k, device=dev) for k in (4, 3, 3, 1)]
    out = [torch.zeros(w * h * k, device=dev) for k in (4, 3, 3, 1)]
    ptr = [b.data_ptr() for b in bufs]
    optr = [b.data_ptr() for b in out]
    acc = Accumulator(w, h)
    hip_ctx.accum_reset(w, h)
    start = 0
    for count in (1, 3, 2):
        hip_ctx.render_device(wl.camera, wl.settings, ptr, start, count)
        hip_ctx.accum_add(ptr, w, h, count)
        fb = hip_ctx.render(wl.camera, wl.settings, start, count)
        acc.add(fb, count)
        start += count
    hip_ctx.accum_resolve(optr, w, h)
    torch.cuda.synchronize()
    ref = acc.resolve()
    assert np.array_equal(out[0].cpu().numpy().reshape(-1, 4), ref.beauty)
    assert np.array_equal(out[3].cpu().numpy(), ref.depth)


def test_progressive_batches_vs_reference_fixture(hip_ctx):
    """Two progressive batches (the second with sampleStart > 0, where the
    CPU backend carries its path-guide state) accumulated on the device,
    against the frame the reference's own CPU backend wrote for the same
    scene and batches: beauty relMSE <= 1e-3."""
    import torch
    from golden import golden_cases as cases
    name, wl, batches = cases.accumulated_workload()
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"img_{name}.npz"))
    _log(f"progressive {name} batches {batches}")
    hip_ctx.upload(wl.scene)
    w, h = wl.settings.width, wl.settings.height
    dev = torch.device("cuda:0")
    bufs = [torch.zeros(w * h * k, device=dev) for k in (4, 3, 3, 1)]
    out = [torch.zeros(w * h * k, device=dev) for k in (4, 3, 3, 1)]
    hip_ctx.accum_reset(w, h)
    start = 0
    for count in batches:
        hip_ctx.render_device(wl.camera, wl.settings, [b.data_ptr() for b in bufs], start, count)
        hip_ctx.accum_add([b.data_ptr() for b in bufs], w, h, count)
        start += count
    hip_ctx.accum_resolve([b.data_ptr() for b in out], w, h)
    torch.cuda.synchronize()
    beauty = out[0].cpu().numpy().reshape(-1, 4)
    e = relmse(beauty[:, :3], z["beauty"][:, :3])
    _log(f"relMSE vs reference fixture = {e:.3e}")
    assert e <= 1e-3
    assert np.mean(beauty[:, 3] == z["beauty"][:, 3]) >= 0.999


def test_cancel_returns_cleared_frame(hip_ctx):
    wl = c1_mini(size=32, spp=2)
    hip_ctx.upload(wl.scene)
    flag = C.c_uint8(1)
    fb = hip_ctx.render(wl.camera, wl.settings, cancel=C.pointer(flag))
    assert fb.rc == 3
    assert not fb.beauty.any() and np.all(np.isinf(fb.depth))


def test_probe_scene(hip_ctx):
    """The reference's GPU smoke probe (tools/optix_probe.cpp:7-80): one
    triangle, one directional light, gradient sky, 8x8 at 1 spp; the centre
    pixel must be lit, covered and at finite depth."""
    sc = Scene(environmentTop=(0.2, 0.4, 0.8), environmentBottom=(0.05, 0.05, 0.05))
    m = PbrMaterial()
    m.baseColor = (0.8, 0.2, 0.2)
    m.roughness = 0.4
    sc.materials.append(m)
    sc.meshes.append(TriangleMesh(np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32),
                                  np.array([[0, 0, 1]] * 3, np.float32), np.array([0, 1, 2], np.uint32), 0))
    sc.distantLights.append(DirectionalLight((0.0, -1.0, -1.0), (2.0, 2.0, 2.0)))
    st = RenderSettings(width=8, height=8, samplesPerPixel=1, samplesPerUpdate=1)
    cam = Camera(aspectRatio=1.0)
    hip_ctx.upload(sc)
    fb = hip_ctx.render(cam, st, 0, 1)
    c = fb.beauty[4 * 8 + 4]
    assert c[3] > 0 and (c[0] != 0 or c[1] != 0 or c[2] != 0)
    assert fb.depth[4 * 8 + 4] < 1e10
    o = OracleScene(sc, bvh=False).render(cam, st, 0, 1)
    assert np.allclose(fb.beauty, o.beauty, rtol=1e-5, atol=1e-6)


def test_python_backend_interface(hip_ctx):
    """HipBackend mirrors RenderBackend: status, SupportsScene, RenderSampleBatch."""
    from shiro_amd.hip import HipBackend, RenderRequest
    wl = c1_mini(size=32, spp=2)
    b = HipBackend()
    st = b.GetStatus()
    assert st.available and st.usesGpu and st.name == "hip"
    assert b.SupportsScene(wl.scene)
    fb = b.RenderSampleBatch(RenderRequest(wl.scene, wl.camera, wl.settings, 0, 2))
    o = OracleScene(wl.scene, bvh=True).render(wl.camera, wl.settings, 0, 2)
    assert relmse(fb.beauty[:, :3], o.beauty[:, :3]) <= 1e-3


def test_shiro_adapter_probe_binary():
    """The C++ HipBackend adapter (built against Shiro's own headers) passes
    the reference probe's checks (tools/optix_probe.cpp:53-70)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "shiro_amd", "csrc", "adapter",
                       "_build", "hip_probe")
    if not os.path.exists(exe):
        pytest.skip("adapter probe not built (needs a Shiro source tree at build time)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    _log(r.stdout.strip())
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("probe center")


@pytest.mark.parametrize("args", [("3840", "2160", "4", "1"), ("640", "360", "3", "2"), ("64", "48", "3", "1")],
                         ids=["4k-1spp", "640x360-2spp", "64x48-megakernel"])
def test_shiro_adapter_progressive_binary(args):
    """Progressive batches through HipBackend (C++, Shiro's own types) as
    the Hydra worker drives them (RenderParam.cpp:744-836): the device
    accumulator's resolved frames equal FrameAccumulator's host arithmetic
    over the per-batch frames (Renderer.cpp:47-110), bit for bit, with and
    without a readback per batch.  4K at 1 spp per batch renders through
    the wavefront transport (the adapter picks it for batches of 2^18
    samples or more), 64x48 at 1 spp through the megakernel."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "shiro_amd", "csrc", "adapter",
                       "_build", "hip_progressive")
    if not os.path.exists(exe):
        pytest.skip("adapter not built (needs a Shiro source tree at build time)")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
    _log(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("progressive ok")


@pytest.mark.parametrize("make", MINI, ids=MINI_IDS)
def test_wavefront_equals_megakernel(hip_ctx, make):
    """Both transports evaluate the same per-bounce functions in the same
    order and accumulate samples in order, so their frames are identical."""
    from shiro_amd import abi
    wl = make()
    _log(f"wavefront vs megakernel {wl.name}")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, wl.settings)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    wf = hip_ctx.render(wl.camera, st)
    for k in ("beauty", "albedo", "normal", "depth"):
        a, b = getattr(wf, k), getattr(mk, k)
        assert np.array_equal(a, b), (k, np.max(np.abs(a - b)))


@pytest.mark.parametrize("make", MINI, ids=MINI_IDS)
def test_megakernel_plus_equals_megakernel(hip_ctx, make):
    """The megakernel with path regeneration runs the same bounce body per
    sample and sums each pixel's samples in order: identical frames."""
    from shiro_amd import abi
    wl = make()
    _log(f"megakernel+ vs megakernel {wl.name}")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, wl.settings)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_MEGAKERNEL_PLUS})
    mp = hip_ctx.render(wl.camera, st)
    for k in ("beauty", "albedo", "normal", "depth"):
        a, b = getattr(mp, k), getattr(mk, k)
        assert np.array_equal(a, b), (k, np.max(np.abs(a - b)))


def test_megakernel_plus_sliced_and_progressive(hip_ctx):
    """Slices of one batch (running sums between launches) and batches
    with sampleStart > 0 render the megakernel's frames."""
    from shiro_amd import abi
    wl = c1_mini(size=512, spp=40)  # two launch slices
    hip_ctx.upload(wl.scene)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_MEGAKERNEL_PLUS})
    for start, count in ((0, 40), (7, 3)):
        mk = hip_ctx.render(wl.camera, wl.settings, start, count)
        mp = hip_ctx.render(wl.camera, st, start, count)
        assert np.array_equal(mp.beauty, mk.beauty) and np.array_equal(mp.depth, mk.depth), (start, count)


def test_wavefront_multi_pass(hip_ctx):
    """A batch spanning several wavefront passes (running sums carried
    between passes) equals the megakernel."""
    from shiro_amd import abi
    wl = c1_mini(size=1024, spp=9)  # 1024^2 pixels -> 4 samples per pass
    _log("wavefront multi-pass c1 1024x1024 9spp")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, wl.settings)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    wf = hip_ctx.render(wl.camera, st)
    assert np.array_equal(wf.beauty, mk.beauty)
    assert np.array_equal(wf.depth, mk.depth)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_wavefront_fetch_modes_identical(hip_ctx, mode):
    """The traversal work distribution (tuning) does not change the frame."""
    from shiro_amd import abi
    wl = c3_mini(width=64, spp=4)
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, wl.settings)
    saved = hip_ctx.get_tuning()
    try:
        hip_ctx.set_tuning(fetchMode=mode, pathsPerPass=1 << 14)
        st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
        wf = hip_ctx.render(wl.camera, st)
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k in ("pathsPerPass", "fetchMode", "traceBlocks",
                                                               "stageBlocks")})
    assert np.array_equal(wf.beauty, mk.beauty)


@pytest.mark.parametrize("make", [c2_mini, c3_mini, emissive_mini], ids=["c2", "c3", "emissive"])
def test_shade_order_identical(hip_ctx, make):
    """Shading the hit queue in shading-class order permutes where paths are
    shaded, not what they compute: frames equal those without the ordering
    and those of the megakernel, bit for bit."""
    from shiro_amd import abi
    wl = make()
    _log(f"shade order {wl.name}")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, wl.settings)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    saved = hip_ctx.get_tuning()
    frames = []
    try:
        for order in (0, 1):
            hip_ctx.set_tuning(shadeOrder=order, pathsPerPass=1 << 16)
            frames.append(hip_ctx.render(wl.camera, st))
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})
    for k in ("beauty", "albedo", "normal", "depth"):
        assert np.array_equal(getattr(frames[0], k), getattr(frames[1], k)), k
        assert np.array_equal(getattr(frames[1], k), getattr(mk, k)), k


@pytest.mark.parametrize("make", [c2_mini, c3_mini, emissive_mini], ids=["c2", "c3", "emissive"])
def test_bvh_width_identical(hip_ctx, make):
    """The wavefront traversal over the 8-wide tree (same leaves, same
    watertight test, same closest-hit rule) renders the frames of the 4-wide
    megakernel bit for bit, with every fetch mode, with and without
    camera-ray packets."""
    from shiro_amd import abi
    wl = make()
    _log(f"bvh width 8 {wl.name}")
    hip_ctx.upload(wl.scene)
    info = hip_ctx.scene_info()
    assert info.bvh8NodeCount > 0 and info.bvh8Depth <= info.bvhDepth
    mk = hip_ctx.render(wl.camera, wl.settings)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    saved = hip_ctx.get_tuning()
    try:
        for mode in (0, 1, 2):
            for packets in (0, 1, 4, 16, 64):
                hip_ctx.set_tuning(bvhWidth=8, fetchMode=mode, primaryPackets=packets, pathsPerPass=1 << 16)
                wf = hip_ctx.render(wl.camera, st)
                for k in ("beauty", "albedo", "normal", "depth"):
                    assert np.array_equal(getattr(wf, k), getattr(mk, k)), (mode, packets, k)
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})


@pytest.mark.parametrize("make", [c2_mini, c3_mini, emissive_mini], ids=["c2", "c3", "emissive"])
def test_packet_sample_groups_identical(hip_ctx, make):
    """Camera-ray packets of 4 samples of a 4x4 pixel square, of 16 samples
    of a 2x2 square and of 64 samples of one pixel (and the fallback to a
    smaller group when the pass's samples per pixel are not a multiple)
    render the megakernel's frame bit for bit: each lane keeps its own ray's
    closest hit, over the 4-wide and the 8-wide tree."""
    from shiro_amd import abi
    wl = make()
    st = RenderSettings(**{**wl.settings.__dict__, "samplesPerPixel": 64})
    _log(f"packet groups {wl.name} {st.width}x{st.height} 64 spp")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, st)
    wst = RenderSettings(**{**st.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    saved = hip_ctx.get_tuning()
    try:
        # 1 << 8 paths per pass: one sample per pass, so G falls back to 1.
        cases = [(4, 1 << 26, 4), (16, 1 << 26, 4), (64, 1 << 26, 4), (64, 1 << 8, 4), (16, 1 << 26, 8),
                 (64, 1 << 26, 8)]
        for packets, paths, width in cases:
            hip_ctx.set_tuning(primaryPackets=packets, pathsPerPass=paths, bvhWidth=width)
            wf = hip_ctx.render(wl.camera, wst)
            for k in ("beauty", "albedo", "normal", "depth"):
                assert np.array_equal(getattr(wf, k), getattr(mk, k)), (packets, paths, width, k)
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})


def test_autotune_keeps_result(hip_ctx):
    from shiro_amd import abi
    wl = c2_mini(width=128, spp=4)
    hip_ctx.upload(wl.scene)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    ref = hip_ctx.render(wl.camera, st)
    saved = hip_ctx.get_tuning()
    t, report = hip_ctx.autotune(wl.camera, st, 2)
    _log(report)
    assert report.count("ms") >= 10 and t.fetchMode in (0, 1, 2)
    assert "bvhWidth 8" in report and t.bvhWidth in (4, 8)
    out = hip_ctx.render(wl.camera, st)
    hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})
    assert np.array_equal(out.beauty, ref.beauty)


@pytest.mark.parametrize("make", [c2_mini, c3_mini, emissive_mini], ids=["c2", "c3", "emissive"])
def test_traversal_knobs_identical(hip_ctx, make):
    """The leaf-wait thresholds of the closest-hit and any-hit traversal
    change when lanes test their leaves, not which triangles a ray meets:
    the frames equal the megakernel's bit for bit, over both tree widths and
    every fetch mode."""
    from shiro_amd import abi
    wl = make()
    _log(f"traversal knobs {wl.name}")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, wl.settings)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    saved = hip_ctx.get_tuning()
    try:
        for width in (4, 8):
            for mode in (0, 2):
                for wait, swait in ((0, 0), (24, 24), (64, 64), (40, 8), (8, 40)):
                    hip_ctx.set_tuning(bvhWidth=width, fetchMode=mode, leafWaitLanes=wait,
                                       shadowLeafWaitLanes=swait, pathsPerPass=1 << 16)
                    wf = hip_ctx.render(wl.camera, st)
                    for k in ("beauty", "albedo", "normal", "depth"):
                        assert np.array_equal(getattr(wf, k), getattr(mk, k)), (width, mode, wait, swait, k)
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})


@pytest.mark.parametrize("make", [c2_mini, c3_mini], ids=["c2", "c3"])
def test_wavefront_fetch_counts(hip_ctx, make):
    """The counting instantiations of the wavefront traversal kernels trace
    the rays the megakernel traces (same camera, bounce and shadow rays);
    with leaf tests as soon as a lane reaches a leaf the any-hit traversal
    visits exactly the megakernel's nodes and triangles and finds the same
    occluded rays at any leaf-wait threshold; a camera packet fetches each
    node once for its rays."""
    from shiro_amd import abi
    wl = make()
    hip_ctx.upload(wl.scene)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    spp = st.samplesPerPixel
    mkc = hip_ctx.count_fetches(wl.camera, wl.settings, 0, spp)
    saved = hip_ctx.get_tuning()
    try:
        counts = {}
        for swait in (64, 24):
            hip_ctx.set_tuning(shadowLeafWaitLanes=swait, bvhWidth=4, primaryPackets=16, fetchMode=2, tailDepth=0)
            counts[swait] = hip_ctx.count_wavefront(wl.camera, st, 0, spp)
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})
    w0, w1 = counts[64], counts[24]
    _log(f"{wl.name}: rays {list(w0.rays)} vs megakernel {mkc.primaryRays} {mkc.rays - mkc.primaryRays} "
         f"{mkc.shadowRays}; shadow nodes {w0.nodeFetches[2]} vs {mkc.shadowNodeFetches} (leaf wait 24: "
         f"{w1.nodeFetches[2]}); occluded {w0.shadowOccluded}/{w1.shadowOccluded}; camera nodes per ray "
         f"{w0.nodeFetches[0]} per packet {w0.waveNodeFetches[0]}")
    assert w0.samples == st.width * st.height * spp
    assert w0.rays[0] == mkc.primaryRays == w0.samples
    assert w0.rays[1] == mkc.rays - mkc.primaryRays
    assert w0.rays[2] == mkc.shadowRays and w1.rays[2] == w0.rays[2]
    assert w0.nodeFetches[2] == mkc.shadowNodeFetches
    assert w0.triFetches[2] == mkc.shadowTriFetches
    assert w1.shadowOccluded == w0.shadowOccluded
    for w in (w0, w1):
        for k in range(3):
            assert w.waveNodeFetches[k] <= w.nodeFetches[k] and w.waveTriFetches[k] <= w.triFetches[k]
    # Packets of 16 samples of 2x2 pixels: each node once per 64 rays.
    assert w0.waveNodeFetches[0] * 8 < w0.nodeFetches[0]


@pytest.mark.parametrize("make", [c1_mini, c2_mini, c3_mini, emissive_mini, angular_mini, c5_mini],
                         ids=["c1", "c2", "c3", "emissive", "angular", "c5-textured"])
def test_tail_kernel_identical(hip_ctx, make):
    """Finishing a pass's paths from bounce d on in one persistent kernel
    (tuning.tailDepth) runs each path's remaining bounces with the stage
    kernels' operations in their order: frames equal the megakernel's bit
    for bit, from every start depth, over both tree widths and with several
    passes per frame."""
    from shiro_amd import abi
    wl = make()
    _log(f"tail kernel {wl.name}")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, wl.settings)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    saved = hip_ctx.get_tuning()
    try:
        for tail in range(1, st.maxDepth + 1):
            for width, paths in ((4, 1 << 16), (8, 1 << 12)):
                hip_ctx.set_tuning(tailDepth=tail, bvhWidth=width, pathsPerPass=paths, fetchMode=2)
                wf = hip_ctx.render(wl.camera, st)
                for k in ("beauty", "albedo", "normal", "depth"):
                    assert np.array_equal(getattr(wf, k), getattr(mk, k)), (tail, width, k)
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})


@pytest.mark.parametrize("make", [c1_mini, c2_mini, c3_mini, emissive_mini, c5_mini],
                         ids=["c1", "c2", "c3", "emissive", "c5-textured"])
def test_shadow_overlap_identical(hip_ctx, make):
    """A bounce's shadow rays and resolve on a second stream beside the next
    bounce's closest hits (tuning.shadowOverlap) touch arrays the closest
    hits do not, and the next shading waits for both: frames equal the
    serial order's and the megakernel's bit for bit, with and without the
    tail kernel, in one and in several passes per frame, with lane refill
    from pools and with wave rounds.  (The adjoint tests run the forward
    pass with the replay log under the default, overlapped, order.)"""
    from shiro_amd import abi
    wl = make()
    _log(f"shadow overlap {wl.name}")
    hip_ctx.upload(wl.scene)
    mk = hip_ctx.render(wl.camera, wl.settings)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    saved = hip_ctx.get_tuning()
    try:
        for overlap in (0, 1):
            for tail, paths, fetch in ((0, 1 << 16, 2), (abi.TAIL_AUTO, 1 << 14, 0), (3, 1 << 20, 2)):
                hip_ctx.set_tuning(shadowOverlap=overlap, tailDepth=tail, pathsPerPass=paths, fetchMode=fetch)
                wf = hip_ctx.render(wl.camera, st)
                for k in ("beauty", "albedo", "normal", "depth"):
                    assert np.array_equal(getattr(wf, k), getattr(mk, k)), (overlap, tail, paths, k)
    finally:
        hip_ctx.set_tuning(**{k: getattr(saved, k) for k, _ in abi.Tuning._fields_})
