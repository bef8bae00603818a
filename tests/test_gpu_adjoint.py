"""Texture extension and path-replay adjoint on the device (no reference
counterpart), through the C ABI and the PyTorch wrapper.

Bars (tolerances written in each test):
  - textured frames vs the CPU oracle at equal spp: beauty relMSE <= 1e-3,
    both transports; a texture holding the material's own values renders the
    untextured frame bit for bit;
  - device adjoint vs the oracle's adjoint (same samples): relative L2 error
    <= 1e-3 over every texel (float atomics sum in another order);
  - device adjoint vs differences of the frozen-sampling render (texels that
    shading evaluates perturbed, sampling texels fixed) on 64 random texels
    of the full 256x256, 32 spp workload: relative L2 error <= 1e-3;
  - 100 Adam steps on the MSE to a target image rendered from seeded texels
    (same samples) lower the loss by a quarter or more and move the texels
    towards the target; the full 256x256, 32 spp workload with textures and
    dome map as parameters, 100 Adam steps: the loss falls below 0.3 of its
    start.
Dome-map texels as parameters, to the same bars: a map replaced on the
device renders what a fresh upload of it renders, bit for bit; the device
map gradient vs the oracle's; vs central differences on 64 random seen texel
channels of the full workload (importance tables frozen); Adam over textures
and map together lowers the loss.
"""
import numpy as np
import pytest

from helpers import c5_mini, c5_scene_with, relmse
from oracle.bindings import OracleScene
from shiro_amd import abi, workloads
from shiro_amd.scene import RenderSettings

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c5m():
    wl = c5_mini(size=48, spp=4)
    alb, rough = wl.textures_truth
    scene = c5_scene_with(wl, alb, rough)
    return wl, scene


@pytest.mark.parametrize("transport", [abi.TRANSPORT_MEGAKERNEL, abi.TRANSPORT_WAVEFRONT,
                                       abi.TRANSPORT_MEGAKERNEL_PLUS])
def test_textured_frame_vs_oracle(hip_ctx, c5m, transport):
    wl, scene = c5m
    _log(f"textured c5-mini transport {transport}")
    hip_ctx.upload(scene)
    st = wl.settings
    st.transport = transport
    g = hip_ctx.render(wl.camera, st)
    o = OracleScene(scene).render(wl.camera, st, threads=8)
    st.transport = abi.TRANSPORT_MEGAKERNEL
    gb = g.beauty.reshape(-1, 4)[:, :3]
    ob = o.beauty.reshape(-1, 4)[:, :3]
    err = relmse(gb, ob)
    assert err <= 1e-3, err
    assert relmse(g.albedo, o.albedo) <= 1e-6


def test_constant_texture_is_the_untextured_frame(hip_ctx):
    import copy
    wl = c5_mini(size=40, spp=2)
    m = wl.scene.materials[0]
    albedo = np.broadcast_to(np.array(m.baseColor, np.float32), (16, 16, 3)).copy()
    rough = np.full((16, 16, 1), m.roughness, np.float32)
    tex = c5_scene_with(wl, albedo, rough)
    plain = c5_scene_with(wl, albedo, rough)
    pm = copy.copy(m)
    pm.baseColorTexture = pm.roughnessTexture = -1
    plain.materials = [pm] + list(wl.scene.materials[1:])
    _log("constant texture vs untextured")
    hip_ctx.upload(tex)
    a = hip_ctx.render(wl.camera, wl.settings)
    hip_ctx.upload(plain)
    b = hip_ctx.render(wl.camera, wl.settings)
    assert np.array_equal(a.beauty, b.beauty) and np.array_equal(a.albedo, b.albedo)


def _device_adjoint(ctx, wl, weights, ntex):
    import torch
    st = wl.settings
    w = torch.from_numpy(np.ascontiguousarray(weights, np.float32)).cuda()
    grads = [torch.zeros(np.shape(t.texels), dtype=torch.float32, device="cuda") for t in wl_scene_textures(wl)]
    torch.cuda.synchronize()
    ctx.render_adjoint(wl.camera, st, w.data_ptr(), [g.data_ptr() for g in grads[:ntex]])
    return [g.cpu().numpy() for g in grads]


def wl_scene_textures(wl):
    return wl.scene.textures


def test_adjoint_vs_oracle_adjoint(hip_ctx, c5m):
    wl, scene = c5m
    st = wl.settings
    rng = np.random.default_rng(3)
    weights = rng.normal(size=(st.height, st.width, 4)).astype(np.float32)
    weights[..., 3] = 0.0
    _log("adjoint c5-mini vs oracle")
    hip_ctx.upload(scene)
    wl2 = type(wl)(wl.name, scene, wl.camera, st)
    dev = _device_adjoint(hip_ctx, wl2, weights, 2)
    ref = OracleScene(scene).render_adjoint(wl.camera, st, weights, wanted=[0, 1], threads=8)
    for i in (0, 1):
        err = np.linalg.norm(dev[i] - ref[i]) / np.linalg.norm(ref[i])
        _log(f"texture {i}: relative L2 {err:.2e}")
        assert np.linalg.norm(ref[i]) > 0
        assert err <= 1e-3, (i, err)


def test_adjoint_vs_frozen_sampling_differences_c5(hip_ctx):
    """Full c5: MSE loss to the target rendered from the true texels, at the
    initial texels; 64 random (texture, texel, channel) parameters among the
    texels the image sees; central differences with h = 1e-3 of renders whose
    sampling texels stay at the initial values."""
    import torch
    wl = workloads.c5()
    st = wl.settings
    alb_t, rough_t = wl.textures_truth
    alb0, rough0 = wl.textures_init
    _log("c5 target render")
    hip_ctx.upload(c5_scene_with(wl, alb_t, rough_t))
    target = hip_ctx.render(wl.camera, st).beauty.reshape(st.height, st.width, 4)[..., :3].astype(np.float64)
    scene0 = c5_scene_with(wl, alb0, rough0)
    hip_ctx.upload(scene0)
    img0 = hip_ctx.render(wl.camera, st).beauty.reshape(st.height, st.width, 4)[..., :3].astype(np.float64)
    n = img0.size

    def loss(img):
        return float(np.sum((img - target) ** 2) / n)

    dimg = np.zeros((st.height, st.width, 4), np.float32)
    dimg[..., :3] = (2.0 * (img0 - target) / n).astype(np.float32)
    _log("c5 adjoint")
    wl0 = type(wl)(wl.name, scene0, wl.camera, st)
    grads = _device_adjoint(hip_ctx, wl0, dimg, 2)
    rng = np.random.default_rng(5)
    params = []
    for i, g in enumerate(grads):
        flat = np.abs(g.ravel())
        seen = np.nonzero(flat > 1e-3 * flat.max())[0]
        for k in rng.choice(seen, 32, replace=False):
            params.append((i, int(k)))
    base = [alb0, rough0]

    def loss_at(i, k, d):
        t = base[i].copy().ravel()
        t[k] += d
        hip_ctx.update_texture(i, t.reshape(base[i].shape), eval_only=True)
        img = hip_ctx.render(wl.camera, st).beauty.reshape(st.height, st.width, 4)[..., :3]
        return loss(img.astype(np.float64))

    # Central differences with h = 1e-3.  The starting texels are not grey:
    # with grey texels the throughput channels tie and the max inside the
    # reference's clamp_max_component has one-sided derivatives only.
    h = 1e-3
    ad, fd = [], []
    for i, k in params:
        fd.append((loss_at(i, k, h) - loss_at(i, k, -h)) / (2 * h))
        hip_ctx.update_texture(i, base[i], eval_only=False)
        ad.append(float(grads[i].ravel()[k]))
    ad, fd = np.array(ad), np.array(fd)
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    _log(f"c5 adjoint vs central differences (h={h}) over {len(params)} parameters: relative L2 error {err:.2e}")
    assert err <= 1e-3, (err, ad[:5], fd[:5])


def test_replay_log_adjoint_equals_traced_adjoint():
    """The adjoint replaying the forward wavefront pass's hits and shadow-ray
    visibilities (shiro_hip_set_replay_log) against the adjoint that traces
    the rays again, on the half-size c5 workload with texture and dome-map
    gradients: relative L2 difference <= 1e-5 per array (float atomics add in
    another order).  The log is used only for the samples, camera and scene
    content of the logged pass."""
    import torch
    from shiro_amd.hip import Context
    wl = workloads.c5(scale=0.5, spp=8)
    st = RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT})
    alb0, rough0 = wl.textures_init
    scene = c5_scene_with(wl, alb0, rough0)
    ctx = Context(0)
    try:
        ctx.upload(scene)
        n = st.width * st.height
        frame = [torch.empty(n * k, dtype=torch.float32, device="cuda") for k in (4, 3, 3, 1)]
        rng = np.random.default_rng(11)
        w = torch.from_numpy(rng.normal(size=(st.height, st.width, 4)).astype(np.float32)).cuda()

        def adjoint():
            grads = [torch.zeros(np.shape(t.texels), dtype=torch.float32, device="cuda") for t in scene.textures]
            env = [torch.zeros(np.shape(scene.domeLights[0].pixels), dtype=torch.float32, device="cuda")]
            torch.cuda.synchronize()
            ctx.render_adjoint(wl.camera, st, w.data_ptr(), [g.data_ptr() for g in grads],
                               env_grad_ptrs=[e.data_ptr() for e in env])
            return [g.cpu().numpy() for g in grads + env], ctx.last_adjoint_replayed()

        def forward(start=0):
            ctx.render_device(wl.camera, st, [b.data_ptr() for b in frame], start, st.samplesPerPixel)

        _log("c5 half-size: traced adjoint, then the replayed one")
        forward()
        traced, replayed = adjoint()
        assert not replayed
        ctx.set_replay_log(1 << 30)
        forward()
        logged, replayed = adjoint()
        assert replayed
        for k, (a, b) in enumerate(zip(traced, logged)):
            err = np.linalg.norm(a - b) / np.linalg.norm(a)
            _log(f"  array {k}: relative L2 difference {err:.2e}")
            assert np.linalg.norm(a) > 0 and err <= 1e-5, (k, err)
        # Another batch of samples, then a texture update: the log no longer
        # matches and the adjoint traces.
        forward(start=st.samplesPerPixel)
        assert not adjoint()[1]
        forward()
        assert adjoint()[1]
        ctx.update_texture(0, alb0)
        assert not adjoint()[1]
        ctx.set_replay_log(0)
        forward()
        assert not adjoint()[1]
    finally:
        ctx.close()


def test_adam_through_torch_lowers_the_loss():
    import torch
    from shiro_amd.diff import TextureRenderer
    wl = workloads.c5(scale=0.5, spp=8)
    st = wl.settings
    alb_t, rough_t = wl.textures_truth
    alb0, rough0 = wl.textures_init
    _log("c5 half-size Adam")
    r = TextureRenderer(c5_scene_with(wl, alb_t, rough_t), wl.camera, st)
    with torch.no_grad():
        target = r.render(torch.from_numpy(alb_t).cuda(), torch.from_numpy(rough_t).cuda()).clone()
    albedo = torch.from_numpy(alb0.copy()).cuda().requires_grad_(True)
    rough = torch.from_numpy(rough0.copy()).cuda().requires_grad_(True)
    opt = torch.optim.Adam([albedo, rough], lr=0.02)
    losses = []
    # Target and iterates use the same samples, so the loss measures the
    # texel error rather than Monte Carlo noise between sample sets.
    for step in range(100):
        opt.zero_grad()
        img = r.render(albedo, rough)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.0, 1.0)
            rough.clamp_(0.05, 1.0)
        losses.append(loss.item())
    _log(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.75 * losses[0]
    e0 = float(np.abs(alb0 - alb_t).mean())
    e1 = float(torch.abs(albedo - torch.from_numpy(alb_t).cuda()).mean())
    _log(f"mean |base colour - target texels| {e0:.4f} -> {e1:.4f}")
    assert e1 < e0


def test_adam_full_size_c5_textures_and_dome():
    """The full c5 workload as the benchmark runs it: 256x256, 32 spp,
    base-colour and roughness textures and the dome map as parameters,
    100 Adam steps through torch.autograd.  The image loss falls below 0.3
    of its start.  The parameters themselves need not approach the targets:
    a brighter map and darker base colours render alike (the lighting and
    albedo scale ambiguity), so their errors are only reported."""
    import copy
    import torch
    from shiro_amd.diff import TextureRenderer
    from shiro_amd.scene import Texture
    wl = workloads.c5()
    st = wl.settings
    assert (st.width, st.height, st.samplesPerPixel) == (256, 256, 32)
    alb_t, rough_t = wl.textures_truth
    alb0, rough0 = wl.textures_init
    scene = copy.copy(wl.scene)
    scene.textures = [Texture(alb_t), Texture(rough_t)]
    _log("c5 full-size Adam: textures and dome map")
    r = TextureRenderer(scene, wl.camera, st, env_domes=[0])
    with torch.no_grad():
        target = r.render(torch.from_numpy(alb_t).cuda(), torch.from_numpy(rough_t).cuda(),
                          env=[torch.from_numpy(wl.env_truth).cuda()]).clone()
    albedo = torch.from_numpy(alb0.copy()).cuda().requires_grad_(True)
    rough = torch.from_numpy(rough0.copy()).cuda().requires_grad_(True)
    sky = torch.from_numpy(wl.env_init.copy()).cuda().requires_grad_(True)
    opt = torch.optim.Adam([{"params": [albedo, rough], "lr": 0.02}, {"params": [sky], "lr": 0.05}])
    losses = []
    for step in range(100):
        opt.zero_grad()
        img = r.render(albedo, rough, env=[sky])
        loss = ((img - target) ** 2).mean()
        loss.backward()
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.0, 1.0)
            rough.clamp_(0.05, 1.0)
            sky.clamp_(min=0.0)
        losses.append(loss.item())
        if step % 25 == 0:
            _log(f"  step {step}: loss {losses[-1]:.4e}")
    _log(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.3 * losses[0]
    with torch.no_grad():
        ea0 = float(np.abs(alb0 - alb_t).mean())
        ea1 = float(torch.abs(albedo - torch.from_numpy(alb_t).cuda()).mean())
        es0 = float(np.abs(wl.env_init - wl.env_truth).mean())
        es1 = float(torch.abs(sky - torch.from_numpy(wl.env_truth).cuda()).mean())
    _log(f"mean |base colour - target| {ea0:.4f} -> {ea1:.4f}; mean |map - target| {es0:.4f} -> {es1:.4f}")
    assert np.isfinite([ea1, es1]).all()


# ---------------------------------------------------------------------------
# Dome-map texels as parameters.


def _device_env_adjoint(ctx, wl, weights):
    import torch
    st = wl.settings
    w = torch.from_numpy(np.ascontiguousarray(weights, np.float32)).cuda()
    g = torch.zeros(np.shape(wl.scene.domeLights[0].pixels), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.render_adjoint(wl.camera, st, w.data_ptr(), [], env_grad_ptrs=[g.data_ptr()])
    return g.cpu().numpy()


def test_env_update_samples_like_upload(hip_ctx, c5m):
    """A dome map replaced on the device (tables rebuilt there) renders the
    frame a fresh upload of the same map renders, bit for bit, from host and
    from device texels."""
    import copy
    import torch
    from shiro_amd.hip import Context
    from shiro_amd.scene import DomeLight
    wl, scene = c5m
    st = wl.settings
    px = np.array(scene.domeLights[0].pixels, np.float32)
    other = np.ascontiguousarray(px[::-1, ::-1] * 1.7 + 0.05)
    _log("env update vs upload")
    hip_ctx.upload(scene, force=True)
    hip_ctx.update_environment(0, other)
    a = hip_ctx.render(wl.camera, st)
    dev = torch.from_numpy(other).cuda()
    hip_ctx.upload(scene, force=True)
    hip_ctx.update_environment(0, dev.data_ptr())
    torch.cuda.synchronize()
    b = hip_ctx.render(wl.camera, st)
    s2 = copy.copy(scene)
    d = scene.domeLights[0]
    s2.domeLights = [DomeLight(radiance=d.radiance, pixels=other, layout=d.layout, right=d.right, up=d.up,
                               forward=d.forward)]
    fresh = Context(0)
    try:
        fresh.upload(s2)
        c = fresh.render(wl.camera, st)
    finally:
        fresh.close()
    assert np.array_equal(a.beauty, c.beauty) and np.array_equal(b.beauty, c.beauty)
    hip_ctx.upload(scene, force=True)


def test_env_adjoint_vs_oracle_adjoint(hip_ctx, c5m):
    wl, scene = c5m
    st = wl.settings
    rng = np.random.default_rng(13)
    weights = rng.normal(size=(st.height, st.width, 4)).astype(np.float32)
    weights[..., 3] = 0.0
    _log("env adjoint c5-mini vs oracle")
    hip_ctx.upload(scene, force=True)
    dev = _device_env_adjoint(hip_ctx, wl, weights)
    ref = OracleScene(scene).render_adjoint(wl.camera, st, weights, wanted=[], env_wanted=[0], threads=8)[("env", 0)]
    err = np.linalg.norm(dev - ref) / np.linalg.norm(ref)
    _log(f"env gradient relative L2 {err:.2e}")
    assert np.linalg.norm(ref) > 0
    assert err <= 1e-3, err


def test_env_adjoint_vs_frozen_sampling_differences_c5(hip_ctx):
    """Full c5 at the starting parameters: 64 random dome-map texel channels
    among those the image sees; central differences of renders whose
    importance tables stay those of the starting map, with h = 1e-3 times
    the texel's magnitude (at least 1e-3: the map holds radiances up to 40)."""
    import copy
    from shiro_amd.scene import DomeLight
    wl = workloads.c5()
    st = wl.settings
    alb_t, rough_t = wl.textures_truth
    alb0, rough0 = wl.textures_init
    _log("c5 env target render")
    hip_ctx.upload(c5_scene_with(wl, alb_t, rough_t), force=True)
    target = hip_ctx.render(wl.camera, st).beauty.reshape(st.height, st.width, 4)[..., :3].astype(np.float64)
    scene0 = c5_scene_with(wl, alb0, rough0)
    d = scene0.domeLights[0]
    scene0.domeLights = [DomeLight(radiance=d.radiance, pixels=wl.env_init.copy(), layout=d.layout, right=d.right,
                                   up=d.up, forward=d.forward)]
    hip_ctx.upload(scene0, force=True)
    img0 = hip_ctx.render(wl.camera, st).beauty.reshape(st.height, st.width, 4)[..., :3].astype(np.float64)
    n = img0.size

    def loss(img):
        return float(np.sum((img - target) ** 2) / n)

    dimg = np.zeros((st.height, st.width, 4), np.float32)
    dimg[..., :3] = (2.0 * (img0 - target) / n).astype(np.float32)
    wl0 = copy.copy(wl)
    wl0.scene = scene0
    g = _device_env_adjoint(hip_ctx, wl0, dimg)
    rng = np.random.default_rng(6)
    flat = np.abs(g.ravel())
    seen = np.nonzero(flat > 1e-3 * flat.max())[0]
    picks = rng.choice(seen, 64, replace=False)
    base = wl.env_init.copy()

    def loss_at(k, dv):
        t = base.copy().ravel()
        t[k] += dv
        hip_ctx.update_environment(0, t.reshape(base.shape), eval_only=True)
        img = hip_ctx.render(wl.camera, st).beauty.reshape(st.height, st.width, 4)[..., :3]
        return loss(img.astype(np.float64))

    ad, fd = [], []
    for k in picks:
        h = 1e-3 * max(1.0, abs(float(base.ravel()[k])))
        fd.append((loss_at(k, h) - loss_at(k, -h)) / (2 * h))
        ad.append(float(g.ravel()[k]))
    hip_ctx.update_environment(0, base)
    ad, fd = np.array(ad), np.array(fd)
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    _log(f"c5 env adjoint vs central differences over {len(picks)} texel channels: relative L2 error {err:.2e}")
    assert err <= 1e-3, (err, ad[:5], fd[:5])


def test_adam_with_the_dome_map_lowers_the_loss():
    """100 Adam steps on base colour, roughness and the dome map together."""
    import torch
    from shiro_amd.diff import TextureRenderer
    wl = workloads.c5(scale=0.5, spp=8)
    st = wl.settings
    alb_t, rough_t = wl.textures_truth
    alb0, rough0 = wl.textures_init
    _log("c5 half-size Adam with the dome map")
    r = TextureRenderer(c5_scene_with(wl, alb_t, rough_t), wl.camera, st, env_domes=[0])
    env_t = torch.from_numpy(wl.env_truth).cuda()
    with torch.no_grad():
        target = r.render(torch.from_numpy(alb_t).cuda(), torch.from_numpy(rough_t).cuda(), env=[env_t]).clone()
    albedo = torch.from_numpy(alb0.copy()).cuda().requires_grad_(True)
    rough = torch.from_numpy(rough0.copy()).cuda().requires_grad_(True)
    sky = torch.from_numpy(wl.env_init.copy()).cuda().requires_grad_(True)
    opt = torch.optim.Adam([{"params": [albedo, rough], "lr": 0.02}, {"params": [sky], "lr": 0.05}])
    losses = []
    for step in range(100):
        opt.zero_grad()
        img = r.render(albedo, rough, env=[sky])
        loss = ((img - target) ** 2).mean()
        loss.backward()
        assert sky.grad is not None and torch.isfinite(sky.grad).all()
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.0, 1.0)
            rough.clamp_(0.05, 1.0)
            sky.clamp_(min=0.0)
        losses.append(loss.item())
    _log(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.75 * losses[0]


def test_two_forwards_then_one_backward():
    """Two render() calls at different parameters before one backward():
    each image's gradient is taken at its own parameters (the forward pass
    keeps them and the backward pass loads them again when another forward
    pass has replaced them).  The summed gradient equals that of two
    separate forward/backward pairs (relative L2 <= 1e-5: the adjoint's
    float atomics may sum in another order), while the gradient of the first
    image at the second image's parameters is far from it."""
    import torch
    from shiro_amd.diff import TextureRenderer
    wl = c5_mini(size=48, spp=4)
    alb_t, rough_t = wl.textures_truth
    alb0, rough0 = wl.textures_init
    _log("two forwards, one backward")
    r = TextureRenderer(c5_scene_with(wl, alb_t, rough_t), wl.camera, wl.settings, env_domes=[0])
    env_t = torch.from_numpy(wl.env_truth).cuda()
    with torch.no_grad():
        target = r.render(torch.from_numpy(alb_t).cuda(), torch.from_numpy(rough_t).cuda(), env=[env_t]).clone()
    pa = [torch.from_numpy(alb0.copy()).cuda(), torch.from_numpy(rough0.copy()).cuda(),
          torch.from_numpy(wl.env_init.copy()).cuda()]
    pb = [torch.clamp(pa[0] * 0.7 + 0.2, 0, 1), torch.clamp(pa[1] * 0.5 + 0.3, 0.05, 1), pa[2] * 1.5]

    def leaf(ps):
        return [p.clone().requires_grad_(True) for p in ps]

    def loss_of(ps):
        img = r.render(ps[0], ps[1], env=[ps[2]])
        return ((img - target) ** 2).mean()

    a, b = leaf(pa), leaf(pb)
    (loss_of(a) + loss_of(b)).backward()
    a1, b1 = leaf(pa), leaf(pb)
    loss_of(a1).backward()
    loss_of(b1).backward()
    # The old hazard: the first image's adjoint at the second's parameters.
    a2, b2 = leaf(pa), leaf(pb)
    la = loss_of(a2)
    v = r.version
    with torch.no_grad():
        r._load([b2[0], b2[1]], [b2[2]])
    r.version = v  # the context holds b's parameters but claims a's version
    la.backward()
    for k in range(3):
        for x, y in ((a[k].grad, a1[k].grad), (b[k].grad, b1[k].grad)):
            e = float(torch.linalg.norm(x - y) / torch.linalg.norm(y))
            assert e <= 1e-5, (k, e)
        wrong = float(torch.linalg.norm(a2[k].grad - a1[k].grad) / torch.linalg.norm(a1[k].grad))
        _log(f"parameter {k}: gradient at the other pass's parameters is off by {wrong:.2e}")
        assert wrong > 1e-2
