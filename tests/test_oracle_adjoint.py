"""Pins the CPU oracle's texture extension and adjoint (no reference
counterpart) with what the mathematics fixes:
  - bsdf_grad equals central differences of bsdf_eval (base colour per
    channel, roughness) over random tuples of the differentiable class;
  - a texture holding a material's own values renders the untextured frame
    bit for bit (lookups, UV interpolation and per-hit derivation add no
    rounding);
  - bilinear lookups reproduce a linear texel ramp exactly at texel centres;
  - the path-replay adjoint equals central differences of the frozen-
    sampling render (EVAL_ONLY texels) on texels the image depends on,
    through multi-bounce paths, clamps and both textures.
A retyped formula would not fail these; a dropped term, a wrong sign or a
transposed index does."""
import numpy as np
import pytest

from helpers import c5_mini, c5_scene_with
from oracle.bindings import OracleScene, bsdf_grad, eval_bsdf, oracle_fns
from shiro_amd.scene import PbrMaterial, Texture


def _class_materials(rng, n):
    mats = []
    for i in range(n):
        m = PbrMaterial()
        m.baseColor = tuple(rng.uniform(0.05, 0.95, 3))
        m.roughness = float(rng.uniform(0.2, 0.9))
        m.metallic = float(rng.choice([0.0, 0.0, 1.0, rng.uniform(0, 1)]))
        m.specularWeight = float(rng.uniform(0.5, 1.0))
        m.ior = float(rng.uniform(1.3, 1.8))
        m.diffuseRoughness = float(rng.choice([0.0, rng.uniform(0, 1)]))
        m.coatWeight = float(rng.choice([0.0, 0.0, rng.uniform(0, 1)]))
        m.coatRoughness = float(rng.uniform(0.05, 0.4))
        m.coatAffectRoughness = float(rng.choice([0.0, 0.5]))
        m.subsurface = float(rng.choice([0.0, rng.uniform(0, 0.5)]))
        mats.append(m)
    return mats


def _tuples(rng, n):
    def unit(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)
    nn = unit(n)
    v = unit(n)
    l = unit(n)
    v = np.where((np.sum(v * nn, 1) < 0.05)[:, None], v + 2 * nn, v)
    l = np.where((np.sum(l * nn, 1) < 0.05)[:, None], l + 2 * nn, l)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    l /= np.linalg.norm(l, axis=1, keepdims=True)
    return nn.astype(np.float32), v.astype(np.float32), l.astype(np.float32)


def test_bsdf_grad_matches_central_differences():
    rng = np.random.default_rng(7)
    mats = _class_materials(rng, 24)
    n, v, l = _tuples(rng, 400)
    mi = rng.integers(0, len(mats), len(n)).astype(np.uint32)
    f = oracle_fns()
    db, dr = bsdf_grad(mats, mi, n, n, v, l)

    def eval_with(mod):
        ms = [mod(m) for m in mats]
        val, _ = eval_bsdf(f, ms, mi, n, n, v, l)
        return val.astype(np.float64)

    import copy
    errs_b, errs_r = [], []
    for c in range(3):
        h = 1e-2  # f is affine in the base colour away from clamps
        def shift(m, d, c=c):
            m2 = copy.copy(m)
            bc = list(m2.baseColor)
            bc[c] += d
            m2.baseColor = tuple(bc)
            return m2
        fd = (eval_with(lambda m: shift(m, h)) - eval_with(lambda m: shift(m, -h)))[:, c] / (2 * h)
        errs_b.append(np.abs(fd - db[:, c]) / (np.abs(fd) + 1e-2))
    h = 1e-3
    def rshift(m, d):
        m2 = copy.copy(m)
        m2.roughness = m2.roughness + d
        return m2
    fd = (eval_with(lambda m: rshift(m, h)) - eval_with(lambda m: rshift(m, -h))) / (2 * h)
    errs_r.append(np.abs(fd - dr) / (np.abs(fd) + 1e-1))
    eb = np.concatenate(errs_b)
    er = np.concatenate(errs_r).ravel()
    # Central differences in fp32: a few 1e-4 of noise, far below a wrong term.
    assert np.percentile(eb, 99) < 2e-3 and eb.max() < 2e-2, (np.percentile(eb, 99), eb.max())
    assert np.percentile(er, 99) < 5e-3 and er.max() < 5e-2, (np.percentile(er, 99), er.max())
    # The roughness derivative is not trivially zero.
    assert np.mean(np.abs(dr) > 1e-3) > 0.5


def test_constant_texture_renders_the_untextured_frame():
    wl = c5_mini(size=24, spp=2)
    m = wl.scene.materials[0]
    albedo = np.broadcast_to(np.array(m.baseColor, np.float32), (16, 16, 3)).copy()
    rough = np.full((16, 16, 1), m.roughness, np.float32)
    tex_scene = c5_scene_with(wl, albedo, rough)
    plain = c5_scene_with(wl, albedo, rough)
    import copy
    pm = copy.copy(m)
    pm.baseColorTexture = pm.roughnessTexture = -1
    plain.materials = [pm] + list(wl.scene.materials[1:])
    a = OracleScene(tex_scene).render(wl.camera, wl.settings, threads=4)
    b = OracleScene(plain).render(wl.camera, wl.settings, threads=4)
    assert np.array_equal(a.beauty, b.beauty) and np.array_equal(a.albedo, b.albedo)


def test_bilinear_lookup_reproduces_a_linear_ramp():
    # Albedo texture = linear ramp in u; a camera-facing quad spanning u in
    # [0.5/w, 1 - 0.5/w] shows the ramp: the depth-0 albedo AOV (base colour
    # times the diffuse factor 1 for metallic 0, base weight 1) follows it.
    from shiro_amd.scene import Camera, RenderSettings, Scene, TriangleMesh, usd_preview_surface
    w = 8
    ramp = np.zeros((1, w, 3), np.float32)
    ramp[0, :, 0] = np.arange(w) / (w - 1)
    ramp[0, :, 1] = 0.25
    ramp[0, :, 2] = 0.5
    u0, u1 = 0.5 / w, 1 - 0.5 / w
    pos = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32)
    uv = np.array([[u0, 0.5], [u1, 0.5], [u1, 0.5], [u0, 0.5]], np.float32)
    mat = usd_preview_surface((0.5, 0.5, 0.5))
    mat.baseColorTexture = 0
    scene = Scene(materials=[mat], meshes=[TriangleMesh(pos, np.zeros((0, 3), np.float32),
                                                        np.array([0, 1, 2, 0, 2, 3], np.uint32), 0, uv)],
                  textures=[Texture(ramp)])
    cam = Camera.look_at((0, 0, 3), (0, 0, 0), fov=30.0, aspect=1.0)
    st = RenderSettings(width=16, height=16, samplesPerPixel=1, maxDepth=1)
    fb = OracleScene(scene).render(cam, st, threads=1)
    alb = fb.albedo.reshape(16, 16, 3)
    hit = fb.depth.reshape(16, 16) < np.inf
    xs = np.nonzero(hit[8])[0]
    red = alb[8, xs, 0]
    # Monotone ramp in screen x, linear in u: constant steps, G/B constant.
    assert np.all(np.diff(red) > 0)
    steps = np.diff(red)
    assert np.ptp(steps) < 1e-4 * max(1.0, steps.mean()) + 2e-6
    assert np.allclose(alb[8, xs, 1], 0.25 * 1.0, atol=1e-6) and np.allclose(alb[8, xs, 2], 0.5, atol=1e-6)


@pytest.mark.parametrize("which", [0, 1])
def test_adjoint_matches_frozen_sampling_differences(which):
    wl = c5_mini(size=32, spp=4)
    st = wl.settings
    rng = np.random.default_rng(11)
    alb, rough = wl.textures_truth
    scene = c5_scene_with(wl, alb, rough)
    osc = OracleScene(scene)
    weights = rng.normal(size=(st.height, st.width, 4)).astype(np.float32)
    weights[..., 3] = 0.0
    grads = osc.render_adjoint(wl.camera, st, weights, wanted=[0, 1], threads=8)
    g = grads[which]
    base = [alb, rough][which]
    flat = np.abs(g).ravel()
    order = np.argsort(-flat)
    picks = np.concatenate([order[:12], rng.choice(order[:200], 8, replace=False)])
    # Richardson-extrapolated central differences, (4 D(h/2) - D(h)) / 3:
    # fourth-order accurate, with a step large enough that the fp32 rounding
    # of individual path contributions (up to the direct clamp of 24) does
    # not dominate the difference.
    h = 1e-2

    def loss_at(k, d):
        t = base.copy().ravel()
        t[k] += d
        osc.update_texture(which, t.reshape(base.shape), eval_only=True)
        img = osc.render(wl.camera, st, threads=8).beauty.reshape(st.height, st.width, 4)
        return float(np.sum(img[..., :3].astype(np.float64) * weights[..., :3]))

    ad, fd = [], []
    for k in picks:
        d1 = (loss_at(k, h) - loss_at(k, -h)) / (2 * h)
        d2 = (loss_at(k, h / 2) - loss_at(k, -h / 2)) / h
        osc.update_texture(which, base, eval_only=False)
        fd.append((4 * d2 - d1) / 3)
        ad.append(float(g.ravel()[k]))
    ad, fd = np.array(ad), np.array(fd)
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    assert np.linalg.norm(fd) > 0
    print(f"adjoint vs differences: relative L2 error {err:.2e} over {len(picks)} texels")
    assert err < 1e-3, (err, ad[:6], fd[:6])


def test_env_adjoint_matches_frozen_sampling_differences():
    """Dome-map texels: the oracle's adjoint against Richardson-extrapolated
    central differences of the render whose importance tables stay those of
    the unperturbed map (eval-only updates).  The render is linear in the map
    except where the direct-light clamp binds, so a step proportional to the
    texel's magnitude keeps the differences accurate."""
    wl = c5_mini(size=32, spp=4)
    st = wl.settings
    rng = np.random.default_rng(12)
    alb, rough = wl.textures_truth
    scene = c5_scene_with(wl, alb, rough)
    base = np.array(scene.domeLights[0].pixels, np.float32)
    osc = OracleScene(scene)
    weights = rng.normal(size=(st.height, st.width, 4)).astype(np.float32)
    weights[..., 3] = 0.0
    grads = osc.render_adjoint(wl.camera, st, weights, wanted=[], env_wanted=[0], threads=8)
    g = grads[("env", 0)]
    assert g.shape == base.shape
    flat = np.abs(g).ravel()
    order = np.argsort(-flat)
    picks = np.concatenate([order[:12], rng.choice(order[:300], 8, replace=False)])

    def loss_at(k, d):
        t = base.copy().ravel()
        t[k] += d
        osc.update_environment(0, t.reshape(base.shape), eval_only=True)
        img = osc.render(wl.camera, st, threads=8).beauty.reshape(st.height, st.width, 4)
        return float(np.sum(img[..., :3].astype(np.float64) * weights[..., :3]))

    ad, fd = [], []
    for k in picks:
        h = 1e-2 * max(1.0, abs(float(base.ravel()[k])))
        d1 = (loss_at(k, h) - loss_at(k, -h)) / (2 * h)
        d2 = (loss_at(k, h / 2) - loss_at(k, -h / 2)) / h
        fd.append((4 * d2 - d1) / 3)
        ad.append(float(g.ravel()[k]))
    osc.update_environment(0, base, eval_only=False)
    ad, fd = np.array(ad), np.array(fd)
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    assert np.linalg.norm(fd) > 0
    print(f"env adjoint vs differences: relative L2 error {err:.2e} over {len(picks)} texels")
    assert err < 1e-3, (err, ad[:6], fd[:6])


def test_env_update_rebuilds_tables_like_upload():
    """An environment update without eval_only samples exactly as the map
    uploaded with the scene (same tables): frames bit-identical."""
    wl = c5_mini(size=24, spp=2)
    st = wl.settings
    alb, rough = wl.textures_truth
    scene = c5_scene_with(wl, alb, rough)
    px = np.array(scene.domeLights[0].pixels, np.float32)
    other = np.ascontiguousarray(px[::-1, ::-1] * 1.5)
    a = OracleScene(scene)
    a.update_environment(0, other)
    fa = a.render(wl.camera, st, threads=8)
    import copy
    from shiro_amd.scene import DomeLight
    s2 = copy.copy(scene)
    d = scene.domeLights[0]
    s2.domeLights = [DomeLight(radiance=d.radiance, pixels=other, layout=d.layout, right=d.right, up=d.up,
                               forward=d.forward)]
    fb = OracleScene(s2).render(wl.camera, st, threads=8)
    assert np.array_equal(fa.beauty, fb.beauty)
