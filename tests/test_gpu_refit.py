"""BVH refit of a re-uploaded scene with the same meshes and indices
(shiro_hip_refit_scene), through the C ABI and the Python backend mirror.

Bars: refitting the unmoved scene reproduces the frame bit for bit (the
refit rebuilds the same boxes); after a deformation the refitted tree
renders the frame a fresh upload (new SAH tree) renders, up to the order in
which exact ties between triangles are met (>= 99.99% of pixels identical,
beauty relMSE <= 1e-9), and matches the CPU oracle at relMSE <= 1e-3; a
changed topology is rejected and leaves the uploaded scene in place.
"""
import copy

import numpy as np
import pytest

from helpers import c2_mini, relmse
from oracle.bindings import OracleScene
from shiro_amd import abi
from shiro_amd.hip import Context, HipBackend, HipError, RenderRequest
from shiro_amd.scene import RenderSettings, TriangleMesh

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _deformed(scene, amp=0.6):
    out = copy.copy(scene)
    out.meshes = []
    for m in scene.meshes:
        p = np.array(m.positions, np.float32)
        p[:, 1] += (amp * np.sin(0.35 * p[:, 0]) * np.cos(0.25 * p[:, 2])).astype(np.float32)
        p[:, 0] += np.float32(0.3)
        out.meshes.append(TriangleMesh(p, m.normals, m.indices, m.materialIndex, m.uvs))
    return out


@pytest.mark.parametrize("transport", [abi.TRANSPORT_MEGAKERNEL, abi.TRANSPORT_WAVEFRONT])
def test_refit_unmoved_scene_is_identical(hip_ctx, transport):
    wl = c2_mini()
    st = RenderSettings(**{**wl.settings.__dict__, "transport": transport})
    _log(f"refit unmoved, transport {transport}")
    hip_ctx.upload(wl.scene, force=True)
    a = hip_ctx.render(wl.camera, st)
    info = hip_ctx.scene_info()
    hip_ctx.refit(wl.scene)
    b = hip_ctx.render(wl.camera, st)
    assert hip_ctx.scene_info().bvhNodeCount == info.bvhNodeCount
    for k in ("beauty", "albedo", "normal", "depth"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_refit_after_deformation_matches_rebuild_and_oracle(hip_ctx):
    wl = c2_mini()
    moved = _deformed(wl.scene)
    _log("refit deformed c2 mini")
    hip_ctx.upload(wl.scene, force=True)
    hip_ctx.refit(moved)
    r = hip_ctx.render(wl.camera, wl.settings)
    fresh = Context(0)
    try:
        fresh.upload(moved)
        f = fresh.render(wl.camera, wl.settings)
    finally:
        fresh.close()
    same = np.mean(np.all(r.beauty == f.beauty, axis=1))
    _log(f"identical pixels {same:.6f}, relMSE {relmse(r.beauty[:, :3], f.beauty[:, :3]):.3e}")
    assert same >= 0.9999
    assert relmse(r.beauty[:, :3], f.beauty[:, :3]) <= 1e-9
    assert np.array_equal(r.depth, f.depth)
    o = OracleScene(moved, bvh=True).render(wl.camera, wl.settings)
    assert relmse(r.beauty[:, :3], o.beauty[:, :3]) <= 1e-3


def test_refit_rejects_changed_topology(hip_ctx):
    wl = c2_mini()
    hip_ctx.upload(wl.scene, force=True)
    before = hip_ctx.render(wl.camera, wl.settings)
    other = copy.copy(wl.scene)
    other.meshes = list(wl.scene.meshes)
    m = other.meshes[1]
    other.meshes[1] = TriangleMesh(m.positions, m.normals, m.indices[:-3], m.materialIndex, m.uvs)
    with pytest.raises(HipError):
        hip_ctx.refit(other)
    after = hip_ctx.render(wl.camera, wl.settings)
    assert np.array_equal(before.beauty, after.beauty)


def test_backend_refits_deforming_scene():
    """The backend mirror refits when only vertex data changed between
    batches (as the C++ adapter does) and renders what a fresh backend
    renders for the moved scene."""
    wl = c2_mini(width=48, spp=2)
    moved = _deformed(wl.scene, amp=0.3)
    be = HipBackend()
    be.RenderSampleBatch(RenderRequest(wl.scene, wl.camera, wl.settings, 0, 2))
    g = be.RenderSampleBatch(RenderRequest(moved, wl.camera, wl.settings, 0, 2))
    assert be._ctx._topo is not None
    ref = HipBackend().RenderSampleBatch(RenderRequest(moved, wl.camera, wl.settings, 0, 2))
    assert np.mean(np.all(g.beauty == ref.beauty, axis=1)) >= 0.999
