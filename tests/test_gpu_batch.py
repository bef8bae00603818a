"""Batched multi-scene rendering (shiro_hip_render_many) through the C ABI.

Bars: every frame of a batch equals the frame the same scene renders alone
(shiro_hip_render_batch_device, megakernel) bit for bit, in each of the five
batch modes (persistent work-list kernel, hipGraph of per-scene launches,
sequential launches, per-scene wavefront, and the wavefront of every scene
in one captured graph of concurrent branches); a textured scene may sit in
the batch; the captured graph is replayed without a new capture while the
items are unchanged, every replay (with eager work and synchronisations in
between) equals the eager frames, and changed items update it; the batch
autotuner keeps a result that renders the same frames; one batched frame
matches the CPU oracle at relMSE <= 1e-3.
"""
import numpy as np
import pytest

from helpers import relmse
from oracle.bindings import OracleScene
from shiro_amd import abi, workloads
from shiro_amd.hip import Context

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def batch():
    import torch
    wls = [workloads.c4_scene(i, scale=0.12, spp_scale=0.25, tri_scale=0.1, env_size=(128, 64)) for i in range(6)]
    c5 = workloads.c5(scale=0.25, spp=4, frequency=12, tex_size=16, env_size=(64, 32))
    wls.append(c5)
    ctxs, frames = [], []
    for wl in wls:
        c = Context(0)
        c.upload(wl.scene)
        ctxs.append(c)
        n = wl.settings.width * wl.settings.height
        frames.append([torch.empty(n * k, dtype=torch.float32, device="cuda") for k in (4, 3, 3, 1)])
    yield wls, ctxs, frames
    for c in ctxs:
        c.close()


def _singles(wls, ctxs):
    import torch
    out = []
    for wl, c in zip(wls, ctxs):
        n = wl.settings.width * wl.settings.height
        bufs = [torch.empty(n * k, dtype=torch.float32, device="cuda") for k in (4, 3, 3, 1)]
        c.render_device(wl.camera, wl.settings, [b.data_ptr() for b in bufs], 0, wl.settings.samplesPerPixel)
        out.append([b.cpu().numpy() for b in bufs])
    return out


def _items(wls, ctxs, frames):
    return Context.batch_items([(c, wl.camera, wl.settings, [b.data_ptr() for b in f], 0, None)
                                for wl, c, f in zip(wls, ctxs, frames)])


MODES = [abi.MANY_PERSISTENT, abi.MANY_GRAPH, abi.MANY_SEQUENTIAL, abi.MANY_WAVEFRONT, abi.MANY_WAVEFRONT_GRAPH]


@pytest.mark.parametrize("mode", MODES)
def test_batch_frames_equal_single_renders(batch, hip_ctx, mode):
    wls, ctxs, frames = batch
    _log(f"batch mode {mode}: {len(wls)} scenes")
    ref = _singles(wls, ctxs)
    for f in frames:
        for b in f:
            b.fill_(-1.0)
    ms = hip_ctx.render_many(_items(wls, ctxs, frames), mode)
    assert ms > 0.0
    for i, (f, r) in enumerate(zip(frames, ref)):
        for k, (b, rb) in enumerate(zip(f, r)):
            assert np.array_equal(b.cpu().numpy(), rb), (mode, i, k)


@pytest.mark.parametrize("branches", [1, 3, 8])
def test_wavefront_graph_replays_and_updates(batch, branches):
    """The graph of the wavefront batch is captured once and replayed; each
    replay, with eager renders and synchronisations between replays, renders
    the eager frames; new sample ranges for the items capture again and
    update the instantiated graph in place; going back replays the first
    frames again."""
    import torch
    wls, ctxs, frames = batch
    ctx = Context(0)
    ctx.set_tuning(manyBranches=branches)
    _log(f"wavefront graph, {branches} branches")
    ref = _singles(wls, ctxs)
    items = _items(wls, ctxs, frames)
    for rep in range(4):
        for f in frames:
            for b in f:
                b.fill_(-1.0)
        ms = ctx.render_many(items, abi.MANY_WAVEFRONT_GRAPH)
        assert ms > 0.0
        for i, (f, r) in enumerate(zip(frames, ref)):
            for k, (b, rb) in enumerate(zip(f, r)):
                assert np.array_equal(b.cpu().numpy(), rb), (rep, i, k)
        # Eager work between replays.
        ctxs[rep % len(ctxs)].render(wls[rep % len(wls)].camera, wls[rep % len(wls)].settings, 0, 1)
        torch.cuda.synchronize()
    assert ctx.many_graph_stats() == (1, 0)
    moved = Context.batch_items([(c, wl.camera, wl.settings, [b.data_ptr() for b in f], 5, None)
                                 for wl, c, f in zip(wls, ctxs, frames)])
    ctx.render_many(moved, abi.MANY_WAVEFRONT_GRAPH)
    caps, ups = ctx.many_graph_stats()
    _log(f"  captures {caps}, updates {ups}")
    assert caps == 2
    for i, (wl, c, f) in enumerate(zip(wls, ctxs, frames)):
        n = wl.settings.width * wl.settings.height
        bufs = [torch.empty(n * k, dtype=torch.float32, device="cuda") for k in (4, 3, 3, 1)]
        c.render_device(wl.camera, wl.settings, [b.data_ptr() for b in bufs], 5, wl.settings.samplesPerPixel)
        assert np.array_equal(f[0].cpu().numpy(), bufs[0].cpu().numpy()), i
    ctx.render_many(items, abi.MANY_WAVEFRONT_GRAPH)
    for i, (f, r) in enumerate(zip(frames, ref)):
        assert np.array_equal(f[0].cpu().numpy(), r[0]), i
    ctx.close()


def test_batch_autotune_keeps_frames(batch, hip_ctx):
    wls, ctxs, frames = batch
    _log("batch autotune")
    items = _items(wls, ctxs, frames)
    best, report = hip_ctx.autotune_many(items)
    _log(report)
    assert best.batchBlocksPerCU in (1, 2, 4, 8, 16, 32)
    assert hip_ctx.get_tuning().batchBlocksPerCU == best.batchBlocksPerCU
    assert best.manyBranches in (1, 2, 3, 4, 6, 8) and "manyBranches" in report
    ref = _singles(wls, ctxs)
    for mode in (abi.MANY_PERSISTENT, abi.MANY_WAVEFRONT_GRAPH):
        hip_ctx.render_many(items, mode)
        for f, r in zip(frames, ref):
            assert np.array_equal(f[0].cpu().numpy(), r[0]), mode


def test_batched_frame_vs_oracle(batch, hip_ctx):
    wls, ctxs, frames = batch
    hip_ctx.render_many(_items(wls, ctxs, frames), abi.MANY_PERSISTENT)
    i = min(range(len(wls) - 1), key=lambda j: wls[j].samples)
    wl = wls[i]
    _log(f"oracle check of batch item {i} ({wl.settings.width}x{wl.settings.height}, {wl.settings.samplesPerPixel} spp)")
    o = OracleScene(wl.scene).render(wl.camera, wl.settings, threads=8)
    g = frames[i][0].cpu().numpy().reshape(-1, 4)[:, :3]
    err = relmse(g, o.beauty.reshape(-1, 4)[:, :3])
    assert err <= 1e-3, err


def test_batch_rejects_bad_items(hip_ctx):
    from shiro_amd.hip import HipError
    wl = workloads.c1(scale=0.05, spp=1)
    empty = Context(0)
    items = Context.batch_items([(empty, wl.camera, wl.settings, [0, 0, 0, 0], 0, None)])
    with pytest.raises(HipError):
        hip_ctx.render_many(items, abi.MANY_PERSISTENT)
    with pytest.raises(HipError):
        hip_ctx.render_many(items, 7)
    empty.close()
