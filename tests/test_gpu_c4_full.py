"""The full c4 batch (32 scenes of 256^2-1024^2 pixels at 4/16/64 spp with
10k-500k triangles and their own 1024x512 domes: 3.6M triangles and 542M
samples in all) through every batch mode of shiro_hip_render_many.

Bars: every frame of every mode equals the frame its scene renders alone
(megakernel, shiro_hip_render_batch_device) bit for bit; frames are finite,
alpha (hit fraction) in [0, 1], depth > 0 where hit; one item matches the
CPU oracle on a band of rows at relMSE <= 1e-3.
"""
import numpy as np
import pytest

from helpers import relmse
from oracle.bindings import OracleScene
from shiro_amd import abi, workloads
from shiro_amd.hip import Context

pytestmark = pytest.mark.gpu

MODES = {"persistent": abi.MANY_PERSISTENT, "graph": abi.MANY_GRAPH, "sequential": abi.MANY_SEQUENTIAL,
         "wavefront": abi.MANY_WAVEFRONT, "wavefront_graph": abi.MANY_WAVEFRONT_GRAPH}


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def c4():
    import torch
    wls = workloads.c4()
    _log(f"c4: {len(wls)} scenes, {sum(w.scene.triangle_count for w in wls)} triangles, "
         f"{sum(w.samples for w in wls)} samples")
    ctxs, frames, singles = [], [], []
    for wl in wls:
        c = Context(0)
        c.upload(wl.scene)
        ctxs.append(c)
        n = wl.settings.width * wl.settings.height
        frames.append([torch.empty(n * k, dtype=torch.float32, device="cuda") for k in (4, 3, 3, 1)])
        bufs = [torch.empty(n * k, dtype=torch.float32, device="cuda") for k in (4, 3, 3, 1)]
        c.render_device(wl.camera, wl.settings, [b.data_ptr() for b in bufs], 0, wl.settings.samplesPerPixel)
        singles.append([b.cpu().numpy() for b in bufs])
    yield wls, ctxs, frames, singles
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("mode", list(MODES), ids=list(MODES))
def test_c4_batch_equals_single_renders(c4, mode):
    wls, ctxs, frames, singles = c4
    launcher = Context(0)
    items = Context.batch_items([(c, wl.camera, wl.settings, [b.data_ptr() for b in f], 0, None)
                                 for wl, c, f in zip(wls, ctxs, frames)])
    for f in frames:
        for b in f:
            b.fill_(-1.0)
    ms = launcher.render_many(items, MODES[mode])
    _log(f"c4 {mode}: {ms:.1f} ms")
    for i, (f, r) in enumerate(zip(frames, singles)):
        beauty = f[0].cpu().numpy()
        for k, (b, rb) in enumerate(zip(f, r)):
            assert np.array_equal(b.cpu().numpy(), rb), (mode, i, k)
        px = beauty.reshape(-1, 4)
        assert np.isfinite(px).all(), (mode, i)
        assert (px[:, 3] >= 0).all() and (px[:, 3] <= 1).all(), (mode, i)
        d = f[3].cpu().numpy()
        assert (d[np.isfinite(d)] > 0).all(), (mode, i)
    launcher.close()


def test_c4_item_vs_oracle(c4):
    wls, ctxs, frames, singles = c4
    i = min(range(len(wls)), key=lambda j: wls[j].samples)
    wl = wls[i]
    st = wl.settings
    rows = 24
    y0 = st.height // 2 - rows // 2
    _log(f"c4 item {i}: {st.width}x{st.height} {st.samplesPerPixel} spp, rows {y0}-{y0 + rows - 1} vs oracle")
    o = OracleScene(wl.scene, bvh=True).render_region(wl.camera, st, 0, y0, st.width, rows, threads=8)
    band = slice(y0 * st.width, (y0 + rows) * st.width)
    g = singles[i][0].reshape(-1, 4)[band, :3]
    e = relmse(g, o.beauty[band, :3])
    _log(f"  relMSE {e:.3e}")
    assert e <= 1e-3, e
