"""The benchmark workloads at their full sizes, checked through properties
that do not need a full CPU render: both transports give the same frame bit
for bit (133M samples each), the output is finite with alpha in [0, 1] and
depth > 0 where hit, and a band of rows matches the CPU oracle's render of
the same rows at the full sample count (beauty relMSE <= 1e-3).
"""
import numpy as np
import pytest

from helpers import relmse
from oracle.bindings import OracleScene
from shiro_amd import abi, workloads
from shiro_amd.scene import RenderSettings

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module", params=["c2", "c3"])
def full(request):
    wl = workloads.WORKLOADS[request.param]()
    _log(f"full-size {wl.name}: {wl.settings.width}x{wl.settings.height} {wl.settings.samplesPerPixel} spp, "
         f"{wl.scene.triangle_count} triangles")
    return wl


def test_full_size_transports_agree(hip_ctx, full):
    wl = full
    hip_ctx.upload(wl.scene, force=True)
    mk = hip_ctx.render(wl.camera, RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_MEGAKERNEL}))
    wf = hip_ctx.render(wl.camera, RenderSettings(**{**wl.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT}))
    for k in ("beauty", "albedo", "normal", "depth"):
        assert np.array_equal(getattr(mk, k), getattr(wf, k)), k
    b = wf.beauty
    assert np.all(np.isfinite(b)) and np.all(np.isfinite(wf.albedo)) and np.all(np.isfinite(wf.normal))
    assert np.all((b[:, 3] >= 0.0) & (b[:, 3] <= 1.0))
    hit = np.isfinite(wf.depth)
    assert np.all(wf.depth[hit] > 0.0)
    assert np.array_equal(hit, b[:, 3] > 0.0)
    _log(f"{wl.name}: transports identical, {hit.mean():.3f} of pixels hit")

    # A band of full-width rows against the oracle at the full spp.
    st = wl.settings
    rows = 2 if wl.name == "c3" else 3
    y0 = st.height // 2
    o = OracleScene(wl.scene, bvh=True).render_region(wl.camera, st, 0, y0, st.width, rows, threads=8)
    band = slice(y0 * st.width, (y0 + rows) * st.width)
    e = relmse(wf.beauty[band, :3], o.beauty[band, :3])
    _log(f"{wl.name}: rows {y0}-{y0 + rows - 1} relMSE vs oracle {e:.3e}")
    assert e <= 1e-3
