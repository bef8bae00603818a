"""shiro_hip_render_batch_device_noalloc (no counterpart in the reference):
the entry point for callers that cannot allocate device memory.  Its frames
equal those of shiro_hip_render_batch_device bit for bit, on two calls in a
row (the second smaller, so it runs over what the first left in the shared
scratch array), including a frame whose path state exceeds the scratch
array (covered in passes over tile ranges), and its calls allocate no
device memory."""
import ctypes as C

import numpy as np
import pytest

from helpers import c1_mini, c3_mini
from shiro_amd import abi
from shiro_amd.scene import RenderSettings

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _frame(w, h):
    import torch
    return [torch.full((w * h * k,), -1.0, dtype=torch.float32, device="cuda") for k in (4, 3, 3, 1)]


def _cframe(bufs, w, h):
    f = abi.Frame()
    f.beauty, f.albedo, f.normal, f.depth = [C.cast(C.c_void_p(b.data_ptr()), C.POINTER(C.c_float)) for b in bufs]
    f.width, f.height = w, h
    return f


def _render(ctx, fn, wl, st, start, count, bufs):
    f = _cframe(bufs, st.width, st.height)
    cam, cst = wl.camera.to_c(), st.to_c()
    rc = fn(ctx.h, C.byref(cam), C.byref(cst), start, count, C.byref(f), 1)
    if rc != 0:
        raise RuntimeError(f"rc {rc}: {ctx.L.shiro_hip_last_error(ctx.h).decode()}")


def test_noalloc_equals_render_batch_device():
    import torch
    from shiro_amd.hip import Context
    # A fresh context: the regular entry point has not grown its buffers yet,
    # so an allocation by the no-allocation entry point would show in the
    # device's free memory.
    ctx = Context(0)
    try:
        big = c3_mini(width=1280, spp=8, objects=8)  # 1280x720x8: 7.4M paths, beyond the 256 MiB scratch
        small = c1_mini(size=96, spp=2)
        cases = [(big, RenderSettings(**{**big.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT}), 0, 8),
                 (small, RenderSettings(**{**small.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT}), 3, 2)]
        noalloc = ctx.L.shiro_hip_render_batch_device_noalloc
        # Warm-up on a tiny frame: the first launches load the kernels and let
        # the runtime size its private (register spill) segment, which is
        # not memory this library allocates.
        tiny = c1_mini(size=16, spp=1)
        ctx.upload(tiny.scene)
        _render(ctx, noalloc, tiny, RenderSettings(**{**tiny.settings.__dict__, "transport": abi.TRANSPORT_WAVEFRONT}),
                0, 1, _frame(16, 16))
        for wl, st, start, count in cases:
            _log(f"no-alloc {wl.name} {st.width}x{st.height} samples [{start}, {start + count})")
            ctx.upload(wl.scene)
            out = _frame(st.width, st.height)
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            _render(ctx, noalloc, wl, st, start, count, out)
            free1 = torch.cuda.mem_get_info()[0]
            _log(f"  device free memory {free0 >> 20} -> {free1 >> 20} MiB")
            assert free0 - free1 < (64 << 20), (free0, free1)
            ref = _frame(st.width, st.height)
            _render(ctx, ctx.L.shiro_hip_render_batch_device, wl, st, start, count, ref)
            mk = _frame(st.width, st.height)
            _render(ctx, ctx.L.shiro_hip_render_batch_device, wl,
                    RenderSettings(**{**st.__dict__, "transport": abi.TRANSPORT_MEGAKERNEL}), start, count, mk)
            torch.cuda.synchronize()
            for k, (a, b, m) in enumerate(zip(out, ref, mk)):
                a, b, m = a.cpu().numpy(), b.cpu().numpy(), m.cpu().numpy()
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (wl.name, k)
                assert np.array_equal(a.view(np.uint32), m.view(np.uint32)), (wl.name, k)
            assert np.isfinite(out[0].cpu().numpy()).all()
    finally:
        ctx.close()
