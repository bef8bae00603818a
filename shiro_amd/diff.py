"""Differentiable rendering of texture texels with PyTorch (texture
extension; no counterpart in the reference).

`TextureRenderer` owns a libshiro_hip context with an uploaded scene whose
materials read textures; `render(*textures)` returns the beauty image (H, W,
3) as a differentiable function of the given texel tensors (device float32,
shaped like the scene's textures).  The forward pass is the textured render
(megakernel or wavefront, the settings' transport); the backward pass is the
path-replay adjoint (shiro_hip_render_adjoint), the exact gradient of the
frozen-sampling estimator at the current texels with the same samples.

    r = TextureRenderer(scene, camera, settings)
    albedo = torch.full((128, 128, 3), 0.5, device="cuda", requires_grad=True)
    rough = torch.full((128, 128, 1), 0.5, device="cuda", requires_grad=True)
    img = r.render(albedo, rough)
    loss = ((img - target) ** 2).mean(); loss.backward()
"""
from __future__ import annotations

import torch

from .hip import Context
from .scene import Camera, RenderSettings, Scene


class _RenderTextures(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, sample_start, *texels):
        r = renderer
        torch.cuda.current_stream().synchronize()
        for i, t in zip(r.indices, texels):
            r.ctx.update_texture(i, t.detach().contiguous().data_ptr())
        st = r.settings
        r.ctx.render_device(r.camera, st, [b.data_ptr() for b in r.frame], sample_start, st.samplesPerPixel,
                            synchronize=True)
        ctx.renderer = r
        ctx.sample_start = sample_start
        ctx.shapes = [t.shape for t in texels]
        return r.frame[0].view(st.height, st.width, 4)[..., :3].clone()

    @staticmethod
    def backward(ctx, grad_out):
        r = ctx.renderer
        st = r.settings
        g = torch.zeros(st.height, st.width, 4, dtype=torch.float32, device=grad_out.device)
        g[..., :3] = grad_out
        grads = [torch.zeros(s, dtype=torch.float32, device=grad_out.device) for s in ctx.shapes]
        ptrs = [0] * (max(r.indices) + 1)
        for i, gt in zip(r.indices, grads):
            ptrs[i] = gt.data_ptr()
        torch.cuda.current_stream().synchronize()
        r.ctx.render_adjoint(r.camera, st, g.data_ptr(), ptrs, ctx.sample_start, st.samplesPerPixel)
        return (None, None, *grads)


class TextureRenderer:
    """Renders `scene` with texture `indices` (default: all) taken from
    tensors.  One context per renderer; calls are synchronous."""

    def __init__(self, scene: Scene, camera: Camera, settings: RenderSettings, indices=None, device: int = 0):
        self.ctx = Context(device)
        self.ctx.upload(scene)
        self.camera = camera
        self.settings = settings
        self.indices = list(range(len(scene.textures))) if indices is None else list(indices)
        n = settings.width * settings.height
        dev = torch.device("cuda", device)
        self.frame = [torch.empty(n * k, dtype=torch.float32, device=dev) for k in (4, 3, 3, 1)]

    def render(self, *texels, sample_start: int = 0) -> torch.Tensor:
        return _RenderTextures.apply(self, sample_start, *texels)
