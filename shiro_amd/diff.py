"""Differentiable rendering of texture and dome-map texels with PyTorch
(texture extension; no counterpart in the reference).

`TextureRenderer` owns a libshiro_hip context with an uploaded scene whose
materials read textures; `render(*textures, env=...)` returns the beauty
image (H, W, 3) as a differentiable function of the given texel tensors
(device float32, shaped like the scene's textures) and, optionally, of the
map texels of the dome lights listed in `env_domes` ((H, W, 3) tensors).
The forward pass is the textured render (megakernel or wavefront, the
settings' transport; a new map rebuilds its importance tables on the
device); the backward pass is the path-replay adjoint
(shiro_hip_render_adjoint_env), the exact gradient of the frozen-sampling
estimator at the current texels with the same samples.

    r = TextureRenderer(scene, camera, settings, env_domes=[0])
    albedo = torch.full((128, 128, 3), 0.5, device="cuda", requires_grad=True)
    rough = torch.full((128, 128, 1), 0.5, device="cuda", requires_grad=True)
    sky = torch.ones((128, 256, 3), device="cuda", requires_grad=True)
    img = r.render(albedo, rough, env=[sky])
    loss = ((img - target) ** 2).mean(); loss.backward()
"""
from __future__ import annotations

import torch

from .hip import Context
from .scene import Camera, RenderSettings, Scene


class _RenderTextures(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, sample_start, n_tex, *tensors):
        r = renderer
        r._load(tensors[:n_tex], tensors[n_tex:])
        st = r.settings
        r.ctx.render_device(r.camera, st, [b.data_ptr() for b in r.frame], sample_start, st.samplesPerPixel,
                            synchronize=True)
        # The adjoint replays the paths at the parameters of this forward
        # pass: keep them, and the version of the context's contents they
        # were uploaded as.  Autograd's version counters reject a backward
        # after an in-place change of a saved tensor.
        ctx.save_for_backward(*tensors)
        ctx.renderer = r
        ctx.version = r.version
        ctx.sample_start = sample_start
        ctx.n_tex = n_tex
        return r.frame[0].view(st.height, st.width, 4)[..., :3].clone()

    @staticmethod
    def backward(ctx, grad_out):
        r = ctx.renderer
        st = r.settings
        saved = ctx.saved_tensors
        if r.version != ctx.version:
            # Another forward pass loaded other parameters since: put this
            # pass's texels and maps back (the maps' importance tables are
            # rebuilt on the device, as a fresh upload builds them).
            r._load(saved[:ctx.n_tex], saved[ctx.n_tex:], version=ctx.version)
        g = torch.zeros(st.height, st.width, 4, dtype=torch.float32, device=grad_out.device)
        g[..., :3] = grad_out
        grads = [torch.zeros(t.shape, dtype=torch.float32, device=grad_out.device) for t in saved]
        tex_grads, env_grads = grads[:ctx.n_tex], grads[ctx.n_tex:]
        ptrs = [0] * (max(r.indices) + 1 if r.indices else 0)
        for i, gt in zip(r.indices, tex_grads):
            ptrs[i] = gt.data_ptr()
        env_ptrs = [0] * (max(r.env_domes) + 1 if env_grads else 0)
        for d, gt in zip(r.env_domes, env_grads):
            env_ptrs[d] = gt.data_ptr()
        torch.cuda.current_stream().synchronize()
        r.ctx.render_adjoint(r.camera, st, g.data_ptr(), ptrs, ctx.sample_start, st.samplesPerPixel,
                             env_grad_ptrs=env_ptrs)
        return (None, None, None, *grads)


class TextureRenderer:
    """Renders `scene` with texture `indices` (default: all) and the maps of
    the dome lights `env_domes` (default: none) taken from tensors.  One
    context per renderer; calls are synchronous.  Wavefront forward passes
    needing at most `replay_log_bytes` of replay log (20 bytes per pixel,
    sample and depth; 0: never) keep their hits and shadow-ray visibilities
    for the backward pass."""

    def __init__(self, scene: Scene, camera: Camera, settings: RenderSettings, indices=None, device: int = 0,
                 env_domes=(), replay_log_bytes: int = 2 << 30):
        self.ctx = Context(device)
        self.ctx.upload(scene)
        # The backward pass replays the forward pass's hits and shadow-ray
        # visibilities when nothing changed in between (wavefront forward).
        self.ctx.set_replay_log(replay_log_bytes)
        self.camera = camera
        self.settings = settings
        self.indices = list(range(len(scene.textures))) if indices is None else list(indices)
        self.env_domes = list(env_domes)
        n = settings.width * settings.height
        dev = torch.device("cuda", device)
        self.frame = [torch.empty(n * k, dtype=torch.float32, device=dev) for k in (4, 3, 3, 1)]
        # Version of the parameters loaded into the context: every forward
        # pass loads its own and takes a new number.
        self.version = 0
        self._versions = 0

    def _load(self, texels, envs, version=None):
        """Uploads texel and map tensors into the context (synchronously)."""
        torch.cuda.current_stream().synchronize()
        for i, t in zip(self.indices, texels):
            t = t.detach().contiguous()
            self.ctx.update_texture(i, t.data_ptr())
        for d, t in zip(self.env_domes, envs):
            t = t.detach().contiguous()
            self.ctx.update_environment(d, t.data_ptr())
        if version is None:
            self._versions += 1
            version = self._versions
        self.version = version

    def render(self, *texels, env=(), sample_start: int = 0) -> torch.Tensor:
        if len(texels) != len(self.indices) or len(env) != len(self.env_domes):
            raise ValueError("one tensor per texture index and per env dome expected")
        return _RenderTextures.apply(self, sample_start, len(texels), *texels, *env)
