"""Renderer.render / render_lookat and pinhole ray generation — drop-in for
sdf-net/lib/renderer.py:43-125 and sdf-net/lib/geoutils.py:140-206.

Ray generation keeps the reference's layout (meshgrid 'ij': ray index =
ix * H + iy) and formulas.  The reference adds unseeded jitter to the pixel
grid (geoutils.py:150,152); here jitter is optional and takes a generator, so
frames are reproducible.  Shading (matcap, shadows, AO) is a post-trace
consumer and is not part of this package.
"""
import math

import torch
import torch.nn.functional as F

from .tracer import RenderBuffer, setparam


def normalized_grid(width, height, device="cuda", jitter=False, generator=None):
    """grid[x, y] -> normalized window coordinates (geoutils.py:140-154)."""
    window_x = torch.linspace(-1, 1, steps=width, device=device) * (width / height)
    window_y = torch.linspace(1, -1, steps=height, device=device)
    if jitter:
        window_x = window_x + torch.rand(width, device=device, generator=generator) * (1.0 / width)
        window_y = window_y + torch.rand(height, device=device, generator=generator) * (1.0 / height)
    return torch.stack(torch.meshgrid(window_x, window_y, indexing="ij")).permute(1, 2, 0)


def look_at(f, t, width, height, mode="persp", fov=90.0, device="cuda", jitter=False, generator=None):
    """Ray origins and directions of a look-at camera (geoutils.py:180-206)."""
    camera_origin = torch.tensor(f, dtype=torch.float32, device=device)
    camera_view = F.normalize(torch.tensor(t, dtype=torch.float32, device=device) - camera_origin, dim=0)
    camera_right = F.normalize(torch.cross(camera_view, torch.tensor([0.0, 1.0, 0.0], device=device), dim=0), dim=0)
    camera_up = F.normalize(torch.cross(camera_right, camera_view, dim=0), dim=0)
    coord = normalized_grid(width, height, device=device, jitter=jitter, generator=generator)
    tanh = math.tan(math.radians(fov / 2))
    ray_origin = camera_right * coord[..., 0, None] * tanh + camera_up * coord[..., 1, None] * tanh + \
        camera_origin + camera_view
    ray_origin = ray_origin.reshape(-1, 3)
    ray_offset = camera_view.unsqueeze(0).repeat(ray_origin.shape[0], 1)
    if mode == "ortho":
        ray_dir = F.normalize(ray_offset, dim=-1)
    elif mode == "persp":
        ray_dir = F.normalize(ray_origin - camera_origin, dim=-1)
        ray_origin = camera_origin.repeat(ray_dir.shape[0], 1)
    else:
        raise ValueError("Invalid camera mode!")
    return ray_origin, ray_dir


class Renderer:
    """Renderer (sdf-net/lib/renderer.py:38-125): batching, relative depth,
    reshape to [W, H, .]."""

    def __init__(self, tracer, args=None, render_res: list = None, camera_clamp: list = None,
                 camera_proj: str = None, render_batch: int = None, perf: bool = None, device=None):
        self.args = args
        self.camera_clamp = tracer.camera_clamp
        self.render_res = setparam(args, render_res, "render_res") or [512, 512]
        self.camera_proj = setparam(args, camera_proj, "camera_proj") or "persp"
        self.render_batch = setparam(args, render_batch, "render_batch") or 0
        self.perf = setparam(args, perf, "perf") or False
        self.device = setparam(args, device, "device") or "cuda"
        self.width, self.height = self.render_res
        self.tracer = tracer

    def render_lookat(self, net, f=[0, 0, 1], t=[0, 0, 0], fov=30.0, camera_proj="persp", device="cuda", mm=None,
                      jitter=False, generator=None):
        ray_o, ray_d = look_at(f, t, self.width, self.height, fov=fov, mode=camera_proj, device=device,
                               jitter=jitter, generator=generator)
        if mm is not None:
            mm = mm.to(ray_o.device)
            ray_o = torch.mm(ray_o, mm)
            ray_d = torch.mm(ray_d, mm)
        return self.render(net, ray_o, ray_d)

    def render(self, net, ray_o, ray_d):
        with torch.no_grad():
            if self.render_batch > 0:
                rb = RenderBuffer()
                for origin, direction in zip(torch.split(ray_o, self.render_batch),
                                             torch.split(ray_d, self.render_batch)):
                    rb += self.tracer(net, origin, direction)
            else:
                rb = self.tracer(net, ray_o, ray_d)
        rb.relative_depth = torch.clamp(rb.depth, 0.0, self.camera_clamp[1]) / self.camera_clamp[1]
        rb.view = ray_d
        return rb.reshape(self.width, self.height, -1)
