"""Renderer.render / render_lookat and pinhole ray generation — drop-in for
sdf-net/lib/renderer.py:43-125 and sdf-net/lib/geoutils.py:140-206.

Ray generation keeps the reference's layout (meshgrid 'ij': ray index =
ix * H + iy) and formulas.  The reference adds unseeded jitter to the pixel
grid (geoutils.py:150,152); here jitter is optional and takes a generator, so
frames are reproducible; so is the shadow rays' N(0, 0.01) jitter.

Shading after the trace (renderer.py:131-209, :268-331) runs in the library:
ground-plane shadows (nglod_shadow_rays + a second nglod_trace_dense),
ambient occlusion (nglod_ao, 40 FP32 SDF taps per hit), the blurred shadow map
(nglod_shadow_map) and matcap shading (nglod_matcap_shade).
"""
import ctypes
import math

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib
from .tracer import RenderBuffer, setparam

# Light position of the shadow pass (renderer.py:150).
LIGHT = (-1.5, 4.5, -1.5)


def normalized_grid(width, height, device="cuda", jitter=False, generator=None):
    """Window coordinates of a width x height image, [W, H, 2]: entry [i, j]
    is (x_i, y_j) with x_i running over [-W/H, W/H] left to right and y_j over
    [1, -1] top to bottom (geoutils.py:140-154), so flattening it gives ray
    index i * H + j.  With jitter, x_i gains U[0,1)/W and y_j gains U[0,1)/H,
    drawn from `generator` (the reference draws them unseeded)."""
    xs = torch.linspace(-1, 1, steps=width, device=device) * (width / height)
    ys = torch.linspace(1, -1, steps=height, device=device)
    if jitter:
        xs = xs + torch.rand(width, device=device, generator=generator) * (1.0 / width)
        ys = ys + torch.rand(height, device=device, generator=generator) * (1.0 / height)
    return torch.stack((xs[:, None].expand(width, height), ys[None, :].expand(width, height)), dim=-1)


def look_at(f, t, width, height, mode="persp", fov=90.0, device="cuda", jitter=False, generator=None):
    """Rays of a camera at `f` looking at `t` with world up +y
    (geoutils.py:180-206), [W*H, 3] origins and unit directions.

    The camera frame is forward = normalize(t - f), side =
    normalize(forward x up_world), up = normalize(side x forward).  Pixel
    (x, y) of normalized_grid sits on the image plane one unit ahead of the
    eye at side*x*s + up*y*s + f + forward with s = tan(fov/2), summed in
    that order (the order fixes the FP32 bits that nglod_camera_rays
    reproduces).  'persp': every ray starts at f towards its plane point;
    'ortho': rays start at the plane points and all run along forward."""
    eye = torch.tensor(f, dtype=torch.float32, device=device)
    forward = F.normalize(torch.tensor(t, dtype=torch.float32, device=device) - eye, dim=0)
    world_up = torch.tensor([0.0, 1.0, 0.0], device=device)
    side = F.normalize(torch.linalg.cross(forward, world_up), dim=0)
    up = F.normalize(torch.linalg.cross(side, forward), dim=0)
    win = normalized_grid(width, height, device=device, jitter=jitter, generator=generator).reshape(-1, 2)
    s = math.tan(math.radians(fov / 2))
    plane = side * win[:, 0:1] * s + up * win[:, 1:2] * s + eye + forward
    if mode == "persp":
        return eye.expand_as(plane).contiguous(), F.normalize(plane - eye, dim=-1)
    if mode == "ortho":
        return plane, F.normalize(forward.expand_as(plane), dim=-1)
    raise ValueError(f"look_at: camera mode must be 'persp' or 'ortho', got {mode!r}")


def _u8(t):
    """bool / uint8 tensor -> contiguous uint8 view (1 byte per flag)."""
    t = t.reshape(-1).contiguous()
    return t.view(torch.uint8) if t.dtype == torch.bool else t.to(torch.uint8)


def shadow_rays(ray_o, ray_d, rb, ground_y, noise=None, light=LIGHT):
    """Ground-plane hits and shadow rays (renderer.py:131-160).  Updates rb.x,
    rb.depth, rb.hit and rb.normal in place (a plane hit leaves the surface-hit
    set and takes the plane's point, depth and normal (0, 1, 0)); returns
    (plane_hit [N] bool, shadow_o [N,3], shadow_d [N,3], light_hit [N] bool).
    noise: [N,3] added to the light position before normalising (the
    reference's N(0, 0.01) jitter); None = no jitter."""
    _lib.require_gpu(ray_o, ray_d, rb.x)
    o = ray_o.detach().float().contiguous()
    d = ray_d.detach().float().contiguous()
    n = o.shape[0]
    for name in ("x", "depth", "normal"):
        if not getattr(rb, name).is_contiguous():
            setattr(rb, name, getattr(rb, name).contiguous())
    if rb.hit.dtype != torch.bool or not rb.hit.is_contiguous():
        rb.hit = rb.hit.bool().contiguous()
    if noise is not None:
        noise = noise.detach().float().contiguous()
        if noise.shape != (n, 3):
            raise ValueError("noise must be [N, 3]")
    dev = o.device
    plane = torch.empty(n, dtype=torch.bool, device=dev)
    so = torch.empty(n, 3, dtype=torch.float32, device=dev)
    sd = torch.empty(n, 3, dtype=torch.float32, device=dev)
    light_hit = torch.empty(n, dtype=torch.bool, device=dev)
    lt = (ctypes.c_float * 3)(*[float(v) for v in light])
    _lib.check(_lib.load().nglod_shadow_rays(_lib.ptr(o), _lib.ptr(d), n, float(ground_y), lt, _lib.ptr(noise),
                                             _lib.ptr(rb.x), _lib.ptr(rb.depth), _lib.ptr(rb.hit),
                                             _lib.ptr(rb.normal), _lib.ptr(plane), _lib.ptr(so), _lib.ptr(sd),
                                             _lib.ptr(light_hit), _lib.stream_of(o)), "nglod_shadow_rays")
    return plane, so, sd, light_hit


_ao_ws = {}


def ambient_occlusion(net, x, normal, hit, plane_hit=None, lod=None, form="fp32"):
    """Ambient occlusion (renderer.py:184-209): 40 SDF taps along the normal
    of each hit (and plane hit), ao = clamp(1 - acc, 0.1, 1)^2; [N, 1].
    The taps are FP32 in both forms; form 'mixed' takes them from the FP32
    cumulative grid of `lod` (one grid per tap instead of lod + 1).
    The claim counter (workspace) is one per device: calls on different
    streams of one device must not overlap."""
    _lib.require_gpu(x, normal)
    lod = lod if lod is not None else (net.lod if net.lod is not None else net.num_lods - 1)
    x = x.detach().float().reshape(-1, 3).contiguous()
    normal = normal.detach().float().reshape(-1, 3).contiguous()
    n = x.shape[0]
    h = _u8(hit)
    ph = _u8(plane_hit) if plane_hit is not None else None
    ao = torch.empty(n, 1, dtype=torch.float32, device=x.device)
    key = (x.device.type, x.device.index)
    lib = _lib.load()
    ws = _ao_ws.get(key)
    if ws is None:
        ws = torch.empty(max(int(lib.nglod_ao_workspace_size()), 256), dtype=torch.uint8, device=x.device)
        _ao_ws[key] = ws
    fm = _lib.MIXED if form == "mixed" else _lib.FP32
    _lib.check(lib.nglod_ao(net.descriptor(fm), _lib.ptr(x), _lib.ptr(normal), _lib.ptr(h), _lib.ptr(ph), n,
                            int(lod), fm, _lib.ptr(ao), _lib.ptr(ws), ws.numel(), _lib.stream_of(x)), "nglod_ao")
    return ao


def shadow_map(shadow_hit, light_hit, width, height):
    """clamp(1 - shadow + 0.9, 0, 1) blurred by gaussian_filter(sigma=2) over
    the [width, height] frame (renderer.py:301-304); [width, height] float."""
    s = _u8(shadow_hit)
    lh = _u8(light_hit)
    _lib.require_gpu(s, lh)
    n = width * height
    if s.numel() != n or lh.numel() != n:
        raise ValueError("shadow_hit / light_hit must hold width * height flags")
    out = torch.empty(width, height, dtype=torch.float32, device=s.device)
    tmp = torch.empty_like(out)
    _lib.check(_lib.load().nglod_shadow_map(_lib.ptr(s), _lib.ptr(lh), int(width), int(height), _lib.ptr(out),
                                            _lib.ptr(tmp), _lib.stream_of(s)), "nglod_shadow_map")
    return out


def matcap_sampler(path):
    """The matcap texture of geoutils.matcap_sampler (geoutils.py:237-250) as
    a float32 tensor [width, height, channels] (the image transposed to
    x-major, values 0..255), for matcap_shade."""
    from PIL import Image
    img = np.array(Image.open(path))
    if img.ndim != 3 or img.shape[2] < 3:
        raise ValueError("matcap texture must be an RGB or RGBA image")
    return torch.from_numpy(np.ascontiguousarray(img.transpose(1, 0, 2)).astype(np.float32))


def matcap_shade(view, normal, hit, texture, mm=None, shadow=None, ao=None):
    """Matcap shading (renderer.py:279-307): rgb [N, 3] from the view rays and
    normals through spherical_envmap and the bilinear texture lookup; pixels
    without a hit get rgb = normal = 1 (normal is updated in place); rgb is
    scaled by `shadow` (a blurred shadow map) and `ao` when given."""
    v = view.detach().float().reshape(-1, 3).contiguous()
    if not normal.is_contiguous():
        raise ValueError("normal must be contiguous (it is updated in place)")
    nrm = normal.reshape(-1, 3)
    h = _u8(hit)
    _lib.require_gpu(v, nrm, h)
    n = v.shape[0]
    tex = texture.to(device=v.device, dtype=torch.float32).contiguous()
    if tex.dim() != 3:
        raise ValueError("texture must be [width, height, channels]")
    sm = shadow.detach().float().reshape(-1).contiguous() if shadow is not None else None
    a = ao.detach().float().reshape(-1).contiguous() if ao is not None else None
    mmp = None
    if mm is not None:
        mmp = (ctypes.c_float * 9)(*[float(e) for e in torch.as_tensor(mm).float().reshape(-1).tolist()])
    rgb = torch.empty(n, 3, dtype=torch.float32, device=v.device)
    _lib.check(_lib.load().nglod_matcap_shade(_lib.ptr(v), _lib.ptr(nrm), _lib.ptr(h), n, mmp, _lib.ptr(tex),
                                              tex.shape[0], tex.shape[1], tex.shape[2], _lib.ptr(sm), _lib.ptr(a),
                                              _lib.ptr(rgb), _lib.stream_of(v)), "nglod_matcap_shade")
    return rgb


class Renderer:
    """Renderer (sdf-net/lib/renderer.py:38-331): batching, ground-plane
    shadows, relative depth, ambient occlusion, reshape to [W, H, .], and the
    matcap shading of shade_tensor / shade_images.

    The ground plane height comes from `ground_height` (or args.ground_height);
    the reference's fallback (min y of 1M surface samples of `self.sdf_net`,
    renderer.py:80-87) refers to an attribute it never sets, so shadows here
    require the height.  The shadow rays' jitter is drawn from `generator`."""

    def __init__(self, tracer, args=None, render_res: list = None, camera_clamp: list = None,
                 camera_proj: str = None, render_batch: int = None, shading_mode: str = None,
                 matcap_path: str = None, shadow: bool = None, ao: bool = None, perf: bool = None, device=None,
                 ground_height: float = None, matcap=None, generator=None):
        self.args = args
        self.camera_clamp = tracer.camera_clamp
        self.render_res = setparam(args, render_res, "render_res") or [512, 512]
        self.camera_proj = setparam(args, camera_proj, "camera_proj") or "persp"
        self.render_batch = setparam(args, render_batch, "render_batch") or 0
        self.shading_mode = setparam(args, shading_mode, "shading_mode") or "matcap"
        self.matcap_path = setparam(args, matcap_path, "matcap_path")
        self.shadow = setparam(args, shadow, "shadow") or False
        self.ao = setparam(args, ao, "ao") or False
        self.perf = setparam(args, perf, "perf") or False
        self.device = setparam(args, device, "device") or "cuda"
        self.width, self.height = self.render_res
        self.tracer = tracer
        self.generator = generator
        self.min_y = setparam(args, ground_height, "ground_height")
        self._matcap = matcap
        if self.shadow and self.min_y is None:
            raise ValueError("shadows need the ground plane height (ground_height)")

    def render_lookat(self, net, f=[0, 0, 1], t=[0, 0, 0], fov=30.0, camera_proj="persp", device="cuda", mm=None,
                      jitter=False, generator=None):
        ray_o, ray_d = look_at(f, t, self.width, self.height, fov=fov, mode=camera_proj, device=device,
                               jitter=jitter, generator=generator)
        if mm is not None:
            mm = mm.to(ray_o.device)
            ray_o = torch.mm(ray_o, mm)
            ray_d = torch.mm(ray_d, mm)
        return self.render(net, ray_o, ray_d)

    def _trace(self, net, ray_o, ray_d):
        with torch.no_grad():
            if self.render_batch > 0:
                rb = RenderBuffer()
                for origin, direction in zip(torch.split(ray_o, self.render_batch),
                                             torch.split(ray_d, self.render_batch)):
                    rb += self.tracer(net, origin, direction)
                return rb
            return self.tracer(net, ray_o, ray_d)

    def render(self, net, ray_o, ray_d):
        rb = self._trace(net, ray_o, ray_d)
        plane_hit = None
        if self.shadow:
            noise = torch.empty(ray_o.shape[0], 3, device=ray_o.device).normal_(0.0, 0.01, generator=self.generator)
            plane_hit, so, sd, light_hit = shadow_rays(ray_o, ray_d, rb, self.min_y, noise=noise)
            rb.shadow = self._trace(net, so, sd).hit & light_hit
        rb.relative_depth = torch.clamp(rb.depth, 0.0, self.camera_clamp[1]) / self.camera_clamp[1]
        if self.ao:
            form = getattr(self.tracer, "form", "fp32")
            rb.ao = ambient_occlusion(net, rb.x, rb.normal, rb.hit, plane_hit, form=form)
        rb.view = ray_d
        return rb.reshape(self.width, self.height, -1)

    def _matcap_texture(self):
        if self._matcap is None:
            if self.matcap_path is None:
                raise ValueError("matcap shading needs a matcap texture (matcap_path or matcap=)")
            self._matcap = matcap_sampler(self.matcap_path)
        return self._matcap

    def shade_tensor(self, net, f=[0, 0, 1], t=[0, 0, 0], fov=30.0, mm=None, jitter=False):
        """Non-differentiable shading for visualisation (renderer.py:268-308)."""
        rb = self.render_lookat(net, f=f, t=t, fov=fov, mm=mm, jitter=jitter, generator=self.generator)
        if self.shading_mode == "matcap":
            smap = None
            if self.shadow:
                s = rb.shadow.reshape(-1)
                smap = shadow_map(s, s, self.width, self.height)
            rb.normal = rb.normal.contiguous()
            rgb = matcap_shade(rb.view, rb.normal, rb.hit, self._matcap_texture(), mm=mm, shadow=smap,
                               ao=rb.ao if self.ao else None)
            rb.rgb = rgb.reshape(self.width, self.height, 3)
        elif self.shading_mode == "rb":
            if rb.rgb is None:
                raise ValueError("No rgb in buffer; change shading-mode")
        else:
            raise NotImplementedError(self.shading_mode)
        return rb

    def shade_images(self, net, f=[0, 0, 1], t=[0, 0, 0], fov=30.0, aa=1, mm=None):
        """shade_tensor averaged over `aa` jittered frames (renderer.py:310-331),
        on the CPU and transposed to [H, W, .]."""
        if mm is None:
            mm = torch.eye(3)
        if aa > 1:
            rb = RenderBuffer.mean(*[self.shade_tensor(net, f=f, t=t, fov=fov, mm=mm, jitter=True)
                                     for _ in range(aa)])
        else:
            rb = self.shade_tensor(net, f=f, t=t, fov=fov, mm=mm)
        return rb.cpu().transpose()
