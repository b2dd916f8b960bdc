"""CPU restatement of the Renderer's post-trace shading
(sdf-net/lib/renderer.py:131-209 and :268-331, sdf-net/lib/geoutils.py:237-275).

TEST INFRASTRUCTURE ONLY: imported by tests/,
as the checker.  The product package never imports it.

FP32 steps are restated op by op in NumPy float32, in the reference's
evaluation order (each PyTorch elementwise op rounds once).  The matcap lookup
and the blur are the library routines the reference itself calls
(scipy.interpolate.RegularGridInterpolator, scipy.ndimage.gaussian_filter).
"""
import numpy as np
from scipy.interpolate import RegularGridInterpolator
from scipy.ndimage import gaussian_filter

f32 = np.float32
LIGHT = (-1.5, 4.5, -1.5)  # renderer.py:150
AO_TAPS = 40


def shadow_rays(o, d, x, t, hit, normal, ground_y, noise=None, light=LIGHT):
    """renderer.py:131-160.  Returns updated (x, t, hit, normal) copies and
    plane_hit, shadow_o, shadow_d, light_hit."""
    o = np.asarray(o, f32)
    d = np.asarray(d, f32)
    x = np.array(x, f32).reshape(-1, 3)
    t = np.array(t, f32).reshape(-1)
    hit = np.array(hit, bool).reshape(-1)
    normal = np.array(normal, f32).reshape(-1, 3)
    n = o.shape[0]
    with np.errstate(divide="ignore", invalid="ignore"):
        rate = -d[:, 1]                               # :135
        delta = o[:, 1] - f32(ground_y)               # :137
        plane_t = delta / rate                        # :138
    plane = (plane_t > 0) & (plane_t < 500)           # :139 (the |rate| < 1e-5 reset of :136 is overridden here)
    plane &= plane_t < t                              # :140
    hit &= ~plane                                     # :142
    t[plane] = plane_t[plane]                         # :144
    x[plane] = o[plane] + d[plane] * plane_t[plane][:, None]  # :145
    normal[plane] = 0                                 # :146-147
    normal[plane, 1] = 1
    so = x + f32(0.1) * normal                        # :151
    e = np.zeros((n, 3), f32) if noise is None else np.asarray(noise, f32).reshape(n, 3)
    v = (e + np.asarray(light, f32)[None, :]) - so    # :152-153
    nn = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
    nn = np.maximum(nn, f32(1e-12))                   # F.normalize, :154
    sd = v / nn[:, None]
    light_hit = ((sd[:, 0] * normal[:, 0] + sd[:, 1] * normal[:, 1]) + sd[:, 2] * normal[:, 2]) > 0  # :156
    return dict(x=x, t=t, hit=hit, normal=normal, plane=plane, shadow_o=so, shadow_d=sd, light_hit=light_hit)


def ao_dist():
    """Tap distances of renderer.py:189-193, rounded to FP32."""
    return np.array([0.1 * 0.25 * (float(i + 1) / float(AO_TAPS + 1)) ** 1.6 for i in range(AO_TAPS)], f32)


def ambient_occlusion(sdf_fn, x, normal, hit, plane=None):
    """renderer.py:184-209.  sdf_fn(points [m,3] float32) -> [m] float32."""
    x = np.asarray(x, f32).reshape(-1, 3)
    normal = np.asarray(normal, f32).reshape(-1, 3)
    hit = np.asarray(hit, bool).reshape(-1)
    plane = np.zeros_like(hit) if plane is None else np.asarray(plane, bool).reshape(-1)
    acc = np.zeros(x.shape[0], f32)
    for dk in ao_dist():
        q = x + normal * dk                                       # :194
        if hit.any():
            r = np.asarray(sdf_fn(q[hit]), f32).reshape(-1)          # :198
            acc[hit] = acc[hit] + f32(3.5) * np.maximum((dk - r) - f32(0.0015), f32(0))  # :205
        if plane.any():
            r = np.minimum(np.asarray(sdf_fn(q[plane]), f32).reshape(-1), dk)  # :201-203
            acc[plane] = acc[plane] + f32(3.5) * np.maximum((dk - r) - f32(0.0015), f32(0))  # :204
    ao = np.clip(f32(1.0) - acc, f32(0.1), f32(1.0))                  # :208
    return ao * ao                                                    # :209


def shadow_map(shadow, width, height):
    """renderer.py:302-303: clamp(1 - shadow + 0.9, 0, 1) blurred with
    gaussian_filter(sigma=2) over the [width, height] frame."""
    s = np.asarray(shadow, bool).reshape(width, height).astype(f32)
    m = np.clip((f32(1.0) - s) + f32(0.9), f32(0), f32(1)).astype(f32)
    return gaussian_filter(m, sigma=2)


def spherical_envmap(view, normal):
    """geoutils.py:253-275 in FP32."""
    v = np.array(view, f32).reshape(-1, 3)
    nrm = np.asarray(normal, f32).reshape(-1, 3)
    v[:, 2] *= f32(-1)                                                 # :265
    dot = (nrm[:, 0] * v[:, 0] + nrm[:, 1] * v[:, 1]) + nrm[:, 2] * v[:, 2]  # :266
    r = v - (f32(2.0) * dot)[:, None] * nrm                           # :267
    r[:, 2] -= f32(1.0)                                               # :268
    m = f32(2.0) * np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2])  # :269
    with np.errstate(divide="ignore", invalid="ignore"):
        vn = r[:, :2] / m[:, None] + f32(0.5)                         # :270
    vn = f32(1.0) - vn                                                 # :271
    vn = np.clip(vn, f32(0), f32(1))                                   # :273
    vn[np.isnan(vn)] = 0                                               # :274
    return vn


def matcap_shade(view, normal, hit, texture, mm=None, shadow=None, ao=None):
    """renderer.py:279-307 for one frame, rays flattened: returns (rgb [n,3]
    float32, normal with misses set to 1)."""
    view = np.asarray(view, f32).reshape(-1, 3)
    if mm is not None:
        m = np.asarray(mm, f32).reshape(3, 3)
        view = np.stack([(view[:, 0] * m[j, 0] + view[:, 1] * m[j, 1]) + view[:, 2] * m[j, 2] for j in range(3)], 1)
    normal = np.array(normal, f32).reshape(-1, 3)
    hit = np.asarray(hit, bool).reshape(-1)
    tex = np.asarray(texture, np.float64)
    vn = spherical_envmap(view, normal)
    interp = RegularGridInterpolator((np.linspace(0, 1, tex.shape[0]), np.linspace(0, 1, tex.shape[1])), tex)
    rgb = interp(vn.astype(np.float64)).astype(f32)[:, :3] / f32(255.0)  # :289-290
    normal[~hit] = 1.0                                                     # :297
    rgb[~hit] = 1.0                                                        # :298
    if shadow is not None:
        rgb = rgb * np.asarray(shadow, f32).reshape(-1, 1)                 # :304
    if ao is not None:
        rgb = rgb * np.asarray(ao, f32).reshape(-1, 1)                     # :307
    return rgb.astype(f32), normal
