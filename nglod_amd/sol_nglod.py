"""Drop-in for the reference's `sol_nglod` extension module
(sdf-net/lib/extensions/sol_nglod/sol_nglod_kernel.cu:161-192, bound as
`aabb` at :190-192, imported at sdf-net/lib/tracer/SphereTracer.py:37).

    from nglod_amd.sol_nglod import aabb
    x, t, hit = aabb(ray_o, ray_d)

computes the box entry of every ray against [-1, 1]^3 on the GPU with
`nglod_aabb` (the slab test of :77-159, bit-exact).  Same return shapes and
dtypes as the reference: x [N, 3] float32 (the entry point of a hit ray, the
origin otherwise), t [N, 1] float32 (0 where there is no hit), hit [N] bool.
The reference also accepts float64 rays (AT_DISPATCH_FLOATING_TYPES); the
kernel here is FP32, so other dtypes raise TypeError instead of being
converted, which would change the results.
"""
import torch

from . import _lib


def aabb(ray_o, ray_d):
    _lib.require_gpu(ray_o, ray_d)
    if ray_o.dtype != torch.float32 or ray_d.dtype != torch.float32:
        raise TypeError(f"aabb: rays must be float32 (got {ray_o.dtype}, {ray_d.dtype})")
    if ray_o.dim() != 2 or ray_o.shape[1] != 3 or ray_d.shape != ray_o.shape:
        raise ValueError(f"aabb: ray_o and ray_d must both be [N, 3] (got {tuple(ray_o.shape)}, "
                         f"{tuple(ray_d.shape)})")
    o = ray_o.contiguous()
    d = ray_d.contiguous()
    n = o.shape[0]
    x = torch.empty_like(o)
    t = torch.empty(n, 1, dtype=torch.float32, device=o.device)
    hit = torch.empty(n, dtype=torch.bool, device=o.device)
    if n:
        _lib.check(_lib.load().nglod_aabb(_lib.ptr(o), _lib.ptr(d), n, _lib.ptr(x), _lib.ptr(t), _lib.ptr(hit),
                                          _lib.stream_of(o)), "nglod_aabb")
    return [x, t, hit]
