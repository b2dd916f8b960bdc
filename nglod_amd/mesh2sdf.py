"""Drop-in for the reference's `mesh2sdf` extension module
(sdf-net/lib/extensions/mesh2sdf_cuda/mesh2sdf_kernel.cu:895-927, bound as
`mesh2sdf_gpu` at :1009, called from sdf-net/lib/torchgp/compute_sdf.py:38).

    from nglod_amd import mesh2sdf
    dist = mesh2sdf.mesh2sdf_gpu(points, mesh)[0]

points [N, 3] float32 and mesh [F, 3, 3] float32 (the triangle soup V[F]) on
the GPU; returns [dist] with dist [N] float32, the signed distance (negative
inside, by the 13-direction stabbing rule of :341-353 and :584-602).  The
arithmetic is FP64 (`nglod_mesh2sdf_f64`): the reference computes in FP32,
so results agree to the FP32 rounding of the exact distance, not bit for bit.
"""
import torch

from . import _lib


def mesh2sdf_gpu(points, mesh):
    _lib.require_gpu(points, mesh)
    if points.dtype != torch.float32 or mesh.dtype != torch.float32:
        raise TypeError(f"mesh2sdf_gpu: points and mesh must be float32 (got {points.dtype}, {mesh.dtype})")
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError(f"mesh2sdf_gpu: points must be [N, 3] (got {tuple(points.shape)})")
    if mesh.dim() != 3 or tuple(mesh.shape[1:]) != (3, 3):
        raise ValueError(f"mesh2sdf_gpu: mesh must be [F, 3, 3] (got {tuple(mesh.shape)})")
    p = points.contiguous()
    tris = mesh.contiguous()
    n, nf = p.shape[0], tris.shape[0]
    dist = torch.empty(n, dtype=torch.float32, device=p.device)
    if n:
        lib = _lib.load()
        ws = torch.empty(max(1, lib.nglod_mesh2sdf_workspace_size(nf)), dtype=torch.uint8, device=p.device)
        _lib.check(lib.nglod_mesh2sdf_f64(_lib.ptr(p), n, _lib.ptr(tris), nf, _lib.ptr(dist), None, _lib.ptr(ws),
                                          ws.numel(), _lib.stream_of(p)), "nglod_mesh2sdf_f64")
    return [dist]
