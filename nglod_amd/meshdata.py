"""GPU-resident training data for a mesh: SDF targets in FP64 and samplers
(drop-in for the parts of sdf-net/lib/datasets/MeshDataset.py and
sdf-net/lib/torchgp that produce (points, sdf) pairs).

compute_sdf / point_sample keep the reference's names and arguments
(torchgp/compute_sdf.py:26-40, torchgp/point_sample.py:29-57).  Random draws
come from a counter-based SplitMix64 stream keyed by an explicit seed, so a
batch is reproducible (the reference draws from torch's global RNG).
"""
import numpy as np
import torch

from . import _lib, mesh as _mesh


class MeshData:
    """Triangles, face-area CDF and the mesh2sdf workspace on one device."""

    def __init__(self, V, F, device="cuda"):
        V = np.asarray(V, dtype=np.float32)
        F = np.asarray(F, dtype=np.int64)
        tris = _mesh.triangles(V, F).astype(np.float32)
        self.device = torch.device(device)
        self.V = torch.from_numpy(V).to(self.device)
        self.F = torch.from_numpy(F).to(self.device)
        self.tris = torch.from_numpy(np.ascontiguousarray(tris)).to(self.device)
        self.cdf = torch.from_numpy(_mesh.face_cdf(tris)).to(self.device)
        self.nf = tris.shape[0]
        lib = _lib.load()
        self._ws = torch.empty(max(1, lib.nglod_mesh2sdf_workspace_size(self.nf)), dtype=torch.uint8,
                               device=self.device)

    def sdf(self, pts, with_f64=False):
        """Signed distance of pts [N, 3] (FP64 arithmetic); [N] fp32 (and fp64)."""
        _lib.require_gpu(pts)
        p = pts.detach().float().contiguous()
        n = p.shape[0]
        out = torch.empty(n, dtype=torch.float32, device=p.device)
        out64 = torch.empty(n, dtype=torch.float64, device=p.device) if with_f64 else None
        lib = _lib.load()
        _lib.check(lib.nglod_mesh2sdf_f64(_lib.ptr(p), n, _lib.ptr(self.tris), self.nf, _lib.ptr(out),
                                          _lib.ptr(out64), _lib.ptr(self._ws), self._ws.numel(),
                                          _lib.stream_of(p)), "nglod_mesh2sdf_f64")
        return (out, out64) if with_f64 else out

    def sample_uniform(self, n, seed):
        out = torch.empty(n, 3, dtype=torch.float32, device=self.device)
        _lib.check(_lib.load().nglod_sample_uniform(n, seed, _lib.ptr(out), _lib.stream_of(out)),
                   "nglod_sample_uniform")
        return out

    def sample_surface(self, n, seed, sigma=0.0, return_faces=False):
        out = torch.empty(n, 3, dtype=torch.float32, device=self.device)
        face = torch.empty(n, dtype=torch.int32, device=self.device) if return_faces else None
        _lib.check(_lib.load().nglod_sample_surface(_lib.ptr(self.tris), _lib.ptr(self.cdf), self.nf, n, seed,
                                                    float(sigma), _lib.ptr(out), _lib.ptr(face),
                                                    _lib.stream_of(out)), "nglod_sample_surface")
        return (out, face) if return_faces else out

    def training_batch(self, n, seed, sigmas=(0.01, 0.003)):
        """n points: 1/3 uniform in [-1,1]^3, 2/3 near the surface split evenly
        over `sigmas`, with their FP64 signed distances.  Returns (pts, sdf [n, 1])."""
        nu = (n + 2) // 3
        rest = n - nu
        parts = [self.sample_uniform(nu, seed * 16 + 1)]
        k = len(sigmas)
        for j, s in enumerate(sigmas):
            m = rest // k + (1 if j < rest % k else 0)
            parts.append(self.sample_surface(m, seed * 16 + 2 + j, sigma=s))
        pts = torch.cat(parts, 0)
        return pts, self.sdf(pts)[:, None]


def compute_sdf(V, F, points):
    """torchgp.compute_sdf(V, F, points) -> [N] signed distances."""
    md = MeshData(V.detach().cpu().numpy(), F.detach().cpu().numpy(), points.device)
    return md.sdf(points)


def point_sample(V, F, techniques, num_samples, seed=0):
    """torchgp.point_sample: num_samples per technique in 'rand', 'near'
    (sigma 0.01), 'trace' (on the surface), concatenated in order."""
    md = MeshData(V.detach().cpu().numpy(), F.detach().cpu().numpy(), V.device)
    out = []
    for j, t in enumerate(techniques):
        s = seed * 1024 + j
        if t == "rand":
            out.append(md.sample_uniform(num_samples, s))
        elif t == "near":
            out.append(md.sample_surface(num_samples, s, sigma=0.01))
        elif t == "trace":
            out.append(md.sample_surface(num_samples, s, sigma=0.0))
        else:
            raise ValueError(f"unknown sampling technique {t!r}")
    return torch.cat(out, 0)
