"""Restatement of the reference's PyTorch path in plain torch ops, for the CPU.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline; see oracle/oracle.py).

It keeps the reference's operation structure, including its costs: per-LOD
F.grid_sample (sdf-net/lib/models/OctreeSDF.py:46-57), every level from 0 up
to `lod` decoded, each decoder applied twice (OctreeSDF.py:124, :142), the
SphereTracer loop with a global any() test and boolean-mask compaction per
step (sdf-net/lib/tracer/SphereTracer.py:74-115), central-difference normals
(sdf-net/lib/diffutils.py:61-70).  The box entry is the reference's CUDA
kernel (sol_nglod_kernel.cu:77-159), which cannot run on a CPU; it is taken
from the C oracle (oracle.aabb), the same arithmetic.
"""
import numpy as np
import torch
import torch.nn.functional as F

from . import oracle


class TorchOctreeSDF:
    """Parameters as tensors in the reference layout: fm [1, C, S, S, S]."""

    def __init__(self, fms, w0, b0, w1, b1, pos_invariant, dtype=torch.float32):
        self.fms = [f.to(dtype).contiguous() for f in fms]
        self.w0 = [w.to(dtype) for w in w0]
        self.b0 = [b.to(dtype) for b in b0]
        self.w1 = [w.to(dtype).reshape(1, -1) for w in w1]
        self.b1 = [b.to(dtype).reshape(-1) for b in b1]
        self.pos_invariant = pos_invariant
        self.num_lods = len(self.fms)
        self.dtype = dtype
        self.lod = None

    @classmethod
    def from_module(cls, net, dtype=torch.float32):
        fms = [f.fm.detach().cpu().contiguous() for f in net.features]
        L = net.louts
        return cls(fms, [l[0].weight.detach().cpu() for l in L], [l[0].bias.detach().cpu() for l in L],
                   [l[2].weight.detach().cpu() for l in L], [l[2].bias.detach().cpu() for l in L],
                   net.pos_invariant, dtype)

    def volume(self, i, x):
        n = x.shape[0]
        g = F.grid_sample(self.fms[i], x.reshape(1, n, 1, 1, 3), align_corners=True, padding_mode="border")
        return g[0, :, :, 0, 0].transpose(0, 1)

    def decoder(self, i, e):
        di = 0 if len(self.w0) == 1 else i
        h = F.relu(F.linear(e, self.w0[di], self.b0[di]))
        return F.linear(h, self.w1[di], self.b1[di])

    def sdf(self, x, lod=None, return_lst=False):
        if lod is None:
            lod = self.lod
        preds = []
        feats = None
        for i in range(self.num_lods):
            s = self.volume(i, x)
            feats = s if feats is None else s + feats
            e = feats if self.pos_invariant else torch.cat([x, feats], dim=-1)
            d = self.decoder(i, e)      # first application (OctreeSDF.py:124)
            d = self.decoder(i, e)      # second application (OctreeSDF.py:142)
            if lod is not None and lod == i:
                return d
            preds.append(d)
        return preds if return_lst else preds[-1]

    def __call__(self, x):
        return self.sdf(x)


def fd_gradient(x, f, h=1.0 / (64.0 * 3.0)):
    ex = torch.tensor([h, 0.0, 0.0], dtype=x.dtype)
    ey = torch.tensor([0.0, h, 0.0], dtype=x.dtype)
    ez = torch.tensor([0.0, 0.0, h], dtype=x.dtype)
    g = torch.cat([f(x + ex) - f(x - ex), f(x + ey) - f(x - ey), f(x + ez) - f(x - ez)], dim=-1)
    return g / (h * 2.0)


def aabb(ray_o, ray_d):
    prec = "f32" if ray_o.dtype == torch.float32 else "f64"
    x, t, hit = oracle.aabb(ray_o.numpy(), ray_d.numpy(), prec)
    return torch.from_numpy(x), torch.from_numpy(t).reshape(-1, 1), torch.from_numpy(hit)


def sphere_trace(net, ray_o, ray_d, num_steps=256, min_dis=3e-4, step_size=1.0, far_clip=10.0):
    """SphereTracer.forward, batched with masks as in the reference; `steps`
    counts each ray's in-loop SDF evaluations."""
    x, t, cond = aabb(ray_o, ray_d)
    with torch.no_grad():
        d = net(x)
        dprev = d.clone()
        hit = torch.zeros_like(d, dtype=torch.bool)[:, 0]
        steps = torch.zeros(x.shape[0], dtype=torch.int32)
        for _ in range(num_steps):
            hit = (torch.abs(t) < far_clip)[:, 0]
            cond = cond & (torch.abs(d) > min_dis)[:, 0]
            cond = cond & (torch.abs((d + dprev) / 2.0) > min_dis * 3)[:, 0]
            cond = cond & hit
            if not cond.any():
                break
            x = torch.where(cond.view(-1, 1), torch.addcmul(ray_o, ray_d, t), x)
            dprev = torch.where(cond.unsqueeze(1), d, dprev)
            d[cond] = net(x[cond]) * step_size
            steps += cond.int()
            t = torch.where(cond.view(-1, 1), t + d, t)
        hit = hit & ~(torch.abs(x) > 1.0).any(dim=-1)
        normal = torch.zeros_like(x)
        if hit.any():
            g = fd_gradient(x[hit], net)
            normal[hit] = F.normalize(g, p=2, dim=-1, eps=1e-5)
    return dict(x=x, t=t, hit=hit, normal=normal, steps=steps)
