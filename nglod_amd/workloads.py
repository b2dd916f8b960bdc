"""Seeded models, rays and query points for the benchmark and test workloads.

Models follow the benchmark description of the path: 5 LODs of dense
feature grids at r = 4, 8, 16, 32, 64 ((r+1)^3 corners), 32 channels drawn
from N(0, 0.01), one 32(+3) -> 128 -> 1 ReLU decoder per LOD with
Xavier-uniform weights and zero biases, all from a torch.Generator seed.
"""
import math

import torch

from .octree_sdf import OctreeSDF, default_args
from .renderer import look_at


def _xavier_(w, g):
    fan_out, fan_in = w.shape
    bound = math.sqrt(6.0 / (fan_in + fan_out))
    w.copy_((torch.rand(w.shape, generator=g) * 2.0 - 1.0) * bound)


def make_model(num_lods=5, base_lod=2, feature_dim=32, hidden_dim=128, pos_invariant=True, joint_decoder=False,
               seed=0, device="cuda"):
    """Random-init OctreeSDF: fm ~ N(0, 0.01), Xavier decoders, from `seed`."""
    args = default_args(num_lods=num_lods, base_lod=base_lod, feature_dim=feature_dim, hidden_dim=hidden_dim,
                        pos_invariant=pos_invariant, joint_decoder=joint_decoder)
    net = OctreeSDF(args)
    g = torch.Generator().manual_seed(seed)
    with torch.no_grad():
        for f in net.features:
            f.fm.copy_(torch.randn(f.fm.shape, generator=g) * 0.01)
        for l in net.louts:
            _xavier_(l[0].weight, g)
            l[0].bias.zero_()
            _xavier_(l[2].weight, g)
            l[2].bias.zero_()
    return net.to(device)


def sphere_field(res, radius):
    """SDF of a sphere of `radius` at the (res+1)^3 grid corners, [z, y, x]."""
    c = torch.linspace(-1.0, 1.0, res + 1, dtype=torch.float64)
    z, y, x = torch.meshgrid(c, c, c, indexing="ij")
    return (torch.sqrt(x * x + y * y + z * z) - radius).float()


def make_sphere_model(radius=0.6, num_lods=5, base_lod=2, pos_invariant=True, seed=0, device="cuda",
                      noise=1e-3):
    """A model whose SDF is (up to `noise`-scale perturbations) the trilinear
    interpolant of a sphere's SDF on the finest grid: channel 0 of the finest
    volume holds the sphere field, channel 0 of the other volumes is zero, and
    decoder row 0 passes channel 0 through the ReLU with a +2 offset.  The
    remaining channels and rows keep small seeded random values, so every
    weight still takes part in the arithmetic."""
    net = make_model(num_lods=num_lods, base_lod=base_lod, pos_invariant=pos_invariant, seed=seed, device="cpu")
    g = torch.Generator().manual_seed(seed + 1000)
    off = 0 if pos_invariant else 3
    with torch.no_grad():
        for i, f in enumerate(net.features):
            f.fm[0, 0].zero_()
            if i == num_lods - 1:
                f.fm[0, 0].copy_(sphere_field(f.fsize, radius))
        for l in net.louts:
            w0 = l[0].weight
            w0.mul_(noise)
            w0[0].zero_()
            w0[0, off] = 1.0
            w0[:, off].zero_()
            w0[0, off] = 1.0
            l[0].bias.copy_(torch.rand(l[0].bias.shape, generator=g) * noise)
            l[0].bias[0] = 2.0
            w1 = l[2].weight
            w1.mul_(noise)
            w1[0, 0] = 1.0
            l[2].bias.fill_(-2.0)
    return net.to(device)


def camera_rays(width, height, eye=(0.0, 0.0, 3.0), at=(0.0, 0.0, 0.0), fov=30.0, device="cuda"):
    """Pinhole rays, no jitter, reference layout (ray index = ix * height + iy)."""
    return look_at(list(eye), list(at), width, height, mode="persp", fov=fov, device=device)


def uniform_points(n, seed=1, device="cuda"):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(n, 3, generator=g) * 2.0 - 1.0).to(device)


BENCH_MODEL_NAME = "5-LOD OctreeSDF (r 4..64, C 32, 32->128->1), sphere field r=0.6 in the finest grid + seeded noise"


def bench_model(device="cuda"):
    """The model the frame benchmark traces."""
    return make_sphere_model(radius=0.6, num_lods=5, device=device)
