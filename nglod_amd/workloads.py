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


def _telescoped_levels(field_fn, resolutions):
    """Per-level corner values whose level sums telescope: level i holds
    F - I_{i-1}(F) at its corners, where I_{i-1} is the trilinear interpolant
    of F on level i-1.  A trilinear function on a coarse cell is trilinear on
    every finer sub-cell, so sum_{i<=L} interp_i(level_i) = I_L(F): the model
    at lod L is the level-L interpolant of F."""
    import torch.nn.functional as F

    levels, prev = [], None
    for r in resolutions:
        full = field_fn(r).double()  # [z, y, x]
        if prev is None:
            levels.append(full)
        else:
            c = torch.linspace(-1.0, 1.0, r + 1, dtype=torch.float64)
            z, y, x = torch.meshgrid(c, c, c, indexing="ij")
            grid = torch.stack([x, y, z], -1).reshape(1, -1, 1, 1, 3)
            coarse = F.grid_sample(prev[None, None], grid, align_corners=True, padding_mode="border")
            levels.append(full - coarse.reshape(r + 1, r + 1, r + 1))
        prev = full
    return [l.float() for l in levels]


def make_sphere_model(radius=0.6, num_lods=5, base_lod=2, pos_invariant=True, seed=0, device="cpu",
                      noise=1e-3):
    """A model whose SDF at lod L is, up to `noise`-scale perturbations, the
    trilinear interpolant of a sphere's SDF on the level-L grid.  Channel 0 of
    the volumes holds telescoped sphere values (see _telescoped_levels) and
    decoder row 0 passes channel 0 through the ReLU with a +2 offset.  The
    other channels and rows keep small seeded random values, so every weight
    still takes part in the arithmetic."""
    net = make_model(num_lods=num_lods, base_lod=base_lod, pos_invariant=pos_invariant, seed=seed, device="cpu")
    g = torch.Generator().manual_seed(seed + 1000)
    off = 0 if pos_invariant else 3
    levels = _telescoped_levels(lambda r: sphere_field(r, radius), [f.fsize for f in net.features])
    with torch.no_grad():
        for f, lv in zip(net.features, levels):
            f.fm[0, 0].copy_(lv)
        for l in net.louts:
            w0 = l[0].weight
            w0.mul_(noise)
            w0[0].zero_()
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
