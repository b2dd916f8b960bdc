"""Shared builders for the tests (seeded, CPU-side)."""
import numpy as np
import torch

from nglod_amd import workloads
from oracle import oracle


def small_model(num_lods=3, pos_invariant=True, joint=False, seed=0):
    """Random-init OctreeSDF on the CPU (grids 4, 8, 16 by default)."""
    return workloads.make_model(num_lods=num_lods, pos_invariant=pos_invariant, joint_decoder=joint, seed=seed,
                                device="cpu")


def rand_points(n, seed, lo=-1.0, hi=1.0):
    g = np.random.default_rng(seed)
    return g.uniform(lo, hi, size=(n, 3)).astype(np.float32)


def arrays(net):
    return oracle.ModelArrays.from_module(net)


def rand_rays(n, seed, dist=(1.5, 4.0)):
    """Rays from a shell around the box, aimed near the origin (random)."""
    g = np.random.default_rng(seed)
    v = g.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    r = g.uniform(*dist, size=(n, 1))
    o = v * r
    target = g.uniform(-0.8, 0.8, size=(n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o.astype(np.float32), d.astype(np.float32)


def ambiguous_rays(orc, other_steps, num_steps, margin=2e-6, cap_is_ambiguous=True):
    """Threshold-ambiguous rays of a trace comparison: a termination input of
    the oracle within `margin` of its threshold, a different number of steps
    in the two implementations (a flipped termination decision), or the step
    cap reached (an unconverged, grazing trajectory whose end point is not a
    surface point)."""
    amb = (orc["margin"] < margin) | (orc["steps"] != other_steps)
    if cap_is_ambiguous:
        amb |= orc["steps"] >= num_steps
    return amb
