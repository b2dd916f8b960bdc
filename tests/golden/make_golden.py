"""Writes the golden fixtures under tests/golden/ (run from the repo root:
`python tests/golden/make_golden.py`).

The reference holds no golden vectors, so these come from independent
sources, never from the code under test:
  query.npz     a 2-LOD OctreeSDF's parameters, query points (inside, on grid
                vertices, outside the box) and its sdf per LOD computed by the
                very library calls the reference makes
                (torch.nn.functional.grid_sample, align_corners=True,
                padding_mode='border', then nn.Linear / ReLU; OctreeSDF.py:46-155)
                in float64 on the CPU.
  aabb.npz      rays vs [-1,1]^3: entry distance and hit flag from an exact
                rational slab test (fractions.Fraction on the float32 inputs);
                no ray enters exactly through an edge or corner, where the
                reference's strict face tests report a miss.
  mesh2sdf.npz  points vs a small closed mesh (octahedron): unsigned distance
                as the minimum over triangles of the plane projection (when it
                falls inside) and the three edge segments, sign by the analytic
                inside test of the convex solid.
  nuggets.npz   a level-4 octree (cells overlapping a sphere shell, decided
                in float64) and rays: each ray's hit cells, in entry order,
                from an exact rational slab test per cell.
"""
import os
import sys
from fractions import Fraction

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def query_fixture():
    g = torch.Generator().manual_seed(1234)
    res = [4, 8]
    fms = [torch.randn(1, 32, r + 1, r + 1, r + 1, generator=g, dtype=torch.float64).float() * 0.1 for r in res]
    w0 = [(torch.rand(128, 35, generator=g, dtype=torch.float64) * 2 - 1).float() * 0.3 for _ in res]
    b0 = [(torch.rand(128, generator=g, dtype=torch.float64) * 2 - 1).float() * 0.1 for _ in res]
    w1 = [(torch.rand(1, 128, generator=g, dtype=torch.float64) * 2 - 1).float() * 0.1 for _ in res]
    b1 = [(torch.rand(1, generator=g, dtype=torch.float64) * 2 - 1).float() * 0.1 for _ in res]
    rng = np.random.default_rng(7)
    inside = rng.uniform(-1, 1, (200, 3))
    verts = -1 + 2 * rng.integers(0, 9, (40, 3)) / 8.0  # vertices of the res-8 grid
    outside = rng.uniform(-1.6, 1.6, (60, 3))
    x = np.concatenate([inside, verts, outside]).astype(np.float32)
    xt = torch.from_numpy(x).double()
    s = torch.zeros(x.shape[0], 32, dtype=torch.float64)
    out = []
    for i in range(len(res)):
        fm = fms[i].double()
        f = torch.nn.functional.grid_sample(fm, xt.view(1, -1, 1, 1, 3), mode="bilinear", padding_mode="border",
                                            align_corners=True)[0, :, :, 0, 0].t()
        s = s + f
        e = torch.cat([xt, s], dim=-1)
        h = torch.relu(e @ w0[i].double().t() + b0[i].double())
        out.append((h @ w1[i].double().t() + b1[i].double())[:, 0].numpy())
    arr = {"x": x, "sdf": np.stack(out), "res": np.array(res)}
    for i in range(len(res)):
        arr[f"fm{i}"] = fms[i].numpy()
        arr[f"w0_{i}"] = w0[i].numpy()
        arr[f"b0_{i}"] = b0[i].numpy()
        arr[f"w1_{i}"] = w1[i].numpy()
        arr[f"b1_{i}"] = b1[i].numpy()
    return arr


def _slab_exact(o, d, lo, hi):
    """Exact entry distance of the ray o + t d (t >= 0) into the box [lo, hi]
    (closed box), or None."""
    t0, t1 = Fraction(0), None
    for k in range(3):
        ok, dk = Fraction(float(o[k])), Fraction(float(d[k]))
        a, b = Fraction(float(lo[k])), Fraction(float(hi[k]))
        if dk == 0:
            if ok < a or ok > b:
                return None
            continue
        ta, tb = (a - ok) / dk, (b - ok) / dk
        if ta > tb:
            ta, tb = tb, ta
        t0 = max(t0, ta)
        t1 = tb if t1 is None else min(t1, tb)
    if t1 is not None and t0 > t1:
        return None
    return t0


def aabb_fixture():
    rng = np.random.default_rng(11)
    v = rng.normal(size=(160, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    o = v * rng.uniform(1.8, 4.0, (160, 1))
    tgt = rng.uniform(-1.3, 1.3, (160, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    axis = np.array([[0, 0, 3, 0, 0, -1], [3, 0, 0, -1, 0, 0], [0, -3, 0, 0, 1, 0], [2, 0.5, 0.25, -1, 0, 0],
                     [0.3, 0.2, 4, 0, 0, -1], [5, 5, 5, -1, -1.02, -0.98], [0, 2, 0, 0, 1, 0]], dtype=np.float64)
    o = np.concatenate([o, axis[:, :3]]).astype(np.float32)
    d = np.concatenate([d, axis[:, 3:] / np.linalg.norm(axis[:, 3:], axis=1, keepdims=True)]).astype(np.float32)
    t = np.zeros(o.shape[0])
    hit = np.zeros(o.shape[0], dtype=bool)
    for i in range(o.shape[0]):
        if np.all(np.abs(o[i]) <= 1):
            continue  # origins inside are skipped by the reference kernel
        te = _slab_exact(o[i], d[i], (-1, -1, -1), (1, 1, 1))
        if te is not None:
            hit[i] = True
            t[i] = float(te)
    return {"o": o, "d": d, "t": t, "hit": hit}


def _tri_dist(p, a, b, c):
    """Exact-projection point-triangle distance (float64): the minimum over
    the plane projection (when inside) and the three edge segments."""
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    # barycentric inside test of the projection
    def same_side(p1, p2, u, v):
        return np.dot(np.cross(v - u, p1 - u), np.cross(v - u, p2 - u)) >= 0
    best = np.inf
    if same_side(q, a, b, c) and same_side(q, b, a, c) and same_side(q, c, a, b):
        best = abs(np.dot(p - a, n))
    for u, v in ((a, b), (b, c), (c, a)):
        e = v - u
        s = np.clip(np.dot(p - u, e) / np.dot(e, e), 0.0, 1.0)
        best = min(best, np.linalg.norm(p - (u + s * e)))
    return best


def mesh_fixture():
    V = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64) * 0.7
    F = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    tris = V[F].astype(np.float32)
    rng = np.random.default_rng(5)
    pts = rng.uniform(-1.2, 1.2, (300, 3)).astype(np.float32)
    dist = np.empty(pts.shape[0])
    T = tris.astype(np.float64)
    for i, p in enumerate(pts.astype(np.float64)):
        u = min(_tri_dist(p, *T[f]) for f in range(T.shape[0]))
        inside = np.abs(p).sum() < 0.7  # the octahedron |x|+|y|+|z| <= 0.7
        dist[i] = -u if inside else u
    return {"tris": tris, "points": pts, "dist": dist}


def nugget_fixture():
    level = 4
    R = 1 << level
    cells = []
    for x in range(R):
        for y in range(R):
            for z in range(R):
                lo = np.array([x, y, z]) / R * 2 - 1
                hi = lo + 2.0 / R
                near = np.linalg.norm(np.clip(0.0, lo, hi))
                far = np.linalg.norm(np.maximum(np.abs(lo), np.abs(hi)))
                if near <= 0.6 <= far:
                    cells.append((x, y, z))
    cells = np.array(cells, dtype=np.int64)
    rng = np.random.default_rng(21)
    v = rng.normal(size=(48, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    o = (v * rng.uniform(2.0, 3.5, (48, 1))).astype(np.float32)
    tgt = rng.uniform(-0.7, 0.7, (48, 3))
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays, hits = [], []
    for i in range(o.shape[0]):
        entries = []
        for c in cells:
            lo = c / R * 2 - 1
            te = _slab_exact(o[i], d[i], lo, lo + 2.0 / R)
            if te is not None:
                entries.append((te, tuple(c)))
        entries.sort()
        for te, c in entries:
            rays.append(i)
            hits.append(c)
    return {"level": np.array(level), "cells": cells, "o": o, "d": d, "ray": np.array(rays, dtype=np.int64),
            "cell": np.array(hits, dtype=np.int64).reshape(-1, 3)}


def main():
    np.savez_compressed(os.path.join(HERE, "query.npz"), **query_fixture())
    np.savez_compressed(os.path.join(HERE, "aabb.npz"), **aabb_fixture())
    np.savez_compressed(os.path.join(HERE, "mesh2sdf.npz"), **mesh_fixture())
    np.savez_compressed(os.path.join(HERE, "nuggets.npz"), **nugget_fixture())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    sys.exit(main())
