"""The committed golden fixtures (tests/golden/, written by make_golden.py from
torch's grid_sample/Linear in float64 and exact rational geometry) against
the oracle on the CPU and the HIP path on the GPU."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _query_arrays(q):
    n = len(q["res"])
    grids = [q[f"fm{i}"][0].transpose(1, 2, 3, 0) for i in range(n)]
    return oracle.ModelArrays(grids, [int(r) for r in q["res"]], [q[f"w0_{i}"] for i in range(n)],
                              [q[f"b0_{i}"] for i in range(n)], [q[f"w1_{i}"] for i in range(n)],
                              [q[f"b1_{i}"] for i in range(n)], pos_invariant=False)


def _nugget_cells(pts, pyr, level, nug):
    lv = pts[pyr[1, level]:pyr[1, level] + pyr[0, level], :3].astype(np.int64)
    return nug[:, 0].astype(np.int64), lv[nug[:, 1]]


# ---- CPU: the oracle -----------------------------------------------------------
def test_oracle_query_golden():
    q = _load("query.npz")
    m = _query_arrays(q)
    for lod in range(len(q["res"])):
        np.testing.assert_allclose(oracle.sdf(m, q["x"], lod, "f64"), q["sdf"][lod], rtol=0, atol=1e-12)
        np.testing.assert_allclose(oracle.sdf(m, q["x"], lod, "f32"), q["sdf"][lod], rtol=0, atol=1e-5)


def test_oracle_aabb_golden():
    a = _load("aabb.npz")
    _, t, hit = oracle.aabb(a["o"], a["d"], "f64")
    np.testing.assert_array_equal(hit, a["hit"])
    np.testing.assert_allclose(t[a["hit"]], a["t"][a["hit"]], rtol=0, atol=1e-12)
    _, t32, hit32 = oracle.aabb(a["o"], a["d"], "f32")
    np.testing.assert_array_equal(hit32, a["hit"])
    np.testing.assert_allclose(t32[a["hit"]], a["t"][a["hit"]], rtol=0, atol=1e-6)


def test_oracle_mesh2sdf_golden():
    g = _load("mesh2sdf.npz")
    np.testing.assert_allclose(oracle.mesh2sdf(g["points"], g["tris"]), g["dist"], rtol=0, atol=1e-12)


def test_oracle_nuggets_golden():
    g = _load("nuggets.npz")
    level = int(g["level"])
    c = g["cells"]
    codes = np.unique(np.array([oracle.morton(int(a), int(b), int(z)) for a, b, z in c], dtype=np.uint64))
    octree, _ = oracle.octree_from_codes(codes, level)
    pts, pyr, _ = oracle.octree_points(octree)
    nug = oracle.spc_raytrace(octree, pts, pyr, level, g["o"], g["d"], level)
    ray, cell = _nugget_cells(pts, pyr, level, nug)
    np.testing.assert_array_equal(ray, g["ray"])
    np.testing.assert_array_equal(cell, g["cell"])


# ---- GPU: the HIP path -----------------------------------------------------------
@pytest.mark.gpu
def test_hip_query_golden(dev):
    from nglod_amd import OctreeSDF
    q = _load("query.npz")
    n = len(q["res"])
    net = OctreeSDF(num_lods=n, base_lod=2, feature_dim=32, hidden_dim=128, pos_invariant=False)
    with torch.no_grad():
        for i in range(n):
            net.features[i].fm.copy_(torch.from_numpy(q[f"fm{i}"]))
            net.louts[i][0].weight.copy_(torch.from_numpy(q[f"w0_{i}"]))
            net.louts[i][0].bias.copy_(torch.from_numpy(q[f"b0_{i}"]))
            net.louts[i][2].weight.copy_(torch.from_numpy(q[f"w1_{i}"]))
            net.louts[i][2].bias.copy_(torch.from_numpy(q[f"b1_{i}"]))
    net = net.to(dev)
    x = torch.from_numpy(q["x"]).to(dev)
    with torch.no_grad():
        for lod in range(n):
            got = net.sdf(x, lod=lod)[:, 0].cpu().numpy()
            np.testing.assert_allclose(got, q["sdf"][lod], rtol=0, atol=1e-5)


@pytest.mark.gpu
def test_hip_aabb_golden(dev):
    from nglod_amd import _lib
    a = _load("aabb.npz")
    o = torch.from_numpy(a["o"]).to(dev)
    d = torch.from_numpy(a["d"]).to(dev)
    n = o.shape[0]
    x = torch.empty(n, 3, device=dev)
    t = torch.empty(n, device=dev)
    hit = torch.empty(n, dtype=torch.uint8, device=dev)
    _lib.check(_lib.load().nglod_aabb(_lib.ptr(o), _lib.ptr(d), n, _lib.ptr(x), _lib.ptr(t), _lib.ptr(hit),
                                      _lib.stream_of(o)), "aabb")
    h = hit.cpu().numpy().astype(bool)
    np.testing.assert_array_equal(h, a["hit"])
    np.testing.assert_allclose(t.cpu().numpy()[h], a["t"][h], rtol=0, atol=1e-6)


@pytest.mark.gpu
def test_hip_mesh2sdf_golden(dev):
    from nglod_amd import _lib
    g = _load("mesh2sdf.npz")
    p = torch.from_numpy(g["points"]).to(dev)
    m = torch.from_numpy(g["tris"]).to(dev)
    n, nf = p.shape[0], m.shape[0]
    lib = _lib.load()
    dist = torch.empty(n, device=dev)
    dist64 = torch.empty(n, dtype=torch.float64, device=dev)
    ws = torch.empty(lib.nglod_mesh2sdf_workspace_size(nf), dtype=torch.uint8, device=dev)
    _lib.check(lib.nglod_mesh2sdf_f64(_lib.ptr(p), n, _lib.ptr(m), nf, _lib.ptr(dist), _lib.ptr(dist64),
                                      _lib.ptr(ws), ws.numel(), _lib.stream_of(p)), "mesh2sdf")
    np.testing.assert_allclose(dist64.cpu().numpy(), g["dist"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(dist.cpu().numpy(), g["dist"], rtol=0, atol=1e-6)


@pytest.mark.gpu
def test_hip_nuggets_golden(dev):
    from nglod_amd.spc import SPC
    g = _load("nuggets.npz")
    level = int(g["level"])
    c = g["cells"]
    codes = np.unique(np.array([oracle.morton(int(a), int(b), int(z)) for a, b, z in c], dtype=np.uint64))
    spc = SPC.from_codes(torch.from_numpy(codes.astype(np.int64)).to(dev), level, dev)
    nug, offs = spc.raytrace(torch.from_numpy(g["o"]).to(dev), torch.from_numpy(g["d"]).to(dev), level)
    ray, cell = _nugget_cells(spc.points.cpu().numpy(), spc.pyramid.numpy(), level, nug.cpu().numpy())
    np.testing.assert_array_equal(ray, g["ray"])
    np.testing.assert_array_equal(cell, g["cell"])


@pytest.mark.gpu
def test_dropin_sol_nglod_aabb_golden(dev):
    """nglod_amd.sol_nglod.aabb has the reference extension's signature and
    return shapes (sol_nglod_kernel.cu:161-188: [x [N,3], t [N,1], hit [N]
    bool], x = origin where there is no hit) and matches the fixture and the
    FP32 oracle bit for bit."""
    from nglod_amd.sol_nglod import aabb
    a = _load("aabb.npz")
    o = torch.from_numpy(a["o"]).to(dev)
    d = torch.from_numpy(a["d"]).to(dev)
    out = aabb(o, d)
    assert isinstance(out, list) and len(out) == 3
    x, t, hit = out
    assert x.shape == (o.shape[0], 3) and x.dtype == torch.float32
    assert t.shape == (o.shape[0], 1) and t.dtype == torch.float32
    assert hit.shape == (o.shape[0],) and hit.dtype == torch.bool
    h = hit.cpu().numpy()
    np.testing.assert_array_equal(h, a["hit"])
    np.testing.assert_allclose(t[:, 0].cpu().numpy()[h], a["t"][h], rtol=0, atol=1e-6)
    xo, to, ho = oracle.aabb(a["o"], a["d"], "f32")
    np.testing.assert_array_equal(x.cpu().numpy(), xo)
    np.testing.assert_array_equal(t[:, 0].cpu().numpy(), to)
    np.testing.assert_array_equal(x.cpu().numpy()[~h], a["o"][~h])
    with pytest.raises(TypeError):
        aabb(o.double(), d.double())
    e = aabb(o[:0], d[:0])
    assert e[0].shape == (0, 3) and e[1].shape == (0, 1) and e[2].shape == (0,)


@pytest.mark.gpu
def test_dropin_mesh2sdf_gpu_golden(dev):
    """nglod_amd.mesh2sdf.mesh2sdf_gpu(points, mesh) -> [dist [N] f32] as
    mesh2sdf_kernel.cu:895-927 returns it, against the FP64 fixture."""
    from nglod_amd import mesh2sdf
    g = _load("mesh2sdf.npz")
    out = mesh2sdf.mesh2sdf_gpu(torch.from_numpy(g["points"]).to(dev), torch.from_numpy(g["tris"]).to(dev))
    assert isinstance(out, list) and len(out) == 1
    dist = out[0]
    assert dist.shape == (g["points"].shape[0],) and dist.dtype == torch.float32
    np.testing.assert_allclose(dist.cpu().numpy(), g["dist"], rtol=0, atol=1e-6)
    np.testing.assert_array_equal(np.sign(dist.cpu().numpy()), np.sign(g["dist"].astype(np.float32)))
