"""GPU parity of the sparse (SOL) path against the oracle: corner lists,
corner features and trinkets bit-exact; the sparse SDF and the sparse march
(S semantics: 50 steps, 3e-4 / 1.5e-3 stops, per-cell re-entry, normals in
the hit cell) within FP32 tolerance; the npz export round trip; the GPU
occupancy of traced surface points."""
import numpy as np
import pytest
import torch

from nglod_amd import SphereTracer, mesh, workloads
from nglod_amd.sol import NPZ_KEYS, SolNGLOD, voxel_sparsify
from nglod_amd.spc import SPC
from oracle import oracle
from tests.helpers import arrays, sol_arrays

pytestmark = pytest.mark.gpu

RADIUS = 0.6


def _sphere_tris(level=4):
    V, F = mesh.icosphere(level)
    return mesh.triangles(V * RADIUS, F)


@pytest.fixture(scope="module")
def sphere(dev):
    """Sphere model (5 LODs, xyz-first decoders) and its level-6 octree."""
    net = workloads.make_sphere_model(radius=RADIUS, num_lods=5, pos_invariant=False, device="cpu")
    m = arrays(net)
    spc = SPC.from_mesh(_sphere_tris(), 6, dev)
    return net.to(dev), m, spc


def test_sample_directions(dev):
    from nglod_amd import _lib
    n = 100_000
    out = torch.empty(n, 3, device=dev)
    _lib.check(_lib.load().nglod_sample_directions(n, 11, _lib.ptr(out), _lib.stream_of(out)), "dirs")
    ref = oracle.sample_directions(11, n)
    g = out.cpu().numpy()
    np.testing.assert_allclose(g, ref, rtol=0, atol=2e-7)
    np.testing.assert_allclose(np.linalg.norm(g.astype(np.float64), axis=1), 1.0, atol=1e-6)


def test_codes_from_points(dev):
    import ctypes
    from nglod_amd import _lib
    from nglod_amd.spc import _call_with_workspace
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (20000, 3)).astype(np.float32)
    x[:8] = np.array([[1, 1, 1], [-1, -1, -1], [1, -1, 0.5], [0, 0, 0], [0.5, 0.5, 0.5], [-0.5, 1, -1],
                      [np.nextafter(np.float32(1), np.float32(0))] * 3, [0.25, -0.25, 0.125]], dtype=np.float32)
    mask = (rng.uniform(size=20000) < 0.7).astype(np.uint8)
    mask[:8] = 1
    level = 6
    xt = torch.from_numpy(x).to(dev)
    mt = torch.from_numpy(mask).to(dev)
    codes = torch.empty(20000, dtype=torch.int32, device=dev)
    count = ctypes.c_int64(0)
    lib = _lib.load()
    s = _lib.stream_of(xt)
    rc = _call_with_workspace(lambda *a: lib.nglod_spc_codes_from_points(*a, s), dev, _lib.ptr(xt), _lib.ptr(mt),
                              20000, level, _lib.ptr(codes), 20000, ctypes.byref(count))
    _lib.check(rc, "codes")
    got = codes[:count.value].cpu().numpy().astype(np.uint32).astype(np.uint64)
    ref = oracle.codes_from_points(x[mask.astype(bool)], level)
    np.testing.assert_array_equal(got, ref)


def test_corners_features_trinkets_bit_exact(sphere):
    net, m, spc = sphere
    sol = SolNGLOD.from_net(net, spc, fp16=False)
    pts = spc.points.cpu().numpy()
    pyr = spc.pyramid.numpy()
    sol_ref, blocks, _ = sol_arrays(m, spc.octree.cpu().numpy(), 4)
    np.testing.assert_array_equal(oracle.octree_points(spc.octree.cpu().numpy())[0], pts)
    cc = sol.cc.cpu().numpy()
    assert sol.counts == [b.shape[0] for b in blocks]
    np.testing.assert_array_equal(cc, np.concatenate(blocks))
    np.testing.assert_array_equal(sol.cf.cpu().numpy(), sol_ref.cf)
    beg, end = pyr[1, 2], pyr[1, 6] + pyr[0, 6]
    tv = sol.trinket_v.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(tv[beg:end], sol_ref.tv[beg:end])
    np.testing.assert_array_equal(sol.parent.cpu().numpy()[beg:end], sol_ref.parent[beg:end])


def test_trinkets_reject_missing_corners(sphere, dev):
    net, m, spc = sphere
    sol = SolNGLOD.from_net(net, spc, fp16=False)
    cc = sol.cc.clone()
    counts = list(sol.counts)
    # drop the last corner of the level-2 block: a trinket corner is then missing
    keep = torch.ones(cc.shape[0], dtype=torch.bool, device=dev)
    keep[counts[0] - 1] = False
    counts[0] -= 1
    with pytest.raises(Exception, match="missing"):
        SolNGLOD(spc, cc[keep], counts, sol.cf[keep], sol.w0, sol.b0, sol.w1, sol.b1)


def test_sparse_sdf_parity(sphere, dev):
    net, m, spc = sphere
    sol = SolNGLOD.from_net(net, spc, fp16=False)
    sol_ref, _, _ = sol_arrays(m, spc.octree.cpu().numpy(), 4)
    pyr = spc.pyramid.numpy()
    lv = spc.points_at(6).cpu().numpy()[:, :3].astype(np.float64)
    rng = np.random.default_rng(2)
    k = rng.integers(0, lv.shape[0], 50000)
    x = (((lv[k] + rng.uniform(0.0, 1.0, (50000, 3))) / 64) * 2 - 1).astype(np.float32)
    got = sol.sdf(torch.from_numpy(x).to(dev), torch.from_numpy(k.astype(np.int32)).to(dev))[:, 0].cpu().numpy()
    ref = oracle.sol_eval(sol_ref, x, k)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5)
    dense = oracle.sdf(m, x, 4, "f64")
    np.testing.assert_allclose(got, dense, rtol=0, atol=1e-5)
    assert pyr[0, 6] == lv.shape[0]


def _trace_pair(sol, sol_ref, spc, o, d, dev, lod=4):
    ot, dt = torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev)
    nug, offs = spc.raytrace(ot, dt, lod + 2)
    iters = torch.empty(o.shape[0], dtype=torch.int32, device=dev)
    rb = sol.trace(ot, dt, lod=lod, nuggets=(nug, offs), iters=iters)
    torch.cuda.synchronize()
    ref = oracle.sol_trace(sol_ref, o, d, offs.cpu().numpy(), nug.cpu().numpy())
    got = {"x": rb.x.cpu().numpy(), "t": rb.depth[:, 0].cpu().numpy(), "hit": rb.hit.cpu().numpy(),
           "normal": rb.normal.cpu().numpy(), "iters": iters.cpu().numpy(), "offs": offs.cpu().numpy()}
    return got, ref


def _check_trace(got, ref, o):
    same = got["iters"] == ref["iters"]
    assert same.mean() > 0.99, same.mean()
    assert (got["hit"] == ref["hit"])[same].all()
    both = same & got["hit"]
    assert both.sum() > 100
    assert np.abs(got["t"][both] - ref["t"][both]).max() < 1e-4
    assert np.abs(got["x"][both] - ref["x"][both]).max() < 1e-4
    cos = np.sum(got["normal"][both].astype(np.float64) * ref["normal"][both], 1)
    ang = np.degrees(np.arccos(np.clip(cos, -1, 1)))
    # The decoder's first layer is an FP32 MFMA product, summed in another
    # order than the reference's addmm.  A tap's SDF is a 128-term sum of O(1)
    # terms (rounding ~1e-6 absolute) and the sparse normal differences two
    # taps only 2e-3 apart (eps 1e-3, no division), so the expected angle error
    # is ~1e-6 / 2e-3 = 5e-4 rad = 0.03 degree; and a hit point that differs by
    # up to 1e-4 can move a tap across a ReLU kink of the decoder (a few rays
    # of a frame, up to a few degrees).
    assert np.median(ang) < 0.05, np.median(ang)
    assert np.mean(ang > 0.5) < 1e-3, np.mean(ang > 0.5)
    assert ang.max() < 5.0, ang.max()
    assert np.all(got["normal"][~got["hit"]] == 1.0)
    none = got["offs"][1:] == got["offs"][:-1]
    assert np.all(got["t"][none] == 0.0)
    np.testing.assert_array_equal(got["x"][none], o[none])
    return same


def test_trace_sparse_parity_camera(sphere, dev):
    net, m, spc = sphere
    sol = SolNGLOD.from_net(net, spc, fp16=False)
    sol_ref, _, _ = sol_arrays(m, spc.octree.cpu().numpy(), 4)
    o, d = workloads.camera_rays(256, 256, device="cpu")
    o, d = o.numpy(), d.numpy()
    got, ref = _trace_pair(sol, sol_ref, spc, o, d, dev)
    _check_trace(got, ref, o)
    assert got["hit"].sum() > 10000


def test_trace_sparse_parity_random_rays_and_lods(sphere, dev):
    from tests.helpers import rand_rays
    net, m, spc = sphere
    sol = SolNGLOD.from_net(net, spc, fp16=True)
    for lod in (2, 3, 4):
        sol_ref, _, _ = sol_arrays(m, spc.octree.cpu().numpy(), lod, fp16=True)
        o, d = rand_rays(20000, 40 + lod)
        got, ref = _trace_pair(sol, sol_ref, spc, o, d, dev, lod=lod)
        _check_trace(got, ref, o)


def test_save_load_round_trip(sphere, dev, tmp_path):
    net, m, spc = sphere
    sol = SolNGLOD.from_net(net, spc, fp16=True)
    path = str(tmp_path / "model.npz")
    # levels above 7 do not fit the uint8 corner format; this octree is level 6
    sol.save(path)
    with np.load(path, allow_pickle=False) as f:
        assert sorted(f.files) == sorted(NPZ_KEYS)
        assert f["cc"].dtype == np.uint8 and f["cf"].dtype == np.float16 and f["w0"].dtype == np.float16
        assert f["w0"].shape == (5, 128, 35) and f["w1"].shape == (5, 1, 128) and f["b1"].shape == (5, 1)
        assert f["pyramid"].tolist() == sol.counts
    back = SolNGLOD.load(path, dev)
    o, d = workloads.camera_rays(128, 128, device=dev)
    a = sol.trace(o, d)
    b = back.trace(o, d)
    torch.cuda.synchronize()
    for k in ("x", "depth", "hit", "normal"):
        assert torch.equal(getattr(a, k), getattr(b, k)), k


def test_voxel_sparsify_covers_surface(sphere, dev):
    net, m, spc = sphere
    got = voxel_sparsify(200_000, net, 3, seed=1)
    assert got.level == 5
    lv = got.points_at(5).cpu().numpy()[:, :3].astype(np.float64)
    centers = (lv + 0.5) / 32 * 2 - 1
    # every occupied cell holds a point with |sdf| < 1/32 (half a cell)
    assert np.abs(np.linalg.norm(centers, axis=1) - RADIUS).max() < (1 + np.sqrt(3)) / 32 + 5e-3
    # and the surface is covered: cells of dense surface samples are present
    dirs = oracle.sample_directions(3, 20000).astype(np.float64) * RADIUS
    cells = np.floor((dirs + 1) / 2 * 32).astype(np.int64)
    have = {tuple(c) for c in lv.astype(np.int64)}
    covered = np.mean([tuple(c) in have for c in cells])
    assert covered > 0.99, covered


def test_sparse_agrees_with_dense_tracer(sphere, dev):
    net, m, spc = sphere
    sol = SolNGLOD.from_net(net, spc, fp16=False)
    o, d = workloads.camera_rays(200, 200, device=dev)
    net.lod = 4
    dense = SphereTracer(num_steps=128)(net, o, d)
    sparse = sol.trace(o, d)
    torch.cuda.synchronize()
    od = o.double().cpu().numpy()
    dd = d.double().cpu().numpy()
    b = np.sum(od * dd, 1)
    disc = b * b - (np.sum(od * od, 1) - RADIUS ** 2)
    clear = disc > 0.05
    assert sparse.hit.cpu().numpy()[clear].all() and dense.hit.cpu().numpy()[clear].all()
    diff = (sparse.depth - dense.depth)[:, 0].abs().cpu().numpy()[clear]
    assert diff.max() < 4e-3
