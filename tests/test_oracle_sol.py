"""Pins the oracle's sparse (SOL) path: corner lists and trinkets by their
defining coordinates, the sparse SDF against the dense FP64 oracle (in an
occupied cell the two are the same trilinear sum), and the sparse march
against the analytic sphere and the dense model's central differences, plus
its special cases (no cell on the ray, all cells passed, origin inside a
cell)."""
import numpy as np
import pytest

from nglod_amd import mesh, workloads
from oracle import oracle
from tests.helpers import arrays, full_octree, level_index, small_model, sol_arrays

OFFS = np.array([[(j & 4) >> 2, (j & 2) >> 1, j & 1] for j in range(8)])


def test_unique_corners_small_cases():
    one = oracle.unique_corners(np.array([[3, 4, 5, 0]], dtype=np.int16))
    np.testing.assert_array_equal(one, np.array([3, 4, 5]) + OFFS[np.lexsort(OFFS.T[::-1])])
    two = oracle.unique_corners(np.array([[0, 0, 0, 0], [1, 0, 0, 0]], dtype=np.int16))
    assert two.shape == (12, 3)  # two face neighbours share 4 corners
    key = two[:, 0] * 10000 + two[:, 1] * 100 + two[:, 2]
    assert np.all(np.diff(key) > 0)


def _sphere_octree(radius, level):
    V, F = mesh.icosphere(4)
    T = mesh.triangles(V * radius, F)
    codes = oracle.occupancy(T, level)
    return oracle.octree_from_codes(codes, level)[0]


def test_trinkets_point_at_corners_and_parents():
    net = small_model(num_lods=4, pos_invariant=False)
    m = arrays(net)
    oc = _sphere_octree(0.6, 5)
    sol, blocks, (pts, pyr, L) = sol_arrays(m, oc, 3)
    for l in range(2, 6):
        beg, cnt = pyr[1, l], pyr[0, l]
        p = pts[beg:beg + cnt, :3].astype(np.int64)
        tv = sol.tv[beg:beg + cnt].astype(np.int64)
        assert np.all(tv != 0xFFFFFFFF)
        for j in range(8):
            np.testing.assert_array_equal(sol.cc[tv[:, j]], p + OFFS[j])
        par = sol.parent[beg:beg + cnt]
        if l == 2:
            assert np.all(par == -1)
        else:
            assert np.all((par >= pyr[1, l - 1]) & (par < pyr[1, l - 1] + pyr[0, l - 1]))
            np.testing.assert_array_equal(pts[par, :3].astype(np.int64), p // 2)


@pytest.mark.parametrize("pos_invariant", [False, True])
def test_sparse_sdf_equals_dense_in_occupied_cells(pos_invariant):
    net = small_model(num_lods=3, pos_invariant=pos_invariant, seed=4)
    m = arrays(net)
    rng = np.random.default_rng(1)
    # full octree: every point is in a cell
    sol, _, (pts, pyr, _) = sol_arrays(m, full_octree(4), 2)
    x = rng.uniform(-0.999, 0.999, (3000, 3)).astype(np.float32)
    cells = np.floor((x.astype(np.float64) + 1) / 2 * 16).astype(np.int64)
    pi = level_index(pts, pyr, 4, cells)
    assert np.all(pi >= 0)
    np.testing.assert_allclose(oracle.sol_eval(sol, x, pi), oracle.sdf(m, x, 2, "f64"), rtol=0, atol=2e-6)
    # partial octree: points inside occupied cells only
    sol, _, (pts, pyr, _) = sol_arrays(m, _sphere_octree(0.5, 4), 2)
    lv = pts[pyr[1, 4]:pyr[1, 4] + pyr[0, 4], :3].astype(np.float64)
    k = rng.integers(0, lv.shape[0], 3000)
    x = (((lv[k] + rng.uniform(0.01, 0.99, (3000, 3))) / 16) * 2 - 1).astype(np.float32)
    np.testing.assert_allclose(oracle.sol_eval(sol, x, k), oracle.sdf(m, x, 2, "f64"), rtol=0, atol=2e-6)


def _sphere_sol(lod=3, radius=0.6):
    net = workloads.make_sphere_model(radius=radius, num_lods=lod + 1, pos_invariant=False, device="cpu")
    m = arrays(net)
    return sol_arrays(m, _sphere_octree(radius, lod + 2), lod) + (m,)


def test_sparse_march_hits_the_sphere():
    radius = 0.6
    sol, _, (pts, pyr, L), m = _sphere_sol(3, radius)
    o, d = workloads.camera_rays(96, 96, device="cpu")
    o, d = o.numpy(), d.numpy()
    nug = oracle.spc_raytrace(oracle_octree := _sphere_octree(radius, 5), *oracle.octree_points(oracle_octree)[:2],
                              5, o, d, 5)
    offs = oracle.nugget_offsets(nug, o.shape[0])
    r = oracle.sol_trace(sol, o, d, offs, nug)
    od = o.astype(np.float64)
    dd = d.astype(np.float64)
    b = np.sum(od * dd, 1)
    disc = b * b - (np.sum(od * od, 1) - radius ** 2)
    t_true = -b - np.sqrt(np.maximum(disc, 0))
    clear_hit = disc > 0.05
    clear_miss = disc < -0.05
    assert clear_hit.sum() > 1000 and clear_miss.sum() > 1000
    assert r["hit"][clear_hit].all()
    assert not r["hit"][clear_miss].any()
    # the S stop rule allows |d + dprev|/2 < 1.5e-3 at the last step
    assert np.abs(r["t"][clear_hit] - t_true[clear_hit]).max() < 4e-3
    xh = r["x"][clear_hit].astype(np.float64)
    n_true = xh / np.linalg.norm(xh, axis=1, keepdims=True)
    ang = lambda a, b: np.degrees(np.arccos(np.clip(np.sum(a * b, 1), -1, 1)))
    # the level-5 interpolant's normals are within a few degrees of the sphere's
    assert ang(r["normal"][clear_hit], n_true).max() < 6.0
    # and equal the dense model's central differences (h = 1e-3) at x
    g = np.zeros_like(xh)
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1e-3
        g[:, k] = oracle.sdf(m, xh + e, 3, "f64") - oracle.sdf(m, xh - e, 3, "f64")
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    # both taps use the hit cell's corners (SDF.cu:218-295), so the two agree
    # where x +- 1e-3 stays inside that cell
    u = (xh + 1.0) / 2.0 * 32
    interior = np.all(np.abs(u - np.round(u)) > 1e-3 * 16 * 1.01, axis=1)
    assert interior.sum() > 0.8 * interior.size
    assert ang(r["normal"][clear_hit].astype(np.float64)[interior], g[interior]).max() < 0.05
    # rays without cells keep t = 0, x = origin (SDF.cu:323-326); rays whose
    # cells are all passed get t = 100; misses have normal (1, 1, 1)
    no_cells = offs[1:] == offs[:-1]
    assert no_cells.sum() > 100
    assert np.all(r["t"][no_cells] == 0.0) and np.all(r["iters"][no_cells] == 0)
    np.testing.assert_array_equal(r["x"][no_cells], o[no_cells])
    passed = ~no_cells & ~r["hit"] & (r["iters"] < 50)  # not stopped by the 50-step cap
    assert passed.sum() > 100 and np.all(r["t"][passed] == 100.0)
    assert np.all(r["normal"][~r["hit"]] == 1.0)


def test_sparse_march_origin_inside_a_cell():
    radius = 0.6
    sol, _, (pts, pyr, L), _ = _sphere_sol(3, radius)
    # origins just outside the sphere surface, marching inward along -x
    rng = np.random.default_rng(3)
    v = rng.normal(size=(64, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    o = (v * (radius + 0.01)).astype(np.float32)
    d = (-v).astype(np.float32)
    octree = _sphere_octree(radius, 5)
    p2, pyr2, _ = oracle.octree_points(octree)
    nug = oracle.spc_raytrace(octree, p2, pyr2, 5, o, d, 5)
    offs = oracle.nugget_offsets(nug, o.shape[0])
    r = oracle.sol_trace(sol, o, d, offs, nug)
    inside = level_index(pts, pyr, 5, np.floor((o.astype(np.float64) + 1) / 2 * 32).astype(np.int64)) >= 0
    assert inside.sum() > 32
    # first evaluation at the origin (t unchanged by the entry), surface 0.01 ahead
    assert r["hit"][inside].all()
    assert np.abs(r["t"][inside] - 0.01).max() < 4e-3
