"""GPU parity of the sparse octree path against the oracle, bit-exact:
occupancy codes, octree bytes, points and pyramid, and the nugget lists of
a 1280x720 frame at the benchmark levels."""
import numpy as np
import pytest
import torch

from nglod_amd import mesh, workloads
from nglod_amd.spc import SPC
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def shape():
    V, F = mesh.displaced_icosphere(0)
    return mesh.triangles(V, F)


@pytest.mark.parametrize("level", [3, 7])
def test_octree_build_bitexact(dev, shape, level):
    spc = SPC.from_mesh(torch.from_numpy(shape), level, dev)
    codes = oracle.occupancy(shape, level)
    octree, counts = oracle.octree_from_codes(codes, level)
    np.testing.assert_array_equal(spc.octree.cpu().numpy(), octree)
    pts, pyr, L = oracle.octree_points(octree)
    assert spc.level == L == level
    np.testing.assert_array_equal(spc.pyramid.numpy(), pyr)
    np.testing.assert_array_equal(spc.points.cpu().numpy(), pts)
    if level == 7:
        assert 2e4 < counts[-1] < 3e5  # on the order of 10^5 leaves


def test_spc_from_octree_bytes(dev, shape):
    codes = oracle.occupancy(shape, 5)
    octree, _ = oracle.octree_from_codes(codes, 5)
    spc = SPC(torch.from_numpy(octree), dev)
    pts, pyr, L = oracle.octree_points(octree)
    np.testing.assert_array_equal(spc.points.cpu().numpy(), pts)
    np.testing.assert_array_equal(spc.pyramid.numpy(), pyr)


@pytest.mark.parametrize("target", [6, 7])
def test_raytrace_frame_bitexact(dev, shape, target):
    spc = SPC.from_mesh(torch.from_numpy(shape), 7, dev)
    o, d = workloads.camera_rays(1280, 720, device="cpu")
    nug, offs = spc.raytrace(o.to(dev), d.to(dev), target)
    want = oracle.spc_raytrace(spc.octree.cpu().numpy(), spc.points.cpu().numpy(), spc.pyramid.numpy(), spc.level,
                               o.numpy(), d.numpy(), target)
    assert want.shape[0] > 100000
    np.testing.assert_array_equal(nug.cpu().numpy(), want)
    cnt = np.bincount(want[:, 0], minlength=o.shape[0])
    np.testing.assert_array_equal(offs.cpu().numpy(), np.concatenate([[0], np.cumsum(cnt)]))
    info = SPC.mark_unique_rays(nug)
    np.testing.assert_array_equal(info.cpu().numpy(), oracle.mark_unique_rays(want))


def test_raytrace_random_rays_and_edge_cases(dev, shape):
    spc = SPC.from_mesh(torch.from_numpy(shape), 6, dev)
    rng = np.random.default_rng(3)
    o = rng.uniform(-3, 3, size=(20000, 3)).astype(np.float32)   # includes origins inside the box
    d = rng.normal(size=(20000, 3))
    d[:50, 1:] = 0.0                                               # axis-aligned directions
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    for target in (0, 3, 6):
        nug, _ = spc.raytrace(torch.from_numpy(o).to(dev), torch.from_numpy(d).to(dev), target)
        want = oracle.spc_raytrace(spc.octree.cpu().numpy(), spc.points.cpu().numpy(), spc.pyramid.numpy(), 6, o, d,
                                   target)
        np.testing.assert_array_equal(nug.cpu().numpy(), want)
    empty, offs = spc.raytrace(torch.zeros(0, 3, device=dev), torch.zeros(0, 3, device=dev), 3)
    assert empty.shape == (0, 2)
