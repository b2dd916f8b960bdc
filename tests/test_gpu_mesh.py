"""GPU parity of the training-target path against the oracle: FP64
mesh2sdf (distance to 1e-12 relative, sign exact) and the samplers."""
import numpy as np
import pytest
import torch

from nglod_amd import mesh
from nglod_amd.meshdata import MeshData
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def md(dev):
    V, F = mesh.displaced_icosphere(0)
    return MeshData(V, F, dev), mesh.triangles(V, F)


def test_mesh2sdf_f64_vs_oracle(md, dev):
    data, T = md
    rng = np.random.default_rng(5)
    pts = np.concatenate([rng.uniform(-1.2, 1.2, size=(6000, 3)),
                          oracle.sample_surface(T, mesh.face_cdf(T), 11, 6000, 0.003)[0]]).astype(np.float32)
    d32, d64 = data.sdf(torch.from_numpy(pts).to(dev), with_f64=True)
    want = oracle.mesh2sdf(pts, T)
    got = d64.cpu().numpy()
    np.testing.assert_array_equal(np.sign(got), np.sign(want))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(d32.cpu().numpy(), want.astype(np.float32), rtol=2e-7, atol=0)


def test_mesh2sdf_empty_and_single(md, dev):
    data, T = md
    out = data.sdf(torch.zeros(0, 3, device=dev))
    assert out.shape == (0,)
    c = data.sdf(torch.zeros(1, 3, device=dev))
    assert c.item() < 0  # centre is inside


def test_uniform_sampler_bitexact(md, dev):
    data, _ = md
    got = data.sample_uniform(50001, 77).cpu().numpy()
    np.testing.assert_array_equal(got, oracle.sample_uniform(77, 50001))


def test_surface_sampler_vs_oracle(md, dev):
    data, T = md
    cdf = mesh.face_cdf(T)
    for sigma in (0.0, 0.01):
        p, f = data.sample_surface(40000, 123, sigma=sigma, return_faces=True)
        po, fo = oracle.sample_surface(T, cdf, 123, 40000, sigma)
        np.testing.assert_array_equal(f.cpu().numpy(), fo)
        np.testing.assert_allclose(p.cpu().numpy(), po, rtol=0, atol=2e-6)


def test_training_batch_composition(md, dev):
    data, T = md
    pts, gts = data.training_batch(3000, seed=3)
    assert pts.shape == (3000, 3) and gts.shape == (3000, 1)
    p = pts.cpu().numpy()
    assert np.all(np.abs(p[:1000]) <= 1.0)
    want = oracle.mesh2sdf(p, T)
    np.testing.assert_allclose(gts[:, 0].cpu().numpy(), want.astype(np.float32), rtol=2e-7)
    # near-surface part is close to the surface
    assert np.median(np.abs(want[1000:])) < 0.02
