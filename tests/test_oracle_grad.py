"""Pins the oracle's position gradients (autodiff, tetrahedron) and its
get_min restatement: against torch autograd through F.grid_sample and
nn.Linear in FP64 (the reference's operators), against central differences
and closed forms, and get_min's march against the pinned trace."""
import numpy as np
import pytest
import torch

from oracle import oracle
from oracle.torch_ref import TorchOctreeSDF
from tests.helpers import arrays, rand_points, rand_rays, small_model


def _special_points(res, seed):
    """Random points inside and outside the box, points on grid planes of
    every level, and points exactly on the box faces."""
    g = np.random.default_rng(seed)
    pts = [g.uniform(-1.3, 1.3, size=(1500, 3))]
    for r in res:
        p = g.uniform(-1.0, 1.0, size=(r + 1, 3))
        p[:, g.integers(0, 3)] = -1.0 + 2.0 * np.arange(r + 1) / r
        pts.append(p)
    f = g.uniform(-1.0, 1.0, size=(60, 3))
    f[np.arange(60), np.arange(60) % 3] = np.where(np.arange(60) % 2 == 0, -1.0, 1.0)
    pts.append(f)
    return np.concatenate(pts)


def _torch_grad(net, x, lod):
    ref = TorchOctreeSDF.from_module(net, dtype=torch.float64)
    xt = torch.from_numpy(x).double().requires_grad_(True)
    y = ref.sdf(xt, lod=lod)
    (g,) = torch.autograd.grad(y.sum(), xt)
    return y[:, 0].detach().numpy(), g.numpy()


@pytest.mark.parametrize("pos_invariant", [True, False])
@pytest.mark.parametrize("joint", [False, True])
def test_autodiff_gradient_matches_torch_autograd(pos_invariant, joint):
    net = small_model(num_lods=3, pos_invariant=pos_invariant, joint=joint, seed=31)
    with torch.no_grad():
        for f in net.features:
            f.fm.mul_(30.0)
    m = arrays(net)
    x = _special_points(m.res, 5)
    for lod in range(3):
        y, g = oracle.sdf_grad(m, x, lod)
        yt, gt = _torch_grad(net, x, lod)
        np.testing.assert_allclose(y, yt, rtol=0, atol=1e-12)
        np.testing.assert_allclose(g, gt, rtol=0, atol=1e-10 * max(1.0, np.abs(gt).max()))


def test_autodiff_gradient_closed_forms():
    """Central differences (FP64) agree away from cell planes and ReLU kinks;
    outside the box a position-invariant model has no gradient along the
    clamped axis."""
    net = small_model(num_lods=3, seed=32)
    with torch.no_grad():
        for f in net.features:
            f.fm.mul_(30.0)
    m = arrays(net)
    x = rand_points(400, 33, -0.95, 0.95).astype(np.float64)
    _, g = oracle.sdf_grad(m, x, 2)
    h = 1e-7
    fd = np.stack([(oracle.sdf(m, x + h * e, 2, "f64") - oracle.sdf(m, x - h * e, 2, "f64")) / (2 * h)
                   for e in np.eye(3)], 1)
    err = np.abs(fd - g).max(axis=1)
    # a tap pair straddling a cell plane or a ReLU kink is off by O(h) only there
    assert np.mean(err < 1e-5 * max(1.0, np.abs(g).max())) > 0.97
    out = rand_points(300, 34, -1.0, 1.0).astype(np.float64)
    out[:, 0] = np.where(out[:, 0] > 0, 1.0 + np.abs(out[:, 0]), -1.0 - np.abs(out[:, 0]))
    _, go = oracle.sdf_grad(m, out, 2)
    assert np.all(go[:, 0] == 0.0) and np.abs(go[:, 1:]).max() > 0


def test_tetrahedron_gradient_closed_form_and_fd_agreement():
    """On a model whose SDF is affine in x (identity first layer, biases that
    keep every unit active, an affine feature field) the tetrahedron formula
    returns the affine coefficient exactly (each tap's sign and offset
    matter); on a random model it approximates the analytic gradient (the
    stencil's error is O(h) times the field's mixed second derivatives)."""
    rng = np.random.default_rng(35)
    A = rng.normal(size=(32, 3)) * 0.1
    net = small_model(num_lods=1, seed=36)
    c = np.linspace(-1.0, 1.0, 5)
    z, y, xg = np.meshgrid(c, c, c, indexing="ij")
    with torch.no_grad():
        fm = net.features[0].fm
        for ch in range(32):
            fm[0, ch].copy_(torch.from_numpy(A[ch, 0] * xg + A[ch, 1] * y + A[ch, 2] * z))
        l = net.louts[0]
        W0 = np.zeros((128, 32))
        W0[np.arange(32), np.arange(32)] = 1.0
        l[0].weight.copy_(torch.from_numpy(W0))
        l[0].bias.fill_(10.0)  # every unit active: the SDF is affine
        w1 = rng.normal(size=128)
        l[2].weight.copy_(torch.from_numpy(w1).reshape(1, -1))
        l[2].bias.fill_(0.0)
    m = arrays(net)
    x = rand_points(300, 37, -0.9, 0.9)
    g = oracle.tetra_gradient(m, x, 0, "f64")
    grid = m.grids[0].reshape(5, 5, 5, 32).astype(np.float64)
    A_s = np.stack([(grid[0, 0, 4] - grid[0, 0, 0]) / 2, (grid[0, 4, 0] - grid[0, 0, 0]) / 2,
                    (grid[4, 0, 0] - grid[0, 0, 0]) / 2], 1)
    want = m.w1[0][:32].astype(np.float64) @ A_s
    # the float32 corner values are affine up to their rounding, which the
    # division by 4h amplifies to ~1e-7
    np.testing.assert_allclose(g, np.broadcast_to(want, g.shape), rtol=0, atol=1e-6)
    net2 = small_model(num_lods=2, seed=38)
    m2 = arrays(net2)
    x2 = rand_points(500, 39, -0.9, 0.9)
    gt = oracle.tetra_gradient(m2, x2, 1, "f64")
    _, ga = oracle.sdf_grad(m2, x2, 1)
    rel = np.linalg.norm(gt - ga, axis=1).mean() / np.linalg.norm(ga, axis=1).mean()
    assert rel < 0.2, rel


def test_get_min_march_equals_the_trace():
    """get_min marches with forward's recurrence: for every ray that enters
    the box its x, t and hit equal the oracle trace's; min_x is a point of
    the march (o + d t' for a visited t')."""
    net = small_model(num_lods=3, seed=40)
    m = arrays(net)
    o, d = rand_rays(300, 41)
    ref = oracle.trace(m, o, d, 2, num_steps=64)
    got = oracle.get_min(m, o, d, 2, num_steps=64)
    _, _, entered = oracle.aabb(o, d)
    assert entered.sum() > 100
    np.testing.assert_array_equal(got["x"][entered], ref["x"][entered])
    np.testing.assert_array_equal(got["t"][entered], ref["t"][entered])
    np.testing.assert_array_equal(got["hit"][entered], ref["hit"][entered])
    # min_x lies on the ray: its offset from o is parallel to d
    off = got["min_x"] - o
    cross = np.cross(off.astype(np.float64), d.astype(np.float64))
    assert np.abs(cross).max() < 1e-4 * max(1.0, np.abs(off).max())


def _plane_model(normal, offset):
    """A model whose SDF is exactly affine, d(x) = normal . x + offset: an
    affine feature field (channel 0 = normal . x), identity first layer with
    biases that keep every unit active, second layer reading unit 0."""
    net = small_model(num_lods=1, seed=42)
    c = np.linspace(-1.0, 1.0, 5)
    z, y, xg = np.meshgrid(c, c, c, indexing="ij")
    with torch.no_grad():
        fm = net.features[0].fm
        fm.zero_()
        fm[0, 0].copy_(torch.from_numpy(normal[0] * xg + normal[1] * y + normal[2] * z))
        l = net.louts[0]
        W0 = np.zeros((128, 32))
        W0[np.arange(32), np.arange(32)] = 1.0
        l[0].weight.copy_(torch.from_numpy(W0))
        l[0].bias.fill_(8.0)  # every unit active
        l[2].weight.zero_()
        l[2].weight[0, 0] = 1.0
        l[2].bias.fill_(offset - 8.0)
    return net


def test_get_min_on_a_plane_tracks_the_decreasing_distance():
    """Rays approaching the plane x + 0.5 = 0 see the distance fall at every
    step, so the smallest distance is the last one evaluated and get_min's
    min_x is the final point (a comparison that never moves min_x, or that
    keeps the point where the distance was evaluated, gives the previous one);
    rays leaving the plane keep their first point."""
    m = arrays(_plane_model(np.array([1.0, 0.0, 0.0]), 0.5))
    rng = np.random.default_rng(43)
    n = 200
    o = np.stack([np.full(n, 2.5), rng.uniform(-0.4, 0.4, n), rng.uniform(-0.4, 0.4, n)], 1).astype(np.float32)
    d = np.stack([-np.ones(n), rng.uniform(-0.15, 0.15, n), rng.uniform(-0.15, 0.15, n)], 1)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    got = oracle.get_min(m, o, d, 0, num_steps=64)
    assert got["hit"].all()
    np.testing.assert_allclose(got["x"][:, 0], -0.5, atol=1e-3)
    assert np.array_equal(got["min_x"], got["x"])
    # moving away from the plane from inside the box: the first point stays
    o2 = np.stack([np.full(n, -0.2), rng.uniform(-0.4, 0.4, n), rng.uniform(-0.4, 0.4, n)], 1).astype(np.float32)
    d2 = -d
    away = oracle.get_min(m, o2, d2, 0, num_steps=64)
    ent, _, _ = oracle.aabb(o2, d2)
    assert np.array_equal(away["min_x"], ent)
