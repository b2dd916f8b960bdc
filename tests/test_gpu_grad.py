"""Position gradients on the GPU (nglod_gradient, nglod_sdf_grad, autograd
with respect to the query points), the tracers' grad_method options,
diffutils.gradient and SphereTracer.get_min / sample_surface, against the
CPU oracle."""
import numpy as np
import pytest
import torch

from nglod_amd import SphereTracer, _lib, workloads
from nglod_amd.diffutils import gradient
from oracle import oracle
from tests.helpers import arrays, rand_points, rand_rays

pytestmark = pytest.mark.gpu


def gpu(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _model(dev, pos_invariant=True, seed=51, lods=5):
    net = workloads.make_model(num_lods=lods, pos_invariant=pos_invariant, seed=seed, device="cpu")
    with torch.no_grad():
        for f in net.features:
            f.fm.mul_(30.0)
    return net, arrays(net), net.to(dev)


def _points(seed):
    g = np.random.default_rng(seed)
    x = g.uniform(-1.1, 1.1, size=(6000, 3)).astype(np.float32)
    x[:64, 0] = 1.0  # on a box face: no gradient through the clamp along x
    return x


@pytest.mark.parametrize("pos_invariant", [True, False])
def test_autodiff_gradient_vs_oracle(dev, pos_invariant):
    net, m, netg = _model(dev, pos_invariant)
    x = _points(52)
    for lod in (0, 2, 4):
        got = netg.gradient_hip(gpu(x, dev), lod, "autodiff").cpu().numpy().astype(np.float64)
        y, want = oracle.sdf_grad(m, x, lod)
        scale = np.abs(want).max()
        err = np.abs(got - want).max(axis=1)
        # FP32 against FP64: a ReLU unit whose pre-activation is within FP32
        # rounding of 0 may switch, so a few points may differ more
        assert np.mean(err <= 1e-4 * scale) >= 0.999, (lod, np.sort(err)[-5:], scale)
        if pos_invariant:
            assert np.all(got[:64, 0] == 0.0)
        yg, gg = netg.sdf_grad_hip(gpu(x, dev), lod)
        assert torch.equal(gg.cpu(), torch.from_numpy(got.astype(np.float32)))
        assert np.abs(yg[:, 0].cpu().numpy() - y).max() <= 1e-5 * max(1.0, np.abs(y).max())


def test_tetrahedron_and_fd_vs_oracle(dev):
    net, m, netg = _model(dev)
    x = rand_points(4000, 53, -0.95, 0.95)
    for lod in (1, 4):
        got = netg.gradient_hip(gpu(x, dev), lod, "tetrahedron").cpu().numpy()
        want = oracle.tetra_gradient(m, x, lod, "f32")
        # the taps' FP32 SDF values differ from the oracle's by ~1e-7; / 4h
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 2e-4 * scale + 5e-5, np.abs(got - want).max()
        fd = netg.gradient_hip(gpu(x, dev), lod, "finitediff", normalize=True).cpu().numpy()
        want_fd = oracle.normals(m, x, lod, "f32")
        cos = (fd * want_fd).sum(1)
        assert np.degrees(np.arccos(np.clip(cos, -1, 1))).max() < 0.05


def test_mask_and_normalize(dev):
    net, m, netg = _model(dev)
    x = rand_points(1000, 54)
    mask = torch.from_numpy(np.arange(1000) % 3 == 0).to(dev)
    raw = netg.gradient_hip(gpu(x, dev), 3, "autodiff")
    nrm = netg.gradient_hip(gpu(x, dev), 3, "autodiff", mask=mask, normalize=True)
    assert torch.all(nrm[~mask] == 0)
    want = raw[mask] / raw[mask].norm(dim=1, keepdim=True).clamp_min(1e-5)
    assert (nrm[mask] - want).abs().max().item() < 1e-6


def test_autograd_position_gradient(dev):
    """x.requires_grad: d(sum sdf)/dx from the custom autograd Function equals
    the oracle's FP64 autodiff gradient; parameter gradients still flow."""
    net, m, netg = _model(dev, pos_invariant=False)
    x = rand_points(3000, 55)
    xt = gpu(x, dev).requires_grad_(True)
    y = netg.sdf(xt, lod=3)
    y.sum().backward()
    _, want = oracle.sdf_grad(m, x, 3)
    err = np.abs(xt.grad.cpu().numpy() - want).max(axis=1)
    assert np.mean(err <= 1e-4 * np.abs(want).max()) >= 0.999
    assert netg.louts[3][0].weight.grad is not None
    # every level at once (lod None): the gradient of the sum over levels
    netg.zero_grad()
    xt2 = gpu(x, dev).requires_grad_(True)
    saved = netg.lod
    netg.lod = None
    try:
        preds = netg.sdf(xt2, return_lst=True)
        sum(p.sum() for p in preds).backward()
    finally:
        netg.lod = saved
    want_all = sum(oracle.sdf_grad(m, x, l)[1] for l in range(5))
    err = np.abs(xt2.grad.cpu().numpy() - want_all).max(axis=1)
    assert np.mean(err <= 1e-4 * np.abs(want_all).max()) >= 0.999


def test_diffutils_gradient_drop_in(dev):
    net, m, netg = _model(dev)
    netg.lod = 2
    x = gpu(rand_points(500, 56), dev)
    with torch.no_grad():
        for meth in ("autodiff", "tetrahedron", "finitediff"):
            g = gradient(x, netg, meth)
            assert g.shape == (500, 3) and torch.equal(g, netg.gradient_hip(x, 2, meth))
        ml = gradient(x, netg, "multilayer")
        assert len(ml) == 3 and torch.equal(ml[2], netg.gradient_hip(x, 2, "autodiff"))
    # with parameters that require grad: taps stay differentiable, autodiff refuses a second order
    g = gradient(x, netg, "finitediff")
    g.sum().backward()
    assert netg.features[2].fm.grad is not None
    ga = gradient(x, netg, "autodiff")
    with pytest.raises(NotImplementedError):
        ga.sum().backward()
    # any other callable: the reference's torch arithmetic
    f = lambda q: (q * q).sum(-1, keepdim=True)
    assert torch.allclose(gradient(x, f, "autodiff"), 2 * x)


@pytest.mark.parametrize("grad_method", ["tetrahedron", "autodiff"])
@pytest.mark.parametrize("form", ["fp32", "mixed"])
def test_tracer_grad_methods(dev, grad_method, form):
    net = workloads.make_sphere_model(radius=0.5, num_lods=5, device="cpu", seed=3)
    m = arrays(net)
    netg = net.to(dev)
    o, d = rand_rays(3000, 57)
    fd = SphereTracer(num_steps=128, form=form)(netg, gpu(o, dev), gpu(d, dev))
    rb = SphereTracer(num_steps=128, form=form, grad_method=grad_method)(netg, gpu(o, dev), gpu(d, dev))
    assert torch.equal(rb.hit, fd.hit) and torch.equal(rb.depth, fd.depth) and torch.equal(rb.x, fd.x)
    hit = rb.hit.cpu().numpy()
    assert hit.sum() > 500
    assert torch.all(rb.normal[~rb.hit] == 0)
    xh = rb.x.cpu().numpy()[hit]
    if grad_method == "autodiff":
        want = oracle.sdf_grad(m, xh, 4)[1]
    else:
        want = oracle.tetra_gradient(m, xh, 4, "f64")
    want /= np.maximum(np.linalg.norm(want, axis=1, keepdims=True), 1e-5)
    got = rb.normal.cpu().numpy()[hit]
    ang = np.degrees(np.arccos(np.clip((got * want).sum(1), -1, 1)))
    assert np.median(ang) < 0.01 and np.percentile(ang, 99.9) < 0.5, (np.median(ang), ang.max())


def test_get_min_vs_oracle(dev):
    net = workloads.make_sphere_model(radius=0.55, num_lods=3, device="cpu", seed=4, noise=5e-3)
    m = arrays(net)
    netg = net.to(dev)
    o, d = rand_rays(4000, 58)
    d[:200] = -d[:200]  # rays pointing away from the box: they march from their origins
    tr = SphereTracer(num_steps=64)
    rb = tr.get_min(netg, gpu(o, dev), gpu(d, dev))
    want = oracle.get_min(m, o, d, 2, num_steps=64)
    hit = rb.hit.cpu().numpy()
    assert (hit == want["hit"]).mean() > 0.995
    both = hit & want["hit"]
    assert both.sum() > 1000
    x = rb.x.cpu().numpy()
    t = rb.depth[:, 0].cpu().numpy()
    assert np.median(np.abs(t[both] - want["t"][both])) < 1e-5
    mx = rb.min_x.cpu().numpy()
    same = np.abs(mx - want["min_x"]).max(axis=1) < 1e-4
    assert same.mean() > 0.99, same.mean()
    # normals: the raw (unnormalised) gradient by grad_method at the hits, zero elsewhere
    assert torch.all(rb.normal[~rb.hit] == 0)
    g = netg.gradient_hip(rb.x, 2, "finitediff", mask=rb.hit)
    assert torch.equal(rb.normal, g)
    assert np.median(np.abs(x[both] - want["x"][both])) < 1e-5


def test_sample_surface(dev):
    net = workloads.make_sphere_model(radius=0.5, num_lods=3, device="cpu", seed=5)
    netg = net.to(dev)
    tr = SphereTracer(num_steps=64)
    pts = tr.sample_surface(2000, netg)
    assert pts.shape[0] >= 2000 and pts.shape[1] == 3
    assert pts.abs().max().item() <= 1.0
