"""GPU parity of the dense query path (libnglod_hip.so through the C ABI)
against the CPU oracle on the same seeded inputs.

Bars: box entry bit-exact; SDF within 1e-5 abs of the FP32 oracle; traced
hit depths within 1e-4 and hit flags equal on every ray that is not
threshold-ambiguous (tests/helpers.ambiguous_rays), ambiguous rays < 1%."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from nglod_amd import Renderer, RenderBuffer, SphereTracer, _lib, workloads
from oracle import oracle
from tests.helpers import ambiguous_rays, arrays, rand_points, rand_rays, small_model

pytestmark = pytest.mark.gpu


def gpu(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def special_rays():
    o = np.array([[0, 0, 3], [0.5, 0.5, 0.5], [3, 0.5, 0.25], [0.2, -5, 0.1], [0, 2, 0], [2, 2, 2], [1, 1, 1],
                  [-1, 0.3, 0.2], [0, 0, 600], [0, 0, -3], [1e-8, 0, 3], [0.99999, 0.99999, 3]], np.float32)
    d = np.array([[0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 1], [-1, -1, -1],
                  [1, 0, 0], [0, 0, -1], [0, 0, 1], [0, 0, -1], [0, 0, -1]], np.float32)
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    return o, d.astype(np.float32)


def test_aabb_bitexact(dev):
    from nglod_amd import _lib
    o1, d1 = rand_rays(50000, 21, dist=(0.2, 5.0))
    o2, d2 = special_rays()
    o = np.concatenate([o1, o2])
    d = np.concatenate([d1, d2])
    d[5, 1] = 0.0  # zero direction components: +-inf reciprocals
    d[7, 0] = -0.0
    xo, to, ho = oracle.aabb(o, d, "f32")
    og, dg = gpu(o, dev), gpu(d, dev)
    n = o.shape[0]
    x = torch.empty(n, 3, device=dev)
    t = torch.empty(n, device=dev)
    h = torch.empty(n, dtype=torch.uint8, device=dev)
    lib = _lib.load()
    _lib.check(lib.nglod_aabb(_lib.ptr(og), _lib.ptr(dg), n, _lib.ptr(x), _lib.ptr(t), _lib.ptr(h),
                              _lib.stream_of(og)), "nglod_aabb")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h.cpu().numpy().astype(bool), ho)
    np.testing.assert_array_equal(t.cpu().numpy(), to)
    np.testing.assert_array_equal(x.cpu().numpy(), xo)


def test_feature_volume_matches_grid_sample(dev):
    net = small_model(num_lods=3, seed=22)
    x = rand_points(20000, 23, -1.2, 1.2)
    netg = net.to(dev)
    for i, f in enumerate(netg.features):
        got = f(gpu(x, dev)).detach().cpu().numpy().astype(np.float64)
        fm = net.features[i].fm.detach().cpu().double().contiguous()
        want = F.grid_sample(fm, torch.from_numpy(x).double().reshape(1, -1, 1, 1, 3), align_corners=True,
                             padding_mode="border")[0, :, :, 0, 0].T.numpy()
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-7)


@pytest.mark.parametrize("pos_invariant,joint", [(True, False), (False, False), (True, True)])
def test_sdf_fwd_vs_oracle(dev, pos_invariant, joint):
    net = workloads.make_model(num_lods=5, pos_invariant=pos_invariant, joint_decoder=joint, seed=24, device="cpu")
    m = arrays(net)
    x = rand_points(30000, 25, -1.2, 1.2)
    netg = net.to(dev)
    xg = gpu(x, dev)
    for lod in range(5):
        got = netg.sdf(xg, lod=lod)[:, 0].detach().cpu().numpy()
        want = oracle.sdf(m, x, lod, "f32")
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-5)
        want64 = oracle.sdf(m, x, lod, "f64")
        np.testing.assert_allclose(got, want64, rtol=0, atol=1e-5)


def test_sdf_return_lst_and_lod_none(dev):
    net = workloads.make_model(num_lods=5, pos_invariant=False, seed=26, device="cpu")
    m = arrays(net)
    x = rand_points(8000, 27)
    netg = net.to(dev).eval()
    with torch.no_grad():
        lst = netg.sdf(gpu(x, dev), return_lst=True)
        last = netg.sdf(gpu(x, dev))
        netg.lod = 2
        via_forward = netg(gpu(x, dev), lod=0)  # BaseLOD.forward ignores its lod argument
    want = oracle.sdf_all(m, x, 4, "f32")
    assert len(lst) == 5
    for i in range(5):
        np.testing.assert_allclose(lst[i][:, 0].cpu().numpy(), want[:, i], rtol=0, atol=1e-5)
    np.testing.assert_array_equal(last.cpu().numpy(), lst[-1].cpu().numpy())
    np.testing.assert_allclose(via_forward[:, 0].cpu().numpy(), want[:, 2], rtol=0, atol=1e-5)


def test_empty_input(dev):
    net = small_model(num_lods=2).to(dev)
    out = net.sdf(torch.empty(0, 3, device=dev), lod=1)
    assert out.shape == (0, 1)


def test_cpu_tensor_is_rejected():
    net = small_model(num_lods=2)
    with pytest.raises(ValueError):
        net.sdf(torch.zeros(4, 3), lod=0)


def _trace_both(net, o, d, lod, num_steps, dev, form="fp32"):
    m = arrays(net)
    res = oracle.trace(m, o, d, lod, num_steps=num_steps, prec="f32")
    netg = net.to(dev)
    netg.lod = lod
    tracer = SphereTracer(num_steps=num_steps, min_dis=3e-4, camera_clamp=[-5, 10], step_size=1.0, form=form)
    steps = torch.zeros(o.shape[0], dtype=torch.int32, device=dev)
    rb = tracer(netg, gpu(o, dev), gpu(d, dev), ray_steps=steps)
    torch.cuda.synchronize()
    g = dict(x=rb.x.cpu().numpy(), t=rb.depth[:, 0].cpu().numpy(), hit=rb.hit.cpu().numpy(),
             normal=rb.normal.cpu().numpy(), steps=steps.cpu().numpy(), evals=int(tracer.last_evals.sum().item()))
    return res, g


def _check_trace(res, g, num_steps, max_amb=0.01, cap_is_ambiguous=True):
    amb = ambiguous_rays(res, g["steps"], num_steps, cap_is_ambiguous=cap_is_ambiguous)
    ok = ~amb
    assert amb.mean() < max_amb, amb.mean()
    np.testing.assert_array_equal(g["hit"][ok], res["hit"][ok])
    hb = ok & res["hit"]
    np.testing.assert_allclose(g["t"][hb], res["t"][hb], rtol=0, atol=1e-4)
    np.testing.assert_allclose(g["x"][hb], res["x"][hb], rtol=0, atol=1e-4)
    np.testing.assert_allclose(g["normal"][hb], res["normal"][hb], rtol=0, atol=2e-3)
    # misses carry zero normals
    assert np.all(g["normal"][~g["hit"]] == 0)
    return amb


def test_trace_sphere_model_camera(dev):
    net = workloads.make_sphere_model(radius=0.6, num_lods=5, device="cpu")
    o, d = workloads.camera_rays(96, 96, device="cpu")
    o, d = o.numpy(), d.numpy()
    res, g = _trace_both(net, o, d, 4, 128, dev)
    _check_trace(res, g, 128)
    assert g["hit"].sum() > 1000
    # every SDF evaluation is counted: the oracle's per-ray total summed
    assert g["evals"] == int(res["evals"].sum())


@pytest.mark.parametrize("pos_invariant", [True, False])
def test_trace_random_rays_vs_oracle(dev, pos_invariant):
    net = workloads.make_sphere_model(radius=0.45, num_lods=5, pos_invariant=pos_invariant, device="cpu", seed=3)
    o, d = rand_rays(12000, 28, dist=(0.3, 4.0))  # includes origins inside the box
    res, g = _trace_both(net, o, d, 4, 100, dev)
    _check_trace(res, g, 100)


@pytest.mark.parametrize("lod", [0, 1, 2, 3])
def test_trace_lower_lods(dev, lod):
    """Coarser levels: the sphere model at lod L is the level-L interpolant."""
    net = workloads.make_sphere_model(radius=0.55, num_lods=5, device="cpu", seed=7, noise=3e-3)
    o, d = workloads.camera_rays(80, 60, eye=(0.3, 2.5, 1.5), device="cpu")
    res, g = _trace_both(net, o.numpy(), d.numpy(), lod, 128, dev)
    _check_trace(res, g, 128)
    assert g["hit"].sum() > 500


def test_trace_same_alone_and_in_the_mix(dev):
    """Per-ray results do not depend on the other rays in the launch."""
    net = workloads.make_sphere_model(radius=0.6, num_lods=5, device=dev)
    net.lod = 4
    o, d = workloads.camera_rays(128, 128, device=dev)
    tracer = SphereTracer(num_steps=128)
    full = tracer(net, o, d)
    g = torch.Generator().manual_seed(5)
    idx = torch.randperm(o.shape[0], generator=g)[:3000].to(dev)
    part = tracer(net, o[idx], d[idx])
    for name in ("x", "depth", "hit", "normal"):
        a = getattr(full, name)[idx]
        b = getattr(part, name)
        assert torch.equal(a, b), name


@pytest.mark.parametrize("form,return_dist", [("fp32", 0.0), ("mixed", 0.0), ("mixed", 1e-2)])
def test_trace_ragged_counts(dev, form, return_dist):
    """Ray counts below, at and just past the wave and claim sizes (some of
    the 8 ray ranges empty, partial 64-id chunks), and none at all: every
    ray's result equals its result in a full frame."""
    net = workloads.make_sphere_model(radius=0.6, num_lods=5, device=dev, seed=8)
    net.lod = 4
    o, d = workloads.camera_rays(96, 96, device=dev)
    kw = {} if form == "fp32" else {"return_dist": return_dist}
    tracer = SphereTracer(num_steps=128, form=form, **kw)
    full = tracer(net, o, d)
    assert int(full.hit.sum()) > 1000
    # a band of rays across the sphere's silhouette (hits, grazing rays and misses)
    start = 48 * 96 + 10
    for n in (0, 1, 7, 8, 9, 63, 64, 65, 127, 129, 1003):
        part = tracer(net, o[start:start + n], d[start:start + n])
        for name in ("x", "depth", "hit", "normal"):
            a = getattr(full, name)[start:start + n]
            b = getattr(part, name)
            assert a.shape == b.shape, (n, name)
            assert torch.equal(a, b), (n, name)


def test_normals_fd_vs_oracle(dev):
    from nglod_amd import _lib
    net = workloads.make_model(num_lods=5, seed=30, device="cpu")
    m = arrays(net)
    x = rand_points(5000, 31, -0.9, 0.9)
    want = oracle.normals(m, x, 4, "f32")
    netg = net.to(dev)
    xg = gpu(x, dev)
    out = torch.empty_like(xg)
    lib = _lib.load()
    _lib.check(lib.nglod_normals_fd(netg.descriptor(), _lib.ptr(xg), x.shape[0], 4, 0, _lib.ptr(out),
                                    _lib.stream_of(xg)), "nglod_normals_fd")
    got = out.cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-3)


def test_renderer_render_lookat(dev):
    net = workloads.make_sphere_model(radius=0.6, num_lods=5, device=dev)
    net.lod = 4
    tracer = SphereTracer(num_steps=128)
    r = Renderer(tracer, render_res=[40, 30], render_batch=0)
    rb = r.render_lookat(net, f=[0, 0, 3], t=[0, 0, 0], fov=30)
    assert rb.x.shape == (40, 30, 3) and rb.hit.shape == (40, 30, 1) and rb.relative_depth.shape == (40, 30, 1)
    r2 = Renderer(tracer, render_res=[40, 30], render_batch=333)
    rb2 = r2.render_lookat(net, f=[0, 0, 3], t=[0, 0, 0], fov=30)
    assert torch.equal(rb.depth, rb2.depth) and torch.equal(rb.hit, rb2.hit)
    assert isinstance(rb2, RenderBuffer)


@pytest.mark.parametrize("form", [_lib.FP32, _lib.MIXED])
@pytest.mark.parametrize("n", [1, 65, 4099, 200003])
def test_sdf_ordered_matches_input_order(dev, form, n):
    """The Morton-ordered batch query returns, in input order, exactly what the
    input-order query returns (points outside [-1, 1] included)."""
    net = workloads.make_model(num_lods=5, seed=11, device=dev)
    g = torch.Generator().manual_seed(n)
    x = (torch.rand(n, 3, generator=g) * 2.4 - 1.2).to(dev)
    a = net.sdf_hip(x, 4, form=form, ordered=False)
    b = net.sdf_hip(x, 4, form=form, ordered=True)
    assert torch.equal(a, b)


@pytest.mark.parametrize("form", ["fp32", "mixed"])
@pytest.mark.parametrize("n", [700, 5000])
def test_solo_step_same_bits(dev, form, n, monkeypatch):
    """A wave with nothing left to claim and few busy lanes evaluates their
    queries one at a time with the whole wave (FP32 passes); forced on for
    every such step (NGLOD_TRACE_SOLO_LANES=64: with these small inputs the
    ray feed empties at once, so nearly every FP32 evaluation takes that
    path) the trace is bit-identical to the wave-wide MFMA path (=0)."""
    for pos_invariant in (True, False):
        net = workloads.make_sphere_model(radius=0.6, num_lods=5, pos_invariant=pos_invariant, device="cpu").to(dev)
        net.lod = 4
        o, d = rand_rays(n, 31 + n)
        o[::7] *= 0.2  # some rays start inside the box
        outs = []
        for solo in ("0", "64"):
            monkeypatch.setenv("NGLOD_TRACE_SOLO_LANES", solo)
            steps = torch.zeros(n, dtype=torch.int32, device=dev)
            rb = SphereTracer(num_steps=128, form=form)(net, gpu(o, dev), gpu(d, dev), ray_steps=steps)
            torch.cuda.synchronize()
            outs.append([t.cpu() for t in (rb.x, rb.depth, rb.hit, rb.normal, steps)])
        for a, b in zip(*outs):
            assert torch.equal(a, b)
        assert outs[0][2].sum() > n // 10
