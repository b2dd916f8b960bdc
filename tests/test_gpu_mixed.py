"""Precision form (FP16 cumulative grids, BF16 MFMA decoder, FP32
near-surface re-evaluation and FP32 normals on the FP32 cumulative grids)
against the FP64 CPU oracle; the cumulative grids bit-exact against the
oracle's.

Bars: SDF max abs error <= 3e-2 and mean <= 3e-3 away from the surface,
mean hit-depth error <= 1e-3 over rays both forms call hits, hit flags
equal on >= 99% of rays, median normal angle error <= 0.5 degree."""
import numpy as np
import pytest
import torch

from nglod_amd import SphereTracer, _lib, workloads
from oracle import oracle
from tests.helpers import arrays, rand_points, rand_rays

pytestmark = pytest.mark.gpu


def gpu(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("pos_invariant", [True, False])
def test_sdf_mixed_vs_fp64(dev, pos_invariant):
    net = workloads.make_sphere_model(radius=0.55, num_lods=5, pos_invariant=pos_invariant, device="cpu", seed=2,
                                      noise=5e-3)
    m = arrays(net)
    x = rand_points(40000, 41)
    netg = net.to(dev)
    with torch.no_grad():
        for lod in (0, 2, 4):
            got = netg.sdf_hip(gpu(x, dev), lod, form=_lib.MIXED)[:, 0].cpu().numpy().astype(np.float64)
            want = oracle.sdf(m, x, lod, "f64")
            err = np.abs(got - want)
            assert err.max() < 3e-2, (lod, err.max())
            assert err.mean() < 3e-3, (lod, err.mean())


def test_sdf_mixed_random_model_decoder_layout(dev):
    """A random model exercises every W0 fragment, b0 and w1 row: a wrong MFMA
    operand or output layout would show as O(1) relative errors."""
    net = workloads.make_model(num_lods=5, seed=42, device="cpu")
    with torch.no_grad():
        for f in net.features:
            f.fm.mul_(30.0)  # features O(0.3): decoder output well above bf16 noise
    m = arrays(net)
    x = rand_points(20000, 43)
    netg = net.to(dev)
    with torch.no_grad():
        got = netg.sdf_hip(gpu(x, dev), 4, form=_lib.MIXED)[:, 0].cpu().numpy().astype(np.float64)
    want = oracle.sdf(m, x, 4, "f64")
    scale = np.abs(want).mean()
    assert scale > 1e-2
    assert np.abs(got - want).mean() < 0.02 * scale


@pytest.mark.parametrize("pos_invariant", [True, False])
def test_pack_mixed_cumulative_grids_bitexact(dev, pos_invariant):
    """nglod_pack_mixed's cumulative grids (FP32 and FP16) equal the oracle's
    bit for bit at the reference's resolutions 4..64."""
    net = workloads.make_model(num_lods=5, pos_invariant=pos_invariant, seed=44, device="cpu")
    with torch.no_grad():
        for f in net.features:
            f.fm.mul_(20.0)
    want = oracle.cumulative_grids(arrays(net), "f32")
    netg = net.to(dev)
    grids_c, grids_h, _ = netg.mixed_buffers()
    torch.cuda.synchronize()
    for L, (w32, w16) in enumerate(want):
        assert np.array_equal(grids_c[L].cpu().numpy(), w32), L
        assert np.array_equal(grids_h[L].cpu().numpy().view(np.uint16), w16.view(np.uint16)), L


def test_pack_mixed_refuses_grids_that_do_not_nest(dev):
    net = workloads.make_model(num_lods=2, seed=45, device=dev)
    m = net.descriptor(_lib.FP32)
    bufs = [torch.empty(6 ** 3 * 32, dtype=torch.float32, device=dev) for _ in range(2)]
    half = [torch.empty(6 ** 3 * 32, dtype=torch.float16, device=dev) for _ in range(2)]
    w0s = [torch.empty(128 * 32, dtype=torch.int16, device=dev) for _ in range(2)]
    for i in range(2):
        m.grid_cum[i] = bufs[i].data_ptr()
        m.grid_h[i] = half[i].data_ptr()
        m.w0_bf16[i] = w0s[i].data_ptr()
    m.res[0], m.res[1] = 3, 4  # 3 does not divide 4 (grids of at least that size back the pointers)
    lib = _lib.load()
    assert lib.nglod_pack_mixed(m, _lib.stream_of(bufs[0])) == -2  # NGLOD_ERR_UNSUPPORTED


def _angle_deg(a, b):
    c = (a * b).sum(1) / np.maximum(np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1), 1e-12)
    return np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))


@pytest.mark.parametrize("lod,near_dist,return_dist", [(2, 1e-2, 0.0), (4, 1e-2, 0.0), (4, 0.0, 0.0),
                                                      (4, 1e-2, 1e-2)])
def test_trace_mixed_vs_fp64(dev, lod, near_dist, return_dist):
    net = workloads.make_sphere_model(radius=0.6, num_lods=5, device="cpu", seed=4, noise=5e-3)
    m = arrays(net)
    o, d = workloads.camera_rays(160, 120, eye=(0.4, 0.8, 2.8), device="cpu")
    o, d = o.numpy(), d.numpy()
    ref = oracle.trace(m, o, d, lod, num_steps=128, prec="f64")
    netg = net.to(dev)
    netg.lod = lod
    tr = SphereTracer(num_steps=128, form="mixed", near_dist=near_dist, return_dist=return_dist)
    rb = tr(netg, gpu(o, dev), gpu(d, dev))
    torch.cuda.synchronize()
    hit = rb.hit.cpu().numpy()
    assert (hit == ref["hit"]).mean() >= 0.99
    both = hit & ref["hit"]
    assert both.sum() > 2000
    derr = np.abs(rb.depth[:, 0].cpu().numpy()[both] - ref["t"][both])
    assert derr.mean() <= 1e-3, derr.mean()
    ang = _angle_deg(rb.normal.cpu().numpy()[both], ref["normal"][both])
    assert np.median(ang) <= 0.5, np.median(ang)
    ev = tr.last_evals.cpu().numpy()
    assert ev[1] > 0 and ev[0] > 0  # both passes ran


def test_trace_mixed_normals_are_fp32_taps(dev):
    """The mixed trace's normals are central differences of FP32-accurate
    evaluations on the FP32 cumulative grid (the continuation pass's decoder,
    BF16 products of exact three-term splits): against the FP64 oracle's taps
    at the same hit points on that grid the angles stay at FP32 rounding level.
    A BF16-level decoder or gather in those taps gives tenths of a degree."""
    net = workloads.make_sphere_model(radius=0.6, num_lods=5, device="cpu", seed=4, noise=5e-3)
    m = arrays(net)
    L = 4
    cum = oracle.cumulative_grids(m, "f32")[L][0]
    mc = oracle.ModelArrays([cum], [m.res[L]], [m.w0[L]], [m.b0[L]], [m.w1[L]], [m.b1[L]], m.pos_invariant)
    o, d = workloads.camera_rays(160, 120, eye=(0.4, 0.8, 2.8), device="cpu")
    netg = net.to(dev)
    netg.lod = L
    rb = SphereTracer(num_steps=128, form="mixed")(netg, gpu(o.numpy(), dev), gpu(d.numpy(), dev))
    hit = rb.hit.cpu().numpy()
    assert hit.sum() > 2000
    want = oracle.normals(mc, rb.x.cpu().numpy()[hit], 0, prec="f64")
    ang = _angle_deg(rb.normal.cpu().numpy()[hit], want)
    assert np.median(ang) < 5e-3 and np.percentile(ang, 99) < 5e-2, (np.median(ang), np.percentile(ang, 99))


def test_trace_mixed_random_rays_inside_and_outside(dev):
    net = workloads.make_sphere_model(radius=0.45, num_lods=5, pos_invariant=False, device="cpu", seed=6)
    m = arrays(net)
    o, d = rand_rays(20000, 44, dist=(0.3, 4.0))
    ref = oracle.trace(m, o, d, 4, num_steps=100, prec="f64")
    netg = net.to(dev)
    netg.lod = 4
    rb = SphereTracer(num_steps=100, form="mixed")(netg, gpu(o, dev), gpu(d, dev))
    hit = rb.hit.cpu().numpy()
    assert (hit == ref["hit"]).mean() >= 0.99
    both = hit & ref["hit"]
    derr = np.abs(rb.depth[:, 0].cpu().numpy()[both] - ref["t"][both])
    assert derr.mean() <= 1e-3
    assert np.all(rb.normal.cpu().numpy()[~hit] == 0)


def test_trace_mixed_same_alone_and_in_the_mix(dev):
    net = workloads.make_sphere_model(radius=0.6, num_lods=5, device=dev)
    net.lod = 4
    o, d = workloads.camera_rays(128, 128, device=dev)
    tracer = SphereTracer(num_steps=128, form="mixed")
    full = tracer(net, o, d)
    g = torch.Generator().manual_seed(9)
    idx = torch.randperm(o.shape[0], generator=g)[:2500].to(dev)
    part = tracer(net, o[idx], d[idx])
    for name in ("x", "depth", "hit", "normal"):
        assert torch.equal(getattr(full, name)[idx], getattr(part, name)), name
