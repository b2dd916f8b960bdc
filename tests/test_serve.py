"""Serving: ray-pool allocator and request streams on the CPU; on the GPU,
camera rays vs look_at, admission by pool bytes, latency statistics, and the
same result for a request alone and in the mix."""
import math

import numpy as np
import pytest
import torch

from nglod_amd import workloads
from nglod_amd.renderer import look_at
from nglod_amd.serve import RayPool, RenderRequest, TraceServer, request_stream


def test_ray_pool_first_fit_and_merge():
    p = RayPool(4096, "cpu")
    a = p.alloc(1000)
    b = p.alloc(1000)
    c = p.alloc(1000)
    assert (a, b, c) == (0, 1024, 2048)
    assert p.alloc(2000) is None  # 1024 left
    p.free(b)
    assert p.alloc(1024) == 1024  # first fit reuses the hole
    p.free(a)
    p.free(1024)
    p.free(c)
    assert p.free_list == [(0, 4096)] and p.bytes_in_use == 0
    assert p.alloc(4096) == 0


def test_request_stream_is_seeded_and_poisson():
    a = request_stream(2000, 100.0, 8, seed=3)
    b = request_stream(2000, 100.0, 8, seed=3)
    assert [(r.model, r.width, r.lod, r.eye, r.arrival) for r in a] == \
        [(r.model, r.width, r.lod, r.eye, r.arrival) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 0.01) < 0.001  # exponential with mean 1/rate
    assert abs(np.median(gaps) - 0.01 * math.log(2)) < 0.001
    radii = np.array([np.linalg.norm(r.eye) for r in a])
    assert radii.min() >= 2.5 and radii.max() <= 4.0
    assert {r.model for r in a} == set(range(8)) and {r.lod for r in a} == set(range(5))
    burst = request_stream(10, math.inf, 2, seed=0)
    assert all(r.arrival == 0.0 for r in burst)


@pytest.mark.gpu
def test_camera_rays_match_look_at(dev):
    from nglod_amd import _lib
    import ctypes
    lib = _lib.load()
    for (w, h), eye, at, fov in [((64, 48), (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 30.0),
                                 ((37, 91), (1.5, -2.0, 2.5), (0.1, 0.2, -0.3), 45.0),
                                 ((128, 128), (-3.0, 1.0, 0.5), (0.0, 0.0, 0.0), 60.0)]:
        n = w * h
        o = torch.empty(n, 3, device=dev)
        d = torch.empty(n, 3, device=dev)
        _lib.check(lib.nglod_camera_rays((ctypes.c_float * 3)(*eye), (ctypes.c_float * 3)(*at), fov, w, h, 0,
                                         _lib.ptr(o), _lib.ptr(d), _lib.stream_of(o)), "camera")
        ro, rd = look_at(list(eye), list(at), w, h, mode="persp", fov=fov, device=dev)
        torch.cuda.synchronize()
        assert torch.equal(o, ro)
        assert (d - rd).abs().max().item() < 2e-6


@pytest.fixture(scope="module")
def models(dev):
    return [workloads.make_sphere_model(radius=r, num_lods=5, device=dev, seed=i)
            for i, r in enumerate((0.5, 0.6, 0.7))]


def _reqs():
    out = []
    for i in range(10):
        w, h = [(64, 64), (128, 96), (256, 256)][i % 3]
        out.append(RenderRequest(model=i % 3, width=w, height=h, lod=i % 5, eye=(0.3 * i - 1.5, 0.5, 3.0), rid=i))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["pooled", "streams"])
def test_same_result_alone_and_in_the_mix(models, mode):
    server = TraceServer(models, pool_bytes=64 << 20, slots=4, keep_outputs=True, mode=mode)
    reqs = _reqs()
    mix = server.run(reqs)
    assert len(mix.results) == len(reqs) and mix.peak_inflight > 1
    for req, res in zip(reqs, mix.results):
        alone = server.run([RenderRequest(**{**req.__dict__, "arrival": 0.0})]).results[0]
        for k in ("x", "depth", "hit", "normal"):
            assert torch.equal(res.outputs[k], alone.outputs[k]), (req.rid, k)
    # and the same as the tracer called directly on those rays
    from nglod_amd import SphereTracer
    r0 = mix.results[4]
    req = reqs[4]
    net = models[req.model]
    net.lod = req.lod
    rb = SphereTracer(num_steps=128, form="mixed")(net, r0.outputs["ray_o"], r0.outputs["ray_d"])
    torch.cuda.synchronize()
    assert torch.equal(rb.depth, r0.outputs["depth"]) and torch.equal(rb.hit, r0.outputs["hit"])
    assert rb.hit.sum().item() > 100


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["pooled", "streams"])
def test_admission_is_capped_by_pool_bytes(models, mode):
    probe = TraceServer(models, pool_bytes=1 << 20, slots=4, mode=mode)
    one = probe.request_bytes(RenderRequest(0, 256, 256, 4, (0.0, 0.0, 3.0)))
    server = TraceServer(models, pool_bytes=one + one // 2, slots=4, mode=mode)
    reqs = [RenderRequest(i % 3, 256, 256, 4, (0.0, 0.0, 3.0), rid=i) for i in range(6)]
    st = server.run(reqs)
    assert len(st.results) == 6
    assert st.peak_inflight == 1 and st.peak_pool_bytes <= server.pool_nbytes
    with pytest.raises(ValueError):
        server.run([RenderRequest(0, 1280, 720, 4, (0.0, 0.0, 3.0))])


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["pooled", "streams"])
def test_poisson_stream_latencies(models, mode):
    server = TraceServer(models, pool_bytes=256 << 20, slots=4, mode=mode)
    reqs = request_stream(40, 200.0, len(models), seed=1, resolutions=((64, 64), (256, 256), (640, 480)))
    st = server.run(reqs)
    s = st.summary()
    assert s["requests"] == 40 and s["rays"] == sum(r.rays for r in reqs)
    assert 0 < s["p50_ms"] <= s["p99_ms"]
    assert all(r.finish >= r.start >= r.arrival - 1e-6 for r in st.results)


@pytest.fixture(scope="module")
def fitted_models(dev):
    """The serving workload's 8 resident models: seeds 0-7, each a 5-LOD
    OctreeSDF fitted 50 Adam steps to its own displaced icosphere."""
    from nglod_amd import trainer
    return [trainer.fitted_icosphere_model(seed=s, device=dev, iterations=50)[0] for s in range(8)]


@pytest.mark.gpu
def test_serving_workload_full_size(fitted_models):
    """The serving configuration at its own size: 8 fitted models, requests
    over all five resolutions up to 1920x1080 and LODs 0-4 with orbit
    cameras, arriving together so that they overlap on the device.  Every
    request's outputs are bit-identical alone and in the mix, and a
    1280x720 request's depths agree with the FP64 oracle on a strided
    sample (hit agreement >= 0.99, mean hit-depth error <= 1e-3)."""
    from nglod_amd.serve import RESOLUTIONS, orbit_eye
    from oracle import oracle
    from tests.helpers import arrays

    rng = np.random.default_rng(7)
    reqs = []
    for i in range(10):
        w, h = RESOLUTIONS[i % 5]
        reqs.append(RenderRequest(model=i % 8, width=w, height=h, lod=(3 * i) % 5, eye=orbit_eye(rng), rid=i))
    server = TraceServer(fitted_models, pool_bytes=2 << 30, slots=2, keep_outputs=True, mode="pooled")
    mix = server.run(reqs)
    assert len(mix.results) == len(reqs) and mix.peak_inflight > 1 and mix.launches < len(reqs)
    for req, res in zip(reqs, mix.results):
        alone = server.run([RenderRequest(**{**req.__dict__, "arrival": 0.0})]).results[0]
        for k in ("x", "depth", "hit", "normal"):
            assert torch.equal(res.outputs[k], alone.outputs[k]), (req.rid, req.width, k)
    big = [r for r in reqs if (r.width, r.height) == (1280, 720)][0]
    out = mix.results[big.rid].outputs
    assert out["hit"].sum().item() > 1000
    idx = np.arange(0, big.rays, 53)
    m = arrays(fitted_models[big.model])
    ref = oracle.trace(m, out["ray_o"].cpu().numpy()[idx], out["ray_d"].cpu().numpy()[idx], big.lod,
                       num_steps=128, prec="f64")
    hit = out["hit"].cpu().numpy()[idx]
    both = hit & ref["hit"]
    assert (hit == ref["hit"]).mean() >= 0.99
    assert both.sum() > 100
    derr = np.abs(out["depth"][:, 0].cpu().numpy()[idx][both] - ref["t"][both])
    assert derr.mean() <= 1e-3, derr.mean()
    # the pooled batch gives the same bits as one trace call per request on its own stream
    streams = TraceServer(fitted_models, pool_bytes=2 << 30, slots=4, keep_outputs=True, mode="streams")
    other = streams.run(reqs)
    for req, a, b in zip(reqs, mix.results, other.results):
        for k in ("x", "depth", "hit", "normal"):
            assert torch.equal(a.outputs[k], b.outputs[k]), (req.rid, k)
