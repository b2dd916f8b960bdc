"""Post-trace shading on the GPU (shadow rays, AO, shadow map, matcap) against
the shading oracle (oracle/shade.py), through the C ABI wrappers."""
import numpy as np
import pytest
import torch

from nglod_amd import Renderer, SphereTracer, _lib, workloads
from nglod_amd.renderer import ambient_occlusion, matcap_shade, shadow_map, shadow_rays
from oracle import oracle, shade
from tests.helpers import arrays, rand_rays

pytestmark = pytest.mark.gpu

f32 = np.float32


def gpu(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _frame(net, dev, w=64, h=48, f=(0.0, 1.2, 3.0)):
    o, d = workloads.camera_rays(w, h, eye=f, device=dev)
    rb = SphereTracer(num_steps=128)(net, o, d)
    torch.cuda.synchronize()
    return o, d, rb


@pytest.mark.parametrize("n", [1, 7, 64, 1000, 4099])
def test_shadow_rays_bitexact(dev, n):
    """Every output equals the FP32 op-by-op restatement exactly (IEEE
    division and square root, no contraction), including rays parallel to the
    plane (d.y = 0: plane_t = inf or nan) and rays whose surface hit is nearer
    than the plane."""
    g = np.random.default_rng(n)
    o, d = rand_rays(n, n)
    d[::5, 1] = 0.0
    x = g.uniform(-1, 1, (n, 3)).astype(f32)
    t = g.uniform(0.5, 6.0, n).astype(f32)
    hit = g.random(n) < 0.5
    nrm = g.normal(size=(n, 3)).astype(f32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    noise = (0.01 * g.normal(size=(n, 3))).astype(f32)
    want = shade.shadow_rays(o, d, x, t, hit, nrm, -0.7, noise=noise)
    from nglod_amd.tracer import RenderBuffer
    rb = RenderBuffer(x=gpu(x, dev), depth=gpu(t.reshape(-1, 1), dev), hit=gpu(hit, dev), normal=gpu(nrm, dev))
    plane, so, sd, lh = shadow_rays(gpu(o, dev), gpu(d, dev), rb, -0.7, noise=gpu(noise, dev))
    torch.cuda.synchronize()
    assert np.array_equal(plane.cpu().numpy(), want["plane"])
    assert np.array_equal(rb.hit.cpu().numpy(), want["hit"])
    for got, key in ((rb.x, "x"), (rb.depth.reshape(-1), "t"), (rb.normal, "normal"), (so, "shadow_o"),
                     (sd, "shadow_d")):
        np.testing.assert_array_equal(got.cpu().numpy(), want[key], err_msg=key)
    assert np.array_equal(lh.cpu().numpy(), want["light_hit"])
    assert want["plane"].any() or n < 8


@pytest.mark.parametrize("pos_invariant", [True, False])
def test_ao_bitexact_given_the_gpu_sdf(dev, pos_invariant):
    """The AO kernel's arithmetic, fed the same FP32 SDF values (the batch
    query runs the same decoder, so its values are those the AO taps see),
    equals the oracle's formula bit for bit, for surface and plane hits."""
    net = workloads.make_sphere_model(radius=0.5, num_lods=3, pos_invariant=pos_invariant, seed=5,
                                      device="cpu").to(dev)
    net.lod = 2
    o, d = rand_rays(3000, 11)
    rb = SphereTracer(num_steps=96)(net, gpu(o, dev), gpu(d, dev))
    plane, _, _, _ = shadow_rays(gpu(o, dev), gpu(d, dev), rb, -0.5)
    ao = ambient_occlusion(net, rb.x, rb.normal, rb.hit, plane)
    torch.cuda.synchronize()
    x, nrm = rb.x.cpu().numpy(), rb.normal.cpu().numpy()
    hit, pl = rb.hit.cpu().numpy(), plane.cpu().numpy()
    assert hit.sum() > 100 and pl.sum() > 100

    def gpu_sdf(q):
        return net.sdf(gpu(q, dev), lod=2).detach().reshape(-1).cpu().numpy()

    want = shade.ambient_occlusion(gpu_sdf, x, nrm, hit, pl)
    np.testing.assert_array_equal(ao.reshape(-1).cpu().numpy(), want)


def test_ao_vs_oracle_sdf(dev):
    """End to end against the FP32 oracle SDF: the MFMA decoder's first-layer
    sum order differs from the oracle's by about 1e-6 per tap; 40 taps of
    3.5 * relu(.) and the square give at most 40 * 3.5 * 2 * 1e-6 = 3e-4."""
    net = workloads.make_sphere_model(radius=0.5, num_lods=4, device="cpu").to(dev)
    net.lod = 3
    o, d, rb = _frame(net, dev)
    ao = ambient_occlusion(net, rb.x, rb.normal, rb.hit)
    torch.cuda.synchronize()
    m = arrays(net.cpu())
    want = shade.ambient_occlusion(lambda q: oracle.sdf(m, q, 3), rb.x.cpu().numpy(), rb.normal.cpu().numpy(),
                                   rb.hit.cpu().numpy())
    got = ao.reshape(-1).cpu().numpy()
    assert rb.hit.sum().item() > 500
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-4)
    assert np.abs(got - want).mean() < 1e-5


@pytest.mark.parametrize("pos_invariant", [True, False])
def test_ao_mixed_vs_oracle_cumulative_grid(dev, pos_invariant):
    """The mixed form's AO taps evaluate the FP32 cumulative grid with the
    split-BF16 decoder (FP32-level): against the oracle's FP32 SDF on that
    grid, the bar of the FP32 form's AO (per-tap differences near 1e-6)."""
    net = workloads.make_sphere_model(radius=0.5, num_lods=4, pos_invariant=pos_invariant, seed=7,
                                      device="cpu").to(dev)
    net.lod = 3
    o, d, rb = _frame(net, dev)
    plane, _, _, _ = shadow_rays(o, d, rb, -0.45)
    ao = ambient_occlusion(net, rb.x, rb.normal, rb.hit, plane, form="mixed")
    torch.cuda.synchronize()
    m = arrays(net.cpu())
    cum = oracle.cumulative_grids(m, "f32")[3][0]
    mc = oracle.ModelArrays([cum], [m.res[3]], [m.w0[3]], [m.b0[3]], [m.w1[3]], [m.b1[3]], m.pos_invariant)
    want = shade.ambient_occlusion(lambda q: oracle.sdf(mc, q, 0), rb.x.cpu().numpy(), rb.normal.cpu().numpy(),
                                   rb.hit.cpu().numpy(), plane.cpu().numpy())
    got = ao.reshape(-1).cpu().numpy()
    assert rb.hit.sum().item() > 500 and plane.sum().item() > 100
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-4)
    assert np.abs(got - want).mean() < 1e-5


@pytest.mark.parametrize("wh", [(1, 1), (5, 3), (23, 17), (1280, 720)])
def test_shadow_map_vs_scipy(dev, wh):
    """clamp + gaussian_filter(sigma=2) equals SciPy's on the same mask: FP64
    accumulation in SciPy's order, FP32 after each axis (at most one FP32
    rounding step apart if the weights' last bit differs)."""
    w, h = wh
    g = np.random.default_rng(w * 31 + h)
    s = g.random(w * h) < 0.3
    lh = g.random(w * h) < 0.8
    got = shadow_map(gpu(s, dev), gpu(lh, dev), w, h).cpu().numpy()
    want = shade.shadow_map(s & lh, w, h)
    np.testing.assert_allclose(got, want, rtol=0, atol=1.2e-7)
    assert np.mean(got == want) > 0.999


@pytest.mark.parametrize("with_mm", [False, True])
def test_matcap_shade_vs_oracle(dev, with_mm):
    """Envmap in FP32 op by op, bilinear lookup in FP64 as SciPy's
    RegularGridInterpolator, FP32 / 255, segmentation, shadow and AO
    scaling: equal to the oracle (SciPy) up to the last FP64 bit of the
    lookup, i.e. at most one FP32 rounding step."""
    n = 20000
    g = np.random.default_rng(7)
    view = g.normal(size=(n, 3)).astype(f32)
    view /= np.linalg.norm(view, axis=1, keepdims=True)
    nrm = g.normal(size=(n, 3)).astype(f32)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm[::97] = 0.0  # misses carry zero normals into the envmap
    hit = g.random(n) < 0.7
    tex = g.integers(0, 256, size=(64, 48, 4)).astype(f32)
    smap = g.uniform(0.9, 1.0, n).astype(f32)
    ao = g.uniform(0.01, 1.0, n).astype(f32)
    mm = None
    if with_mm:
        q, _ = np.linalg.qr(g.normal(size=(3, 3)))
        mm = q.astype(f32)
    want_rgb, want_n = shade.matcap_shade(view, nrm, hit, tex, mm=mm, shadow=smap, ao=ao)
    n_gpu = gpu(nrm, dev)
    rgb = matcap_shade(gpu(view, dev), n_gpu, gpu(hit, dev), torch.from_numpy(tex),
                       mm=torch.from_numpy(mm) if mm is not None else None, shadow=gpu(smap, dev), ao=gpu(ao, dev))
    got = rgb.cpu().numpy()
    np.testing.assert_array_equal(n_gpu.cpu().numpy(), want_n)
    np.testing.assert_allclose(got, want_rgb, rtol=2.5e-7, atol=0)
    assert np.mean(got == want_rgb) > 0.999


def test_renderer_shadow_ao_matcap(dev):
    """Renderer with shadow + AO + matcap: deterministic for a seeded
    generator, equal to the functional pipeline on the same rays, and the
    plane pixels (no surface hit) are white before shadow and AO."""
    net = workloads.make_sphere_model(radius=0.5, num_lods=4, device="cpu").to(dev)
    net.lod = 3
    tex = torch.from_numpy(np.random.default_rng(1).integers(0, 256, size=(32, 32, 3)).astype(f32))
    tracer = SphereTracer(num_steps=128)

    def run():
        r = Renderer(tracer, render_res=[64, 48], shadow=True, ao=True, shading_mode="matcap", matcap=tex,
                     ground_height=-0.5, generator=torch.Generator(device=dev).manual_seed(3))
        return r.shade_tensor(net, f=[0, 1.0, 3.0], t=[0, 0, 0], fov=30)

    a, b = run(), run()
    assert a.rgb.shape == (64, 48, 3) and a.ao.shape == (64, 48, 1) and a.shadow.shape == (64, 48, 1)
    assert torch.equal(a.rgb, b.rgb) and torch.equal(a.ao, b.ao)
    hit = a.hit.reshape(-1).cpu().numpy()
    assert 200 < hit.sum() < hit.size
    rgb = a.rgb.reshape(-1, 3).cpu().numpy()
    assert np.all((rgb >= 0) & (rgb <= 1))
    # the sphere rests above the plane y = -0.5: some plane pixels are shadowed
    sh = a.shadow.reshape(-1).cpu().numpy()
    assert sh[~hit].any()
    img = Renderer(tracer, render_res=[64, 48], shadow=True, ao=True, matcap=tex, ground_height=-0.5,
                   generator=torch.Generator(device=dev).manual_seed(3)).shade_images(net, f=[0, 1.0, 3.0])
    assert img.rgb.shape == (48, 64, 3)
