"""Pins the shading oracle (oracle/shade.py) with closed forms and library
identities, on the CPU.  Each case is chosen so that a plausible slip in the
restatement (a flipped tap direction, a dropped square, a wrong constant, a
transposed texture axis) fails it."""
import numpy as np
from scipy.ndimage import gaussian_filter

from oracle import shade

f32 = np.float32


def _sphere_sdf(radius):
    return lambda q: (np.linalg.norm(q.astype(np.float64), axis=1) - radius).astype(f32)


def test_ao_convex_sphere_is_unoccluded():
    """On a sphere, sdf(x + n d) = d: every relu argument is -0.0015 and
    AO = 1 exactly; taps taken against the normal (x - n d) would give
    sdf = -d and darken every point."""
    g = np.random.default_rng(0)
    n = g.normal(size=(500, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    x = (0.5 * n).astype(f32)
    ao = shade.ambient_occlusion(_sphere_sdf(0.5), x, n.astype(f32), np.ones(500, bool))
    assert np.all(ao == 1.0)
    flipped = shade.ambient_occlusion(_sphere_sdf(0.5), x, (-n).astype(f32), np.ones(500, bool))
    assert np.all(flipped < 0.9)


def test_ao_zero_sdf_closed_form():
    """sdf = 0 everywhere: acc = sum_k 3.5 * relu(d_k - 0.0015) with
    d_k = 0.025 ((k+1)/41)^1.6 (renderer.py:193) and AO = clamp(1 - acc,
    0.1, 1)^2, summed here in FP64 from the published constants."""
    dk = np.array([0.025 * ((k + 1) / 41.0) ** 1.6 for k in range(40)])
    acc = np.sum(3.5 * np.maximum(dk - 0.0015, 0.0))
    want = np.clip(1.0 - acc, 0.1, 1.0) ** 2
    x = np.zeros((3, 3), f32)
    nrm = np.tile(np.array([[0, 1, 0]], f32), (3, 1))
    hit = np.array([True, False, False])
    plane = np.array([False, True, False])
    ao = shade.ambient_occlusion(lambda q: np.zeros(len(q), f32), x, nrm, hit, plane)
    np.testing.assert_allclose(ao[:2], want, rtol=2e-6)
    assert ao[2] == 1.0  # neither hit: acc stays 0
    # a plane hit takes min(sdf, d_k), which leaves relu(d_k - r - 0.0015)
    # unchanged (for sdf > d_k both are relu(-0.0015) = 0): same AO as a hit
    for c in (-0.5, 0.004, 0.02, 5.0):
        ao_c = shade.ambient_occlusion(lambda q: np.full(len(q), c, f32), x, nrm, hit, plane)
        assert ao_c[0] == ao_c[1]


def test_shadow_rays_plane_geometry():
    """A ray straight down from y = 2 meets the plane y = -1 at t = 3; a ray
    going up never does; a plane behind the surface hit is ignored."""
    o = np.array([[0.2, 2.0, 0.3], [0.0, 0.0, 0.0], [0.1, 2.0, 0.0]], f32)
    d = np.array([[0, -1, 0], [0, 1, 0], [0, -1, 0]], f32)
    x = np.array([[0.2, 0.5, 0.3], [0, 0.9, 0], [0.1, 1.0, 0.0]], f32)
    t = np.array([10.0, 0.9, 1.0], f32)  # ray 2 hits the surface before the plane
    hit = np.array([False, True, True])
    nrm = np.array([[0, 0, 0], [0, -1, 0], [0, 1, 0]], f32)
    r = shade.shadow_rays(o, d, x, t, hit, nrm, -1.0)
    assert r["plane"].tolist() == [True, False, False]
    assert r["t"][0] == 3.0 and np.all(r["x"][0] == np.array([0.2, -1.0, 0.3], f32))
    assert np.all(r["normal"][0] == [0, 1, 0]) and not r["hit"][0]
    assert r["hit"][1] and r["hit"][2]
    # shadow ray from 0.1 above the plane point, towards the light (-1.5, 4.5, -1.5)
    so = np.array([0.2, -0.9, 0.3])
    v = np.array([-1.5, 4.5, -1.5]) - so
    np.testing.assert_allclose(r["shadow_d"][0], v / np.linalg.norm(v), rtol=1e-6)
    assert r["light_hit"][0]          # light is above the plane
    assert not r["light_hit"][1]      # normal (0, -1, 0) faces away from it
    assert np.allclose(np.linalg.norm(r["shadow_d"], axis=1), 1.0, atol=1e-6)


def test_envmap_facing_normal_is_texture_centre():
    """A normal facing the view ray reflects it straight back: vN = (0.5, 0.5).
    For n = (0.6, 0, 0.8): r = (-0.96, 0, -0.28) - (0, 0, 1), m = 2 * 1.6, so
    vN = 1 - (r.xy / m + 0.5) = (0.8, 0.5)."""
    view = np.array([[0, 0, -1], [0, 0, -1]], f32)
    nrm = np.array([[0, 0, 1], [0.6, 0, 0.8]], f32)
    vn = shade.spherical_envmap(view, nrm)
    assert np.all(vn[0] == 0.5)
    np.testing.assert_allclose(vn[1], [0.8, 0.5], rtol=0, atol=1e-6)


def test_matcap_lookup_nodes_and_axes():
    """At grid nodes the bilinear lookup returns the texel; u indexes the
    first (x) axis of the transposed texture."""
    tw, th = 5, 4
    tex = np.zeros((tw, th, 3), np.float64)
    tex[:, :, 0] = np.arange(tw)[:, None] * 10.0
    tex[:, :, 1] = np.arange(th)[None, :] * 20.0
    tex[:, :, 2] = 255.0
    view = np.array([[0, 0, -1]], f32)
    nrm = np.array([[0, 0, 1]], f32)  # vN = (0.5, 0.5): node 2 on x, between nodes 1 and 2 on y
    rgb, _ = shade.matcap_shade(view, nrm, np.array([True]), tex)
    np.testing.assert_allclose(rgb[0], [20.0 / 255.0, 30.0 / 255.0, 1.0], rtol=1e-6)
    rgb2, n2 = shade.matcap_shade(view, nrm, np.array([False]), tex, shadow=np.array([0.5], f32),
                                  ao=np.array([0.25], f32))
    assert np.all(rgb2[0] == f32(0.125)) and np.all(n2[0] == 1.0)


def test_shadow_map_is_scipy_gaussian():
    """gaussian_filter(sigma=2) equals a direct FP64 17-tap convolution with
    half-sample symmetric padding (scipy 'reflect' = numpy 'symmetric')."""
    g = np.random.default_rng(3)
    s = g.random((23, 17)) < 0.3
    m = shade.shadow_map(s.reshape(-1), 23, 17)
    base = np.clip((1.0 - s.astype(np.float64)) + 0.9, 0, 1).astype(f32).astype(np.float64)
    x = np.arange(-8, 9)
    w = np.exp(-0.5 / 4.0 * x ** 2)
    w /= w.sum()
    a = np.pad(base, ((8, 8), (0, 0)), mode="symmetric")
    a = np.stack([np.sum(a[i:i + 17, :] * w[:, None], axis=0) for i in range(23)]).astype(f32).astype(np.float64)
    b = np.pad(a, ((0, 0), (8, 8)), mode="symmetric")
    b = np.stack([np.sum(b[:, j:j + 17] * w[None, :], axis=1) for j in range(17)], axis=1)
    np.testing.assert_allclose(m, b, rtol=0, atol=2e-7)
    assert m.dtype == np.float32
    # clamp(1 - s + 0.9): 0.9 where shadowed, 1 elsewhere, before the blur
    np.testing.assert_allclose(gaussian_filter(np.full((5, 5), 0.9, f32), sigma=2), 0.9, rtol=1e-6)
