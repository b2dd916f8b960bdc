"""Pins the oracle's training-target path: the restated mesh2sdf against an
independent closed-form closest point (Ericson) and the generalized winding
number; the samplers against SplitMix64's published outputs and the
statistics the sampling formulas imply."""
import numpy as np

from nglod_amd import mesh
from oracle import oracle


def shape(seed=0, level=3):
    V, F = mesh.icosphere(level)
    V = V * (1.0 + mesh.radial_noise(V, seed, 0.1))[:, None]
    V = mesh.normalize(V).astype(np.float32)
    return V, F, mesh.triangles(V, F)


def test_icosphere_counts_and_closed():
    V, F = mesh.icosphere(4)
    assert V.shape == (2562, 3) and F.shape == (5120, 3)
    # closed 2-manifold: every edge shared by exactly two faces, Euler V - E + F = 2
    e = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2)
    assert V.shape[0] - cnt.size + F.shape[0] == 2


def test_normalize_unit_radius_centered():
    V, F = mesh.displaced_icosphere(3)
    r = np.sqrt((V.astype(np.float64) ** 2).sum(1))
    assert abs(r.max() - 1.0) < 1e-6
    np.testing.assert_allclose((V.max(0) + V.min(0)) / 2, 0, atol=1e-6)


def test_unsigned_distance_matches_closed_form():
    V, F, T = shape()
    rng = np.random.default_rng(0)
    pts = rng.uniform(-1.3, 1.3, size=(60, 3))
    got = np.abs(oracle.mesh2sdf(pts, T))
    for i, p in enumerate(pts):
        d2 = min(oracle.point_tri_dist2(p, t[0], t[1], t[2]) for t in T.astype(np.float64))
        assert abs(got[i] - np.sqrt(d2)) < 1e-9 * max(1.0, np.sqrt(d2))


def test_sign_matches_winding_number():
    V, F, T = shape()
    rng = np.random.default_rng(1)
    pts = rng.uniform(-1.1, 1.1, size=(300, 3))
    sd = oracle.mesh2sdf(pts, T)
    w = np.array([oracle.winding(p, T) for p in pts])
    assert np.all(np.abs(np.abs(w) - np.round(np.abs(w))) < 1e-6)  # closed mesh: integer winding
    inside = np.abs(w) > 0.5
    assert inside.any() and (~inside).any()
    far = np.abs(sd) > 1e-3
    np.testing.assert_array_equal((sd < 0)[far], inside[far])
    assert oracle.mesh2sdf(np.zeros((1, 3)), T)[0] < 0  # the centre is inside


def test_splitmix64_published_vectors():
    # first outputs of SplitMix64 seeded with 0 (state += golden gamma, then mix)
    assert oracle.splitmix64(0) == 0xE220A8397B1DCDAF
    assert oracle.splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_uniform_sampler_statistics():
    x = oracle.sample_uniform(7, 200000)
    assert x.min() >= -1.0 and x.max() < 1.0
    np.testing.assert_allclose(x.mean(0), 0.0, atol=0.01)
    np.testing.assert_allclose(x.var(0), 1.0 / 3.0, atol=0.01)
    assert not np.array_equal(oracle.sample_uniform(8, 10), x[:10])


def test_surface_sampler_area_weighting_and_noise():
    V, F, T = shape(level=2)
    cdf = mesh.face_cdf(T)
    p, face = oracle.sample_surface(T, cdf, 3, 200000, 0.0)
    # points lie on their face
    d = np.abs(oracle.mesh2sdf(p[:200], T))
    assert d.max() < 1e-6
    # face frequencies follow the areas
    area = np.diff(np.concatenate([[0.0], cdf]))
    freq = np.bincount(face, minlength=T.shape[0]) / face.size
    assert np.abs(freq - area).max() < 0.01
    # near-surface samples: offset ~ N(0, sigma^2) per axis
    p2, face2 = oracle.sample_surface(T, cdf, 3, 200000, 0.01)
    np.testing.assert_array_equal(face2, face)  # same draws, plus an offset
    off = p2 - p
    np.testing.assert_allclose(off.std(0), 0.01, rtol=0.02)
    np.testing.assert_allclose(off.mean(0), 0.0, atol=2e-4)


def _sphere_moments_ok(v, tol=5e-3):
    """What uniformity on S^2 fixes (Archimedes' hat-box theorem): each
    coordinate is uniform on [-1, 1], so E[c] = 0 and E[c^2] = 1/3; the
    coordinates are uncorrelated; the azimuth is uniform on [0, 2 pi).
    Returns the failed checks."""
    v = v.astype(np.float64)
    bad = []
    if np.abs(np.linalg.norm(v, axis=1) - 1.0).max() > 1e-6:
        bad.append("norm")
    for k in range(3):
        if abs(v[:, k].mean()) > tol:
            bad.append(f"E[{k}]")
        if abs((v[:, k] ** 2).mean() - 1.0 / 3.0) > tol:
            bad.append(f"E[{k}^2]")
        # flat histogram of each coordinate over [-1, 1] (20 bins)
        h = np.histogram(v[:, k], bins=20, range=(-1.0, 1.0))[0] / v.shape[0]
        if np.abs(h - 0.05).max() > 2e-3:
            bad.append(f"hist[{k}]")
    for a, b in ((0, 1), (0, 2), (1, 2)):
        if abs((v[:, a] * v[:, b]).mean()) > tol:
            bad.append(f"E[{a}{b}]")
    phi = np.mod(np.arctan2(v[:, 1], v[:, 0]), 2.0 * np.pi)
    h = np.histogram(phi, bins=24, range=(0.0, 2.0 * np.pi))[0] / v.shape[0]
    if np.abs(h - 1.0 / 24.0).max() > 2e-3:
        bad.append("hist[phi]")
    return bad


def test_direction_sampler_is_uniform_on_the_sphere():
    """orc_sample_directions (sdf-net/lib/geoutils.py:31-40: z = 1 - 2u,
    phi = 2 pi v) on 2^20 samples meets every moment and histogram that a
    uniform law on S^2 fixes, and two plausible mistakes built from the same
    uniform draws fail them: a hemisphere (z = 1 - u) and a half turn
    (phi = pi v)."""
    n = 1 << 20
    assert _sphere_moments_ok(oracle.sample_directions(11, n)) == []
    # the stream's draws 0 and 1, which the direction sampler consumes
    u = (oracle.sample_uniform(11, n).astype(np.float64) + 1.0) * 0.5

    def dirs(z, phi):
        r = np.sqrt(np.maximum(1.0 - z * z, 0.0))
        return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)

    good = dirs(1.0 - 2.0 * u[:, 0], 2.0 * np.pi * u[:, 1])
    np.testing.assert_allclose(good, oracle.sample_directions(11, n), rtol=0, atol=1e-6)
    hemisphere = _sphere_moments_ok(dirs(1.0 - u[:, 0], 2.0 * np.pi * u[:, 1]))
    half_turn = _sphere_moments_ok(dirs(1.0 - 2.0 * u[:, 0], np.pi * u[:, 1]))
    assert "E[2]" in hemisphere and "hist[2]" in hemisphere
    assert "E[1]" in half_turn and "hist[phi]" in half_turn
