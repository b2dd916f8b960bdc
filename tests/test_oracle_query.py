"""Pins the CPU oracle's query path (gather, decoder, OctreeSDF.sdf) to the
library routines the reference calls (F.grid_sample, F.linear) and to
closed-form answers of trilinear interpolation."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import oracle
from oracle.torch_ref import TorchOctreeSDF
from tests.helpers import arrays, rand_points, small_model


def torch_sdf_f64(net, x, lod):
    ref = TorchOctreeSDF.from_module(net, dtype=torch.float64)
    return ref.sdf(torch.from_numpy(x).double(), lod=lod)[:, 0].numpy()


@pytest.mark.parametrize("pos_invariant", [True, False])
@pytest.mark.parametrize("joint", [False, True])
def test_sdf_f64_matches_grid_sample_and_linear(pos_invariant, joint):
    net = small_model(num_lods=3, pos_invariant=pos_invariant, joint=joint, seed=3)
    m = arrays(net)
    x = rand_points(2000, 11, -1.2, 1.2)  # includes border-clamped points
    for lod in range(3):
        got = oracle.sdf(m, x, lod, "f64")
        want = torch_sdf_f64(net, x, lod)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_sdf_f32_close_to_torch_f32():
    net = small_model(num_lods=3, seed=4)
    m = arrays(net)
    x = rand_points(4000, 12)
    ref = TorchOctreeSDF.from_module(net, dtype=torch.float32)
    for lod in range(3):
        got = oracle.sdf(m, x, lod, "f32")
        want = ref.sdf(torch.from_numpy(x), lod=lod)[:, 0].numpy()
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-6)


def test_gather_f32_bitexact_vs_grid_sample():
    """A decoder that reads one channel exactly: out = feature_0 (pos_invariant).
    The oracle's FP32 gather follows grid_sample's op order, so the sampled
    value must match torch's CPU kernel bit for bit."""
    net = small_model(num_lods=1, seed=5)
    with torch.no_grad():
        l = net.louts[0]
        l[0].weight.zero_()
        l[0].weight[0, 0] = 1.0
        l[0].bias.zero_()
        l[0].bias[0] = 4.0
        l[2].weight.zero_()
        l[2].weight[0, 0] = 1.0
        l[2].bias.fill_(-4.0)
    m = arrays(net)
    x = rand_points(3000, 13, -1.1, 1.1)
    got = oracle.sdf(m, x, 0, "f32")
    fm = net.features[0].fm.detach().contiguous()
    g = F.grid_sample(fm, torch.from_numpy(x).reshape(1, -1, 1, 1, 3), align_corners=True, padding_mode="border")
    want = ((g[0, 0, :, 0, 0] + 4.0) - 4.0).numpy()
    np.testing.assert_array_equal(got, want)


def _affine_model(res, A, b):
    net = small_model(num_lods=1, seed=6)
    c = np.linspace(-1.0, 1.0, res + 1)
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    with torch.no_grad():
        fm = net.features[0].fm
        for ch in range(fm.shape[1]):
            val = A[ch, 0] * x + A[ch, 1] * y + A[ch, 2] * z + b[ch]
            fm[0, ch].copy_(torch.from_numpy(val))
    return net


def test_affine_field_reproduced_exactly():
    """Trilinear interpolation reproduces an affine field; decoder row k reads
    channel k, so sdf = sum_k w1_k relu(A_k x + b_k + c_k) is known in closed form."""
    rng = np.random.default_rng(7)
    A = rng.normal(size=(32, 3))
    b = rng.normal(size=32)
    net = _affine_model(4, A, b)
    with torch.no_grad():
        l = net.louts[0]
        W0 = np.zeros((128, 32))
        W0[np.arange(32), np.arange(32)] = 1.0
        l[0].weight.copy_(torch.from_numpy(W0))
        c0 = rng.uniform(-0.5, 0.5, size=128)
        l[0].bias.copy_(torch.from_numpy(c0))
        w1 = rng.normal(size=128)
        l[2].weight.copy_(torch.from_numpy(w1).reshape(1, -1))
        l[2].bias.fill_(0.25)
    m = arrays(net)
    x = rand_points(500, 14, -1.0, 1.0).astype(np.float64)
    got = oracle.sdf(m, x, 0, "f64")
    g = m.grids[0].reshape(5, 5, 5, 32).astype(np.float64)
    # affine coefficients recovered from the stored (float32) corner values
    A_s = np.stack([(g[0, 0, 4] - g[0, 0, 0]) / 2, (g[0, 4, 0] - g[0, 0, 0]) / 2, (g[4, 0, 0] - g[0, 0, 0]) / 2], 1)
    feats = x @ A_s.T + (g[0, 0, 0] + A_s.sum(1))
    W0d = m.w0[0].astype(np.float64)
    hid = np.maximum(feats @ W0d.T + m.b0[0].astype(np.float64), 0.0)
    want = hid @ m.w1[0].astype(np.float64) + float(m.b1[0][0])
    # the float32 corner values are not exactly affine; the residual bounds the error
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-5)


def test_grid_vertices_return_stored_values():
    net = small_model(num_lods=1, seed=8)
    m = arrays(net)
    r = m.res[0]
    k = np.random.default_rng(9).integers(0, r + 1, size=(200, 3))
    x = (-1.0 + 2.0 * k / r).astype(np.float64)
    # single-channel readout decoder
    with torch.no_grad():
        l = net.louts[0]
        l[0].weight.zero_()
        l[0].weight[0, 5] = 1.0
        l[0].bias.zero_()
        l[0].bias[0] = 3.0
        l[2].weight.zero_()
        l[2].weight[0, 0] = 1.0
        l[2].bias.fill_(-3.0)
    m = arrays(net)
    got = oracle.sdf(m, x, 0, "f64")
    g = m.grids[0].reshape(r + 1, r + 1, r + 1, 32)
    want = g[k[:, 2], k[:, 1], k[:, 0], 5].astype(np.float64)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_border_clamp():
    net = small_model(num_lods=3, seed=10)
    m = arrays(net)
    x = rand_points(500, 15, -3.0, 3.0).astype(np.float64)
    got = oracle.sdf(m, x, 2, "f64")
    # decoder input xyz is unclamped only when pos_invariant is False; here it is True
    want = oracle.sdf(m, np.clip(x, -1, 1), 2, "f64")
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_sdf_all_consistent_with_single_lod():
    net = small_model(num_lods=3, pos_invariant=False, seed=16)
    m = arrays(net)
    x = rand_points(300, 17)
    allp = oracle.sdf_all(m, x, 2, "f64")
    for lod in range(3):
        np.testing.assert_array_equal(allp[:, lod], oracle.sdf(m, x, lod, "f64"))


def _grid_sample64(g, x):
    """F.grid_sample (align_corners=True, border) of a [C, S, S, S] FP64 grid at x [n, 3]."""
    t = torch.from_numpy(np.ascontiguousarray(g)).double()[None]
    xs = torch.from_numpy(np.ascontiguousarray(x)).double().reshape(1, -1, 1, 1, 3)
    return F.grid_sample(t, xs, align_corners=True, padding_mode="border")[0, :, :, 0, 0].T.numpy()


def _cum_points(res, seed):
    """Random points inside and outside the box, every fine-cell boundary
    plane of the finest grid, and the box corners."""
    g = np.random.default_rng(seed)
    x = [g.uniform(-1.3, 1.3, size=(3000, 3))]
    R = res[-1]
    planes = -1.0 + 2.0 * np.arange(R + 1) / R
    b = g.uniform(-1.0, 1.0, size=(R + 1, 3))
    b[:, g.integers(0, 3)] = planes
    x.append(b)
    x.append(np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=np.float64))
    return np.concatenate(x)


def test_cumulative_grid_sample_equals_the_level_sum():
    """One trilinear sample of the cumulative grid of LOD L equals the sum of
    the per-level samples of levels 0..L (F.grid_sample in FP64, the
    reference's operator), inside the box, on cell boundaries and clamped at
    the border; and a cumulative grid missing a level, or built with x and z
    swapped, does not."""
    net = small_model(num_lods=4, seed=21)  # res 4, 8, 16, 32
    m = arrays(net)
    cum = oracle.cumulative_grids(m, "f64")
    x = _cum_points(m.res, 3)
    C = m.fdim
    grids = [g.reshape(r + 1, r + 1, r + 1, C).transpose(3, 0, 1, 2) for g, r in zip(m.grids, m.res)]
    level_sum = 0.0
    for L, R in enumerate(m.res):
        level_sum = level_sum + _grid_sample64(grids[L], x)
        cg = cum[L].reshape(R + 1, R + 1, R + 1, C).transpose(3, 0, 1, 2)
        got = _grid_sample64(cg, x)
        scale = np.abs(level_sum).max()
        assert np.abs(got - level_sum).max() <= 1e-13 * scale, L
        if L >= 1:
            swapped = _grid_sample64(cg.transpose(0, 3, 2, 1), x)
            assert np.abs(swapped - level_sum).max() > 1e-3 * scale
            dropped = got - _grid_sample64(grids[0], x)  # as if level 0 were left out
            assert np.abs(dropped - level_sum).max() > 1e-3 * scale


def test_cumulative_grid_rounding_and_nesting():
    """The FP32 values are the FP64 sums rounded once; the FP16 values are
    those FP32 values rounded; on the coarse corners the cumulative grid is
    the plain sum of the coincident grid values; grids that do not nest are
    refused."""
    net = small_model(num_lods=3, seed=22)  # res 4, 8, 16
    m = arrays(net)
    c64 = oracle.cumulative_grids(m, "f64")
    c32 = oracle.cumulative_grids(m, "f32")
    C = m.fdim
    for L in range(3):
        f32, f16 = c32[L]
        assert np.array_equal(f32, c64[L].astype(np.float32))
        assert np.array_equal(f16, f32.astype(np.float16))
    # corner (4, 4, 4) of the 16-grid is corner (1, 1, 1) of the 4-grid and (2, 2, 2) of the 8-grid
    want = sum(g.reshape(r + 1, r + 1, r + 1, C)[k, k, k].astype(np.float64)
               for g, r, k in zip(m.grids, m.res, (1, 2, 4)))
    assert np.array_equal(c64[2].reshape(17, 17, 17, C)[4, 4, 4], want)
    bad = oracle.ModelArrays([np.zeros(4 ** 3 * C), np.zeros(5 ** 3 * C)], [3, 4], m.w0[:1] * 2, m.b0[:1] * 2,
                             m.w1[:1] * 2, m.b1[:1] * 2, True)
    with pytest.raises(ValueError):
        oracle.cumulative_grids(bad)
