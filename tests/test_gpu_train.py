"""Hand-written backward (nglod_sdf_bwd through the autograd Function) against
the reference's CPU autograd (F.grid_sample + nn.Linear in FP64) and against
central finite differences (h = 1e-3, 64 random parameters) of the FP64
oracle loss; determinism; a short fit lowers the loss."""
import numpy as np
import pytest
import torch

from nglod_amd import mesh, trainer, workloads
from nglod_amd.meshdata import MeshData
from oracle import oracle
from oracle.torch_ref import TorchOctreeSDF
from tests.helpers import arrays

pytestmark = pytest.mark.gpu


def _batch(dev, n=4096, seed=0):
    V, F = mesh.displaced_icosphere(1, level=3)
    md = MeshData(V, F, dev)
    return md.training_batch(n, seed=seed)


def _loss_grads_gpu(net, pts, gts):
    net.zero_grad()
    preds = net.sdf(pts, return_lst=True)
    loss = sum(((p - gts) ** 2).sum() for p in preds) / pts.shape[0]
    loss.backward()
    return loss.item(), {k: v.grad.detach().clone() for k, v in net.named_parameters() if v.grad is not None}


def _loss_grads_cpu64(net, pts, gts, dtype=torch.float64):
    ref = TorchOctreeSDF.from_module(net, dtype=dtype)
    params = ref.fms + ref.w0 + ref.b0 + ref.w1 + ref.b1
    for p in params:
        p.requires_grad_(True)
    x = pts.detach().cpu().to(dtype)
    gt = gts.detach().cpu().to(dtype)
    preds = ref.sdf(x, return_lst=True)
    loss = sum(((p - gt) ** 2).sum() for p in preds) / x.shape[0]
    loss.backward()
    g = {}
    for i in range(len(ref.fms)):
        g[f"features.{i}.fm"] = ref.fms[i].grad
    for i in range(len(ref.w0)):
        g[f"louts.{i}.0.weight"] = ref.w0[i].grad
        g[f"louts.{i}.0.bias"] = ref.b0[i].grad
        g[f"louts.{i}.2.weight"] = ref.w1[i].grad.reshape(1, -1)
        g[f"louts.{i}.2.bias"] = ref.b1[i].grad.reshape(-1)
    return loss.item(), g


@pytest.mark.parametrize("pos_invariant", [True, False])
def test_backward_vs_cpu_autograd_fp64(dev, pos_invariant):
    net = workloads.make_model(num_lods=5, pos_invariant=pos_invariant, seed=3, device=dev)
    pts, gts = _batch(dev)
    lg, gg = _loss_grads_gpu(net, pts, gts)
    lc, gc = _loss_grads_cpu64(net, pts, gts)
    assert abs(lg - lc) < 1e-5 * max(1.0, abs(lc))
    assert set(gg) == set(gc)
    for k in gc:
        a = gg[k].double().cpu().reshape(gc[k].shape)
        b = gc[k]
        scale = b.abs().max().item()
        assert (a - b).abs().max().item() <= 1e-4 * max(scale, 1e-3), k


def test_backward_vs_central_differences(dev):
    net = workloads.make_model(num_lods=3, pos_invariant=False, seed=4, device=dev)
    with torch.no_grad():
        for f in net.features:
            f.fm.mul_(20.0)  # larger features: decoder in its nonlinear range
    pts, gts = _batch(dev, n=2048, seed=5)
    _, gg = _loss_grads_gpu(net, pts, gts)
    m = arrays(net)
    x = pts.cpu().numpy().astype(np.float64)
    gt = gts[:, 0].cpu().numpy().astype(np.float64)

    def loss():
        p = oracle.sdf_all(m, x, 2, "f64")
        return ((p - gt[:, None]) ** 2).sum() / x.shape[0]

    rng = np.random.default_rng(6)
    h = 1e-3
    checked = 0
    for _ in range(64):
        kind = rng.integers(0, 3)
        lvl = int(rng.integers(0, 3))
        if kind == 0:  # a grid entry at a corner some query touches
            r = m.res[lvl]
            q = x[rng.integers(0, x.shape[0])]
            i0 = np.clip(np.floor((q + 1) / 2 * r), 0, r - 1).astype(int)
            c = int(rng.integers(0, 32))
            idx = ((i0[2] * (r + 1) + i0[1]) * (r + 1) + i0[0]) * 32 + c
            arr, name = m.grids[lvl], f"features.{lvl}.fm"
            g = gg[name].permute(0, 2, 3, 4, 1).reshape(-1)[idx].item()
        elif kind == 1:
            arr, name = m.w0[lvl].reshape(-1), f"louts.{lvl}.0.weight"
            idx = int(rng.integers(0, arr.size))
            g = gg[name].reshape(-1)[idx].item()
        else:
            arr, name = m.w1[lvl], f"louts.{lvl}.2.weight"
            idx = int(rng.integers(0, arr.size))
            g = gg[name].reshape(-1)[idx].item()
        old = arr[idx]
        arr[idx] = old + h
        lp = loss()
        arr[idx] = old - h
        lm = loss()
        arr[idx] = old
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g) <= 2e-3 * max(abs(fd), 1e-2), (name, idx, fd, g)
        checked += 1
    assert checked == 64


def test_backward_deterministic(dev):
    net = workloads.make_model(num_lods=5, seed=7, device=dev)
    pts, gts = _batch(dev, n=20000, seed=8)
    _, g1 = _loss_grads_gpu(net, pts, gts)
    _, g2 = _loss_grads_gpu(net, pts, gts)
    for k in g1:
        assert torch.equal(g1[k], g2[k]), k


def test_per_lod_calls_match_return_lst(dev):
    net = workloads.make_model(num_lods=4, seed=9, device=dev)
    pts, gts = _batch(dev, n=3000, seed=10)
    _, g_all = _loss_grads_gpu(net, pts, gts)
    net.zero_grad()
    loss = sum(((net.sdf(pts, lod=i) - gts) ** 2).sum() for i in range(4)) / pts.shape[0]
    loss.backward()
    for k, v in net.named_parameters():
        torch.testing.assert_close(v.grad, g_all[k], rtol=1e-5, atol=1e-7)


def test_joint_decoder_backward(dev):
    net = workloads.make_model(num_lods=3, joint_decoder=True, seed=11, device=dev)
    pts, gts = _batch(dev, n=2048, seed=12)
    _, gg = _loss_grads_gpu(net, pts, gts)
    _, gc = _loss_grads_cpu64(net, pts, gts)
    # the reference restatement shares decoder 0 across levels: sum its level grads
    for key in ("0.weight", "0.bias", "2.weight", "2.bias"):
        a = gg[f"louts.0.{key}"].double().cpu()
        b = gc[f"louts.0.{key}"].reshape(a.shape)
        assert (a - b).abs().max().item() <= 1e-4 * max(b.abs().max().item(), 1e-3)


def test_short_fit_lowers_loss(dev):
    net = workloads.make_model(num_lods=5, seed=13, device=dev)
    V, F = mesh.displaced_icosphere(0)
    md = MeshData(V, F, dev)
    tr = trainer.Trainer(net)
    pts, gts = md.training_batch(2 ** 14, seed=99)
    first = tr.step_geometry(pts, gts).item()
    for it in range(30):
        p, g = md.training_batch(2 ** 14, seed=it)
        tr.step_geometry(p, g)
    last = tr.step_geometry(pts, gts).item()
    assert last < 0.5 * first


def test_training_step_full_batch_vs_cpu_autograd_fp64(dev):
    """The benchmarked training step at its own size: a 2^19-sample batch of
    the seed-0 displaced level-4 icosphere (5120 triangles; 1/3 uniform, 2/3
    near the surface at sigma 0.01 and 0.003) with FP64 GPU mesh2sdf targets,
    L2 loss summed over 5 LODs (sdf-net/lib/trainer.py:317-339).  Targets on
    a strided sample against the FP64 oracle; the five LODs' gradients
    against FP64 CPU autograd, within 1e-4 of each tensor's largest entry or
    within twice the deviation of the reference's own FP32 CPU autograd
    (the same graph in FP32), whichever is larger: at 2^19 samples a grid
    entry sums thousands of FP32 contributions of both signs, so FP32
    rounding alone can exceed 1e-4 of the largest entry; two runs
    bit-identical; one Adam step moves every parameter."""
    V, F = mesh.displaced_icosphere(0)
    assert F.shape[0] == 5120
    md = MeshData(V, F, dev)
    n = 1 << 19
    pts, gts = md.training_batch(n, seed=3)
    assert pts.shape == (n, 3) and gts.shape == (n, 1)
    sub = slice(0, n, 128)  # 4096 targets against the oracle
    want = oracle.mesh2sdf(pts[sub].cpu().numpy().astype(np.float64), mesh.triangles(V, F))
    np.testing.assert_allclose(gts[sub, 0].cpu().numpy(), want, rtol=0, atol=1e-6)
    net = workloads.make_model(num_lods=5, seed=0, device=dev)
    lg, gg = _loss_grads_gpu(net, pts, gts)
    lg2, gg2 = _loss_grads_gpu(net, pts, gts)
    assert lg == lg2
    for k in gg:
        assert torch.equal(gg[k], gg2[k]), k
    lc, gc = _loss_grads_cpu64(net, pts, gts)
    _, g32 = _loss_grads_cpu64(net, pts, gts, dtype=torch.float32)
    assert abs(lg - lc) < 1e-5 * max(1.0, abs(lc))
    assert set(gg) == set(gc)
    report = {}
    for k in gc:
        a = gg[k].double().cpu().reshape(gc[k].shape)
        b = gc[k]
        scale = b.abs().max().item()
        assert scale > 0.0, k
        err = (a - b).abs().max().item()
        err_ref32 = (g32[k].double().reshape(b.shape) - b).abs().max().item()
        report[k] = (err / scale, err_ref32 / scale)
        assert err <= max(1e-4 * scale, 2.0 * err_ref32), (k, err, err_ref32, scale)
    print("relative max error (GPU, reference FP32 CPU):",
          {k: (f"{v[0]:.1e}", f"{v[1]:.1e}") for k, v in report.items()})
    tr = trainer.Trainer(net)
    before = {k: v.detach().clone() for k, v in net.named_parameters()}
    tr.step_geometry(pts, gts)
    for k, v in net.named_parameters():
        assert not torch.equal(v.detach(), before[k]), k
