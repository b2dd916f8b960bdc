"""Training step of the dense model (Trainer.step_geometry,
sdf-net/lib/trainer.py:294-340) on the HIP path, and the seeded fit that
produces the benchmark model.

step_geometry: zero_grad; predictions of every LOD in loss_lods from one
fused forward (the reference's return_lst branch, :317-319); loss =
sum over LODs of sum (pred - gt)^2, divided by the batch size (:325-337);
backward through nglod_sdf_bwd; Adam (lr 1e-3, :184-185).
"""
import torch

from . import workloads
from .meshdata import MeshData
from . import mesh as _mesh


class Trainer:
    def __init__(self, net, lr=1e-3, loss_lods=None):
        self.net = net
        self.optimizer = torch.optim.Adam(net.parameters(), lr=lr)
        self.loss_lods = list(range(net.num_lods)) if loss_lods is None else list(loss_lods)
        self.last_loss = None

    def step_geometry(self, pts, gts):
        net = self.net
        net.train()
        batch_size = pts.shape[0]
        net.zero_grad()
        preds = net.sdf(pts, return_lst=True)
        preds = [preds[i] for i in self.loss_lods]
        loss = 0.0
        for pred in preds:
            loss = loss + ((pred - gts) ** 2).sum()
        loss = loss / batch_size
        loss.backward()
        self.optimizer.step()
        self.last_loss = loss.detach()
        return self.last_loss


def fit(net, md, iterations=200, batch=2 ** 17, seed=0, lr=1e-3):
    """Seeded fit of `net` to the mesh held by `md` (MeshData)."""
    tr = Trainer(net, lr=lr)
    for it in range(iterations):
        pts, gts = md.training_batch(batch, seed=seed * 100003 + it)
        tr.step_geometry(pts, gts)
    return tr


def fitted_icosphere_model(seed=0, device="cuda", iterations=200, batch=2 ** 17, lr=1e-3, pos_invariant=True):
    """The benchmark model: 5-LOD OctreeSDF (seeded init) fitted for
    `iterations` Adam steps to the seeded displaced level-4 icosphere."""
    net = workloads.make_model(num_lods=5, pos_invariant=pos_invariant, seed=seed, device=device)
    V, F = _mesh.displaced_icosphere(seed)
    md = MeshData(V, F, device)
    fit(net, md, iterations=iterations, batch=batch, seed=seed, lr=lr)
    net.eval()
    return net, md
