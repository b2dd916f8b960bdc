"""torch.autograd.Function around the fused HIP query (OctreeSDF.sdf) with the
hand-written backward: parameter gradients by nglod_sdf_bwd, computed
deterministically (no float atomics), and position gradients (when the query
points require grad) by nglod_sdf_grad, the derivative torch autograd takes
through grid_sample, the decoder and ReLU.  First order only: the backward
itself is not differentiable (gradient penalties through it raise)."""
import torch
from torch.autograd.function import once_differentiable

from . import _lib


def _param_slots(net):
    """(kind, index, name) of every parameter in net.parameters() order."""
    slots = []
    for name, _ in net.named_parameters():
        parts = name.split(".")
        if parts[0] == "features":
            slots.append(("grid", int(parts[1]), None))
        elif parts[0] == "louts":
            key = {("0", "weight"): "w0", ("0", "bias"): "b0", ("2", "weight"): "w1", ("2", "bias"): "b1"}[
                (parts[2], parts[3])]
            slots.append((key, int(parts[1]), None))
        else:
            slots.append((None, -1, name))
    return slots


class _SdfFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, net, x, lod, all_levels, *params):
        ctx.net = net
        ctx.lod = lod
        ctx.all_levels = all_levels
        xf = x.detach().reshape(-1, 3).float().contiguous()
        ctx.x_shape = x.shape
        ctx.x_dtype = x.dtype
        ctx.save_for_backward(xf)
        if all_levels:
            return torch.stack(net.sdf_all_hip(xf, lod), 0)
        return net.sdf_hip(xf, lod)

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_out):
        (x,) = ctx.saved_tensors
        net, L = ctx.net, ctx.lod
        n = x.shape[0]
        dev = x.device
        if ctx.all_levels:
            g = grad_out[..., 0].float().contiguous()  # [L+1, n]
            mask = (1 << (L + 1)) - 1
        else:
            g = torch.zeros(L + 1, n, dtype=torch.float32, device=dev)
            g[L] = grad_out[:, 0].float()
            mask = 1 << L
        grad_x = None
        if ctx.needs_input_grad[1]:
            # sum over the decoded levels of d loss / d d_l * d d_l / d x
            grad_x = torch.zeros(n, 3, dtype=torch.float32, device=dev)
            for lvl in range(L + 1):
                if (mask >> lvl) & 1:
                    grad_x.addcmul_(g[lvl][:, None], net.sdf_grad_hip(x, lvl)[1])
            grad_x = grad_x.to(ctx.x_dtype).reshape(ctx.x_shape)
        if not any(ctx.needs_input_grad[4:]):
            return (None, grad_x, None, None, *([None] * len(ctx.needs_input_grad[4:])))
        used_dec = {0} if net.num_decoder == 1 else {i for i in range(L + 1) if (mask >> i) & 1}
        G = _lib.NglodModelGrads()
        out = {}
        for i, f in enumerate(net.features):
            if i <= L:
                t = torch.empty_like(f.fm, memory_format=torch.channels_last_3d)
                out[("grid", i)] = t
                G.grid[i] = t.data_ptr()
        for d, l in enumerate(net.louts):
            for key, p in (("w0", l[0].weight), ("b0", l[0].bias), ("w1", l[2].weight), ("b1", l[2].bias)):
                t = torch.empty_like(p, memory_format=torch.contiguous_format)
                out[(key, d)] = t
                getattr(G, key)[d] = t.data_ptr()
        m = net.descriptor(_lib.FP32)
        lib = _lib.load()
        nbytes = lib.nglod_sdf_bwd_workspace_size(m, n, L)
        ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
        _lib.check(lib.nglod_sdf_bwd(m, _lib.ptr(x), n, L, _lib.ptr(g), mask, G, _lib.ptr(ws), ws.numel(),
                                     _lib.stream_of(x)), "nglod_sdf_bwd")
        grads = []
        for kind, idx, _ in _param_slots(net):
            if kind == "grid":
                grads.append(out.get(("grid", idx)))
            elif kind is not None and idx in used_dec:
                grads.append(out[(kind, idx)])
            else:
                grads.append(None)
        return (None, grad_x, None, None, *grads)


def sdf_with_grad(net, x, lod, all_levels):
    params = list(net.parameters())
    out = _SdfFunction.apply(net, x, lod, all_levels, *params)
    if all_levels:
        return [out[i] for i in range(out.shape[0])]
    return out
