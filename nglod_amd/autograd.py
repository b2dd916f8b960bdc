"""torch.autograd.Function around the fused HIP query (OctreeSDF.sdf)."""
import torch

from . import _lib


class _SdfFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, net, x, lod, all_levels, *params):
        ctx.net = net
        ctx.lod = lod
        ctx.all_levels = all_levels
        ctx.save_for_backward(x)
        if all_levels:
            return torch.stack(net.sdf_all_hip(x, lod), 0)
        return net.sdf_hip(x, lod)

    @staticmethod
    def backward(ctx, grad_out):
        raise NotImplementedError("backward of the fused query is not available in this build")


def sdf_with_grad(net, x, lod, all_levels):
    params = [p for p in net.parameters()]
    out = _SdfFunction.apply(net, x, lod, all_levels, *params)
    if all_levels:
        return [out[i] for i in range(out.shape[0])]
    return out
