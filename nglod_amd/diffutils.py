"""diffutils.gradient (sdf-net/lib/diffutils.py:29-84) on the HIP path.

gradient(x, f, method) with f an nglod_amd OctreeSDF runs one kernel
(nglod_gradient): 'autodiff' is the derivative torch autograd takes through
grid_sample (align_corners=True, padding_mode='border'), the decoder and
ReLU; 'tetrahedron' and 'finitediff' are the reference's taps (h = 1/192)
combined in the reference's order.  The model's LOD is f.lod (the last LOD
when it is None), as net(x) uses it.

Differences a caller can see:
- 'autodiff' returns a tensor without an autograd graph into the model's
  parameters (the reference builds one with create_graph=True, for gradient
  losses).  When gradients are enabled and a parameter requires grad, the
  result is attached to a node whose backward raises, so a loss built on it
  fails loudly instead of training without that term.
- 'finitediff' and 'tetrahedron' with gradients enabled and parameters that
  require grad evaluate f itself at the taps, as the reference does, so the
  result stays differentiable with respect to the parameters; otherwise one
  fused kernel.
- 'multilayer' returns the autodiff gradients of every level's decoder
  0..lod (the reference's branch is defined for lod None, where it returns
  all of them).
Any other callable f is differentiated the reference's way with torch ops.
"""
import torch

from . import _lib

H = 1.0 / (64.0 * 3.0)


class _NoSecondOrder(torch.autograd.Function):
    @staticmethod
    def forward(ctx, g, *params):
        return g.clone()

    @staticmethod
    def backward(ctx, grad):
        raise NotImplementedError("the autodiff gradient of an nglod_amd model is computed without a graph; "
                                  "losses on it (second-order derivatives) are not supported")


def _is_nglod(f):
    return hasattr(f, "gradient_hip") and hasattr(f, "descriptor")


def _lod_of(f):
    return f.lod if f.lod is not None else f.num_lods - 1


def _params_need_grad(f):
    return torch.is_grad_enabled() and any(p.requires_grad for p in f.parameters())


def _taps(x, f, method):
    """The reference's tap arithmetic with f called at each tap (diffutils.py:38-70)."""
    dev = x.device
    if method == "tetrahedron":
        k = [torch.tensor(v, device=dev) for v in ([1.0, -1.0, -1.0], [-1.0, -1.0, 1.0], [-1.0, 1.0, -1.0],
                                                    [1.0, 1.0, 1.0])]
        hs = [torch.tensor([H * c for c in v], device=dev) for v in ([1.0, -1.0, -1.0], [-1.0, -1.0, 1.0],
                                                                     [-1.0, 1.0, -1.0], [1.0, 1.0, 1.0])]
        terms = [k[j] * f((x + hs[j]).detach()) for j in range(4)]
        return (terms[0] + terms[1] + terms[2] + terms[3]) / (H * 4.0)
    ex = torch.tensor([H, 0.0, 0.0], device=dev)
    ey = torch.tensor([0.0, H, 0.0], device=dev)
    ez = torch.tensor([0.0, 0.0, H], device=dev)
    g = torch.cat([f(x + ex) - f(x - ex), f(x + ey) - f(x - ey), f(x + ez) - f(x - ez)], dim=-1)
    return g / (H * 2.0)


def gradient(x, f, method="autodiff"):
    """Gradient of f at x [N, 3] -> [N, 3] (a list of them for 'multilayer')."""
    if method not in ("autodiff", "tetrahedron", "finitediff", "multilayer"):
        raise NotImplementedError(method)
    if _is_nglod(f):
        _lib.require_gpu(x)
        lod = _lod_of(f)
        if method == "multilayer":
            return [f.gradient_hip(x, l, "autodiff") for l in range(lod + 1)]
        if method == "autodiff":
            g = f.gradient_hip(x, lod, "autodiff")
            if _params_need_grad(f):
                return _NoSecondOrder.apply(g, *[p for p in f.parameters() if p.requires_grad])
            return g
        if _params_need_grad(f):
            return _taps(x, f, method)
        return f.gradient_hip(x, lod, method)
    if method == "autodiff":
        with torch.enable_grad():
            x = x.requires_grad_(True)
            y = f(x)
            return torch.autograd.grad(y, x, grad_outputs=torch.ones_like(y), create_graph=True)[0]
    if method == "multilayer":
        grads = []
        with torch.enable_grad():
            ys = f.sdf(x, return_lst=True)
            for y in ys:
                grads.append(torch.autograd.grad(y, x, grad_outputs=torch.ones_like(y), create_graph=True)[0])
        return grads
    return _taps(x, f, method)
