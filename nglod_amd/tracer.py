"""SphereTracer / RenderBuffer on the HIP path — drop-in for
sdf-net/lib/tracer/{BaseTracer,SphereTracer,RenderBuffer}.py.

`SphereTracer(args)(net, ray_o, ray_d)` returns the same RenderBuffer
(x, depth [N,1], hit [N] bool, normal [N,3]) as the reference, computed by one
persistent HIP kernel (nglod_trace_dense) instead of a Python loop with a
host sync and a mask compaction per step.
"""
import copy
import ctypes
from dataclasses import dataclass, fields
from typing import List, Union

import torch

from . import _lib


def setparam(args, param, paramstr):
    """argparse bridge of the reference (sdf-net/lib/utils.py:33-38)."""
    argsparam = getattr(args, paramstr, None)
    if param is not None or argsparam is None:
        return param
    return argsparam


@dataclass
class RenderBuffer:
    """Per-ray render outputs; the field list and its order are the
    reference's interface (sdf-net/lib/tracer/RenderBuffer.py:28-41), so
    buffers unpack and concatenate like the reference's.

    Contract of the methods (RenderBuffer.py:43-139):
      * iterating yields the fields in declaration order;
      * `a + b` joins two buffers field by field along dim 0; a field set in
        only one of them is taken as is, a field set in neither stays None;
      * cuda/cpu/detach/byte/float/numpy/reshape/transpose map one tensor
        operation over every set field (transpose swaps the first two dims
        of a [W, H, k] field);
      * exrdict() is a copy of the set fields by name, with `rgb` renamed
        `default`;
      * image() is a preview buffer in the 0..255 range: hit and
        relative_depth repeated to three channels, normal moved from [-1, 1]
        to [0, 1], rgb as is, each times 255 (relative_depth goes to `depth`);
      * mean(*buffers) sums each field over the buffers that set it and
        divides by the number of buffers given.
    """

    x: Union[torch.Tensor, None] = None
    min_x: Union[torch.Tensor, None] = None
    hit: Union[torch.Tensor, None] = None
    depth: Union[torch.Tensor, None] = None
    relative_depth: Union[torch.Tensor, None] = None
    normal: Union[torch.Tensor, None] = None
    rgb: Union[torch.Tensor, None] = None
    shadow: Union[torch.Tensor, None] = None
    ao: Union[torch.Tensor, None] = None
    view: Union[torch.Tensor, None] = None
    err: Union[torch.Tensor, None] = None
    albedo: Union[torch.Tensor, None] = None

    def __iter__(self):
        return (getattr(self, f.name) for f in fields(self))

    def __add__(self, other):
        joined = {}
        for f in fields(self):
            mine, theirs = getattr(self, f.name), getattr(other, f.name)
            if mine is None or theirs is None:
                joined[f.name] = theirs if mine is None else mine
            else:
                joined[f.name] = torch.cat((mine, theirs))
        return type(self)(**joined)

    def _apply(self, op):
        return type(self)(**{f.name: op(v) for f, v in zip(fields(self), self) if v is not None})

    def cuda(self):
        return self._apply(torch.Tensor.cuda)

    def cpu(self):
        return self._apply(torch.Tensor.cpu)

    def detach(self):
        return self._apply(torch.Tensor.detach)

    def byte(self):
        return self._apply(torch.Tensor.byte)

    def float(self):
        return self._apply(torch.Tensor.float)

    def numpy(self):
        return self._apply(torch.Tensor.numpy)

    def reshape(self, *dims: List[int]):
        return self._apply(lambda v: v.reshape(*dims))

    def transpose(self):
        return self._apply(lambda v: v.permute(1, 0, 2))

    def exrdict(self):
        out = {f.name: copy.deepcopy(v) for f, v in zip(fields(self), self) if v is not None}
        if "rgb" in out:
            out["default"] = out.pop("rgb")
        return out

    def image(self):
        def scaled(v):
            return None if v is None else v * 255.0

        def grey3(v):
            return None if v is None else torch.cat((v, v, v), dim=-1)

        def unit(v):
            return None if v is None else (v + 1.0) / 2.0

        return type(self)(hit=scaled(grey3(self.hit)), depth=scaled(grey3(self.relative_depth)),
                          normal=scaled(unit(self.normal)), rgb=scaled(self.rgb))

    @staticmethod
    def mean(*rblst):
        count = float(len(rblst))
        out = {}
        for f in fields(RenderBuffer):
            total = None
            for rb in rblst:
                v = getattr(rb, f.name)
                if v is not None:
                    total = v if total is None else total + v
            out[f.name] = None if total is None else total / count
        return RenderBuffer(**out)


class BaseTracer(object):
    """Tracer parameters (sdf-net/lib/tracer/BaseTracer.py:29-45); defaults
    from sdf-net/lib/options.py:209-221 when neither args nor kwargs set them."""

    _DEFAULTS = dict(camera_clamp=[-5, 10], step_size=1.0, grad_method="finitediff", num_steps=256, min_dis=0.0003)

    def __init__(self, args=None, camera_clamp: list = None, step_size: float = None, grad_method: str = None,
                 num_steps: int = None, min_dis: float = None):
        self.args = args
        self.camera_clamp = setparam(args, camera_clamp, "camera_clamp")
        self.step_size = setparam(args, step_size, "step_size")
        self.grad_method = setparam(args, grad_method, "grad_method")
        self.num_steps = setparam(args, num_steps, "num_steps")
        self.min_dis = setparam(args, min_dis, "min_dis")
        for k, v in self._DEFAULTS.items():
            if getattr(self, k) is None:
                setattr(self, k, v)
        self.inv_num_steps = 1.0 / self.num_steps
        self.diagonal = 3 ** 0.5 * 2.0

    def __call__(self, *args, **kwargs):
        return self.forward(*args, **kwargs)

    def forward(self, net, ray_o, ray_d):
        raise NotImplementedError


class SphereTracer(BaseTracer):
    """Sphere tracing with the semantics of SphereTracer.forward
    (sdf-net/lib/tracer/SphereTracer.py:41-132).

    form: 'fp32' (the reference's arithmetic) or 'mixed' (FP16 grids, BF16
    MFMA decoder; when the march stops on a surface test its last near-surface
    evaluations -- at most `refine_steps`, back to the last step of at least
    `near_dist`; near_dist=0 takes all `refine_steps` -- are redone in FP32
    and it goes on in FP32 to the FP32 stop; normals are FP32).

    The tracer keeps one device workspace per device; concurrent calls of one
    tracer object on different streams would share it and must not overlap."""

    def __init__(self, args=None, form: str = None, refine_steps: int = 4, return_dist: float = 0.0,
                 near_dist: float = 1e-2, **kw):
        super().__init__(args, **kw)
        self.form = form or getattr(args, "form", None) or "fp32"
        self.refine_steps = refine_steps
        self.return_dist = return_dist
        self.near_dist = near_dist
        self._ws = {}
        self.last_evals = None
        # When True, each call times its passes with HIP events (and waits
        # for them): last_pass_ms = [march, FP32 continuation (with the
        # mixed form's normals), 0 (no separate normals pass), resets].
        self.time_passes = False
        self.last_pass_ms = None

    def _workspace(self, device, n, form):
        key = (device.type, device.index)
        nbytes = _lib.load().nglod_trace_dense_workspace_size(n, form)
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
            self._ws[key] = ws
        return ws

    def forward(self, net, ray_o, ray_d, ray_steps=None):
        """ray_steps: optional int32 [N] device tensor that receives each ray's
        in-loop SDF evaluation count."""
        if not hasattr(net, "descriptor"):
            raise TypeError("SphereTracer runs nglod_amd models (OctreeSDF); got " + type(net).__name__)
        if self.grad_method not in ("finitediff", "tetrahedron", "autodiff"):
            raise NotImplementedError(f"grad_method {self.grad_method!r}")
        _lib.require_gpu(ray_o, ray_d)
        lod = net.lod if net.lod is not None else net.num_lods - 1
        o = ray_o.detach().float().contiguous()
        d = ray_d.detach().float().contiguous()
        n = o.shape[0]
        dev = o.device
        x = torch.empty(n, 3, device=dev, dtype=torch.float32)
        t = torch.empty(n, 1, device=dev, dtype=torch.float32)
        hit = torch.empty(n, device=dev, dtype=torch.bool)
        normal = torch.empty(n, 3, device=dev, dtype=torch.float32)
        evals = torch.zeros(2, device=dev, dtype=torch.int64)
        form = _lib.MIXED if self.form == "mixed" else _lib.FP32
        p = _lib.NglodTraceParams()
        p.num_steps = int(self.num_steps)
        p.min_dis = float(self.min_dis)
        p.step_size = float(self.step_size)
        p.far_clip = float(self.camera_clamp[1])
        # finite differences run inside the trace; the other methods after it,
        # on the hits (SphereTracer.py:128-130)
        fused_normals = self.grad_method == "finitediff"
        p.normals = 1 if fused_normals else 0
        p.refine_steps = int(self.refine_steps)
        p.return_dist = float(self.return_dist)
        p.near_dist = float(self.near_dist)
        pass_ms = (ctypes.c_float * 4)() if self.time_passes else None
        p.pass_ms = ctypes.cast(pass_ms, ctypes.c_void_p) if pass_ms is not None else None
        m = net.descriptor(form)
        ws = self._workspace(dev, n, form)
        lib = _lib.load()
        _lib.check(lib.nglod_trace_dense(m, _lib.ptr(o), _lib.ptr(d), n, lod, p, form, _lib.ptr(x), _lib.ptr(t),
                                         _lib.ptr(hit), _lib.ptr(normal), _lib.ptr(ws), ws.numel(), _lib.ptr(evals),
                                         _lib.ptr(ray_steps), _lib.stream_of(o)), "nglod_trace_dense")
        if not fused_normals:
            normal = net.gradient_hip(x, lod, self.grad_method, mask=hit, normalize=True, form=form)
        self.last_evals = evals
        self.last_pass_ms = list(pass_ms) if pass_ms is not None else None
        return RenderBuffer(x=x, depth=t, hit=hit, normal=normal)

    def get_min(self, net, ray_o, ray_d):
        """SphereTracer.get_min (SphereTracer.py:134-218), FP32 whatever the
        form: the march of every ray from its box entry (from its origin when
        it has none), keeping in min_x the point the reference's loop pairs
        with the smallest distance seen; normals are the raw (unnormalised)
        gradient by grad_method at the hits, zero elsewhere.  The batch loop's
        stopping rule is kept, so a ray's min_x can depend on the batch, as in
        the reference.  The reference builds its RenderBuffer with a `minx`
        keyword the dataclass does not have (its field is min_x), which raises;
        this returns min_x in that field."""
        if not hasattr(net, "descriptor"):
            raise TypeError("SphereTracer runs nglod_amd models (OctreeSDF); got " + type(net).__name__)
        if self.grad_method not in ("finitediff", "tetrahedron", "autodiff"):
            raise NotImplementedError(f"grad_method {self.grad_method!r}")
        _lib.require_gpu(ray_o, ray_d)
        lod = net.lod if net.lod is not None else net.num_lods - 1
        o = ray_o.detach().float().contiguous()
        d = ray_d.detach().float().contiguous()
        n = o.shape[0]
        dev = o.device
        x = torch.empty(n, 3, device=dev, dtype=torch.float32)
        t = torch.empty(n, 1, device=dev, dtype=torch.float32)
        hit = torch.empty(n, device=dev, dtype=torch.bool)
        min_x = torch.empty(n, 3, device=dev, dtype=torch.float32)
        p = _lib.NglodTraceParams()
        p.num_steps = int(self.num_steps)
        p.min_dis = float(self.min_dis)
        p.step_size = float(self.step_size)
        p.far_clip = float(self.camera_clamp[1])
        lib = _lib.load()
        ws = torch.empty(max(int(lib.nglod_trace_min_workspace_size(n)), 256), dtype=torch.uint8, device=dev)
        _lib.check(lib.nglod_trace_min(net.descriptor(_lib.FP32), _lib.ptr(o), _lib.ptr(d), n, lod, p, _lib.ptr(x),
                                       _lib.ptr(t), _lib.ptr(hit), _lib.ptr(min_x), _lib.ptr(ws), ws.numel(),
                                       _lib.stream_of(o)), "nglod_trace_min")
        normal = net.gradient_hip(x, lod, self.grad_method, mask=hit, normalize=False)
        return RenderBuffer(x=x, depth=t, hit=hit, normal=normal, min_x=min_x)

    def sample_surface(self, n, net, seed=0, max_batches=1000):
        """SphereTracer.sample_surface (SphereTracer.py:220-245): batches of n
        rays with origins uniform in [-1, 1]^3 and directions uniform on the
        sphere, traced until at least n hits are collected; returns every
        collected hit point ([M, 3], M >= n).  The device is the model's (the
        reference reads a self.device it never sets).  As in the reference,
        the tracer reports an origin inside the box as a hit at depth 0 (its
        box entry skips such rays, SphereTracer.py:53 with
        sol_nglod_kernel.cu:109-112), so with these origins the points are
        mostly the origins themselves; sol.voxel_sparsify starts its rays
        outside the box for a sample of the surface.  Rays come from seeded
        counter streams (nglod_sample_uniform / nglod_sample_directions)."""
        dev = net.features[0].fm.device
        lib = _lib.load()
        pts, total = [], 0
        for b in range(max_batches):
            o = torch.empty(n, 3, device=dev, dtype=torch.float32)
            d = torch.empty(n, 3, device=dev, dtype=torch.float32)
            s = _lib.stream_of(o)
            _lib.check(lib.nglod_sample_uniform(n, 2 * (seed * max_batches + b), _lib.ptr(o), s),
                       "nglod_sample_uniform")
            _lib.check(lib.nglod_sample_directions(n, 2 * (seed * max_batches + b) + 1, _lib.ptr(d), s),
                       "nglod_sample_directions")
            rb = self.forward(net, o, d)
            pts.append(rb.x[rb.hit])
            total += pts[-1].shape[0]
            if total >= n:
                break
            if b == 49:
                print("Taking an unusually long time to sample desired # of points.")
        return torch.cat(pts, dim=0)
