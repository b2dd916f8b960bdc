"""OctreeSDF on the HIP path — drop-in for sdf-net/lib/models/OctreeSDF.py.

Same constructor arguments (an argparse namespace, as the reference's
`OctreeSDF(args)`), same parameter names and shapes (`features.{i}.fm`
[1, C, r+1, r+1, r+1], `louts.{i}.{0,2}.{weight,bias}`), same `sdf(x, lod,
return_lst)` / `forward(x, lod)` behaviour.  The one difference a caller can
see is the memory layout of `fm`: it is allocated channels_last_3d, so the
storage is corner-major ([z][y][x][c]) and the HIP kernels read it in place.
Loading a reference state_dict copies into that layout unchanged.
"""
import types

import torch
import torch.nn as nn

from . import _lib


def default_args(**kw):
    """The reference's defaults for the fields OctreeSDF reads
    (sdf-net/lib/options.py:61-89)."""
    a = types.SimpleNamespace(
        feature_dim=32,
        feature_size=4,
        hidden_dim=128,
        pos_invariant=False,
        num_lods=1,
        base_lod=2,
        joint_decoder=False,
        interpolate=None,
        pos_enc=False,
        ff_dim=-1,
        ff_width=16.0,
    )
    for k, v in kw.items():
        setattr(a, k, v)
    return a


def corner_major(fm):
    """fm [1, C, D, H, W] as a corner-major view (channels_last_3d storage)."""
    if fm.is_contiguous(memory_format=torch.channels_last_3d):
        return fm
    return fm.contiguous(memory_format=torch.channels_last_3d)


class FeatureVolume(nn.Module):
    """Dense feature grid of one LOD (sdf-net/lib/models/OctreeSDF.py:38-57)."""

    def __init__(self, fdim, fsize):
        super().__init__()
        self.fsize = fsize
        self.fdim = fdim
        fm = torch.randn(1, fdim, fsize + 1, fsize + 1, fsize + 1) * 0.01
        self.fm = nn.Parameter(fm.contiguous(memory_format=torch.channels_last_3d))
        self.sparse = None

    def forward(self, x):
        """Trilinear sample at x [N, 3] -> [N, C] (OctreeSDF.py:46-51).
        The 2-D query branch (x [N, K, 3], :52-55) flattens to [N*K, 3]."""
        _lib.require_gpu(x)
        shape = x.shape[:-1]
        xf = x.reshape(-1, 3).float().contiguous()
        fm = corner_major(self.fm.detach())
        out = torch.empty(xf.shape[0], self.fdim, device=x.device, dtype=torch.float32)
        lib = _lib.load()
        _lib.check(lib.nglod_grid_sample(_lib.ptr(fm), self.fsize, self.fdim, _lib.ptr(xf), xf.shape[0],
                                         _lib.ptr(out), _lib.stream_of(x)), "nglod_grid_sample")
        return out.reshape(*shape, self.fdim)


class OctreeSDF(nn.Module):
    """Multi-LOD dense-grid neural SDF (sdf-net/lib/models/OctreeSDF.py:59-155)."""

    def __init__(self, args=None, init=None, **overrides):
        super().__init__()
        if args is None:
            args = default_args()
        for k, v in overrides.items():
            setattr(args, k, v)
        self.args = args
        # BaseSDF / BaseLOD fields (sdf-net/lib/models/BaseSDF.py:41-54, BaseLOD.py:32-35)
        if getattr(args, "pos_enc", False) or getattr(args, "ff_dim", -1) > 0:
            raise NotImplementedError("positional / Fourier encodings are not on the OctreeSDF path")
        self.input_dim = 3
        self.out_dim = 1
        self.num_lods = args.num_lods
        self.lod = None

        self.fdim = args.feature_dim
        self.fsize = args.feature_size
        self.hidden_dim = args.hidden_dim
        self.pos_invariant = args.pos_invariant
        self.features = nn.ModuleList(
            [FeatureVolume(self.fdim, 2 ** (i + args.base_lod)) for i in range(args.num_lods)])
        self.interpolate = getattr(args, "interpolate", None)
        if self.interpolate is not None:
            # The reference's interpolation branch references an undefined name
            # (OctreeSDF.py:138) and cannot run; it is not provided here.
            raise NotImplementedError("interpolate mode is broken in the reference (OctreeSDF.py:138)")
        self.sdf_input_dim = self.fdim + (0 if self.pos_invariant else self.input_dim)
        self.num_decoder = 1 if args.joint_decoder else args.num_lods
        self.louts = nn.ModuleList([
            nn.Sequential(
                nn.Linear(self.sdf_input_dim, self.hidden_dim, bias=True),
                nn.ReLU(),
                nn.Linear(self.hidden_dim, 1, bias=True),
            ) for _ in range(self.num_decoder)
        ])
        self.loss_preds = None
        self._mixed = None  # (key, cumulative grids FP32, FP16, w0_bf16) cache of the mixed form
        self._order_ws = {}  # device -> workspace of the spatially ordered batch query

    # -- reference API -----------------------------------------------------
    def encode(self, x):
        return x

    def forward(self, x, lod=None):
        """BaseLOD.forward (BaseLOD.py:37-41): the lod argument is ignored and
        self.lod is used, as in the reference."""
        return self.sdf(self.encode(x))

    def sdf(self, x, lod=None, return_lst=False):
        """OctreeSDF.sdf (OctreeSDF.py:94-155).

        lod given      -> [N, 1], the decoder of level `lod`.
        lod None       -> self.lod if set, else every level is decoded and the
                          last one returned (or the list with return_lst).
        In training mode with lod None, self.loss_preds holds every level."""
        if lod is None:
            lod = self.lod
        _lib.require_gpu(x)
        if lod is not None:
            if not 0 <= lod < self.num_lods:
                raise IndexError(f"lod {lod} out of range for {self.num_lods} LODs")
            if self._needs_grad(x):
                from .autograd import sdf_with_grad
                return sdf_with_grad(self, x, lod, all_levels=False)
            return self.sdf_hip(x, lod)
        if self._needs_grad(x):
            from .autograd import sdf_with_grad
            preds = sdf_with_grad(self, x, self.num_lods - 1, all_levels=True)
        else:
            preds = self.sdf_all_hip(x, self.num_lods - 1)
        if self.training:
            self.loss_preds = preds
        return preds if return_lst else preds[-1]

    # -- HIP path ----------------------------------------------------------
    def _needs_grad(self, x):
        if not torch.is_grad_enabled():
            return False
        return x.requires_grad or any(p.requires_grad for p in self.parameters())

    def decoder_index(self, lod):
        return 0 if self.num_decoder == 1 else lod

    def descriptor(self, form=_lib.FP32):
        """nglod_model_t for the current parameters (device pointers)."""
        m = _lib.NglodModel()
        m.num_lods = self.num_lods
        m.feature_dim = self.fdim
        m.hidden_dim = self.hidden_dim
        m.pos_invariant = 1 if self.pos_invariant else 0
        m.num_decoders = self.num_decoder
        keep = []
        for i, f in enumerate(self.features):
            g = corner_major(f.fm.detach())
            keep.append(g)
            m.res[i] = f.fsize
            m.grid[i] = g.data_ptr()
        for i, l in enumerate(self.louts):
            for name, t in (("w0", l[0].weight), ("b0", l[0].bias), ("w1", l[2].weight), ("b1", l[2].bias)):
                t = t.detach().contiguous()
                keep.append(t)
                getattr(m, name)[i] = t.data_ptr()
        if form == _lib.MIXED:
            grids_c, grids_h, w0s = self.mixed_buffers()
            for i, (gc, gh) in enumerate(zip(grids_c, grids_h)):
                m.grid_cum[i] = gc.data_ptr()
                m.grid_h[i] = gh.data_ptr()
            for i, w in enumerate(w0s):
                m.w0_bf16[i] = w.data_ptr()
            keep += grids_c + grids_h + w0s
        m._keep = keep
        return m

    def mixed_buffers(self):
        """What the precision form reads (nglod_pack_mixed): per LOD the
        cumulative grid (at each corner, the sum of the trilinear samples of
        that level and every coarser one) in FP32 and its FP16 rounding, and
        the packed BF16 first-layer weights; rebuilt when any parameter changed
        since the last call."""
        key = tuple((p.data_ptr(), p._version) for p in self.parameters())
        if self._mixed is not None and self._mixed[0] == key:
            return self._mixed[1], self._mixed[2], self._mixed[3]
        dev = self.features[0].fm.device
        grids_c = [torch.empty(f.fm.numel(), dtype=torch.float32, device=dev) for f in self.features]
        grids_h = [torch.empty(f.fm.numel(), dtype=torch.float16, device=dev) for f in self.features]
        w0s = [torch.empty(self.hidden_dim * self.fdim, dtype=torch.int16, device=dev) for _ in self.louts]
        m = self.descriptor(_lib.FP32)
        for i, (gc, gh) in enumerate(zip(grids_c, grids_h)):
            m.grid_cum[i] = gc.data_ptr()
            m.grid_h[i] = gh.data_ptr()
        for i, w in enumerate(w0s):
            m.w0_bf16[i] = w.data_ptr()
        lib = _lib.load()
        _lib.check(lib.nglod_pack_mixed(m, _lib.stream_of(grids_h[0])), "nglod_pack_mixed")
        self._mixed = (key, grids_c, grids_h, w0s)
        return grids_c, grids_h, w0s

    # Batches of at least this many points are evaluated in Morton order
    # (nglod_sdf_fwd_ordered); the results are identical either way.
    ORDERED_MIN = 1 << 16

    def sdf_hip(self, x, lod, form=_lib.FP32, ordered=None):
        """ordered: None picks by batch size; True / False forces the spatially
        ordered / input-order evaluation."""
        xf = x.detach().reshape(-1, 3).float().contiguous()
        n = xf.shape[0]
        out = torch.empty(n, 1, device=x.device, dtype=torch.float32)
        m = self.descriptor(form)
        lib = _lib.load()
        if ordered is None:
            ordered = n >= self.ORDERED_MIN
        if ordered and n > 0:
            nbytes = int(lib.nglod_sdf_fwd_ordered_workspace_size(n))
            ws = self._order_ws.get(x.device)
            if ws is None or ws.numel() < nbytes:
                ws = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
                self._order_ws[x.device] = ws
            _lib.check(lib.nglod_sdf_fwd_ordered(m, _lib.ptr(xf), n, lod, form, _lib.ptr(out), _lib.ptr(ws),
                                                 ws.numel(), _lib.stream_of(x)), "nglod_sdf_fwd_ordered")
            return out
        _lib.check(lib.nglod_sdf_fwd(m, _lib.ptr(xf), n, lod, form, _lib.ptr(out), _lib.stream_of(x)),
                   "nglod_sdf_fwd")
        return out

    def gradient_hip(self, x, lod, method="autodiff", mask=None, normalize=False, form=_lib.FP32):
        """d sdf(x, lod) / d x [N, 3] (diffutils.gradient, sdf-net/lib/diffutils.py:29-84)
        by 'finitediff', 'tetrahedron' or 'autodiff' (the derivative torch
        autograd computes), FP32.  mask: optional bool [N]; rows where it is
        False are 0.  normalize: F.normalize(eps=1e-5) of each row, as
        SphereTracer.forward applies to its normals.  No autograd graph."""
        if method not in _lib.GRAD_METHODS:
            raise NotImplementedError(f"gradient method {method!r}")
        xf = x.detach().reshape(-1, 3).float().contiguous()
        n = xf.shape[0]
        out = torch.empty(n, 3, device=x.device, dtype=torch.float32)
        mk = None
        if mask is not None:
            mk = mask.reshape(-1).to(device=x.device, dtype=torch.uint8).contiguous()
            if mk.numel() != n:
                raise ValueError("mask must hold one flag per point")
        m = self.descriptor(form)
        lib = _lib.load()
        _lib.check(lib.nglod_gradient(m, _lib.ptr(xf), n, int(lod), form, _lib.GRAD_METHODS[method], _lib.ptr(mk),
                                      1 if normalize else 0, _lib.ptr(out), _lib.stream_of(x)), "nglod_gradient")
        return out

    def sdf_grad_hip(self, x, lod):
        """(sdf [N, 1], d sdf / d x [N, 3]) at `lod` in one pass, FP32 (the
        autodiff gradient)."""
        xf = x.detach().reshape(-1, 3).float().contiguous()
        n = xf.shape[0]
        y = torch.empty(n, 1, device=x.device, dtype=torch.float32)
        g = torch.empty(n, 3, device=x.device, dtype=torch.float32)
        m = self.descriptor(_lib.FP32)
        lib = _lib.load()
        _lib.check(lib.nglod_sdf_grad(m, _lib.ptr(xf), n, int(lod), _lib.FP32, _lib.ptr(y), _lib.ptr(g),
                                      _lib.stream_of(x)), "nglod_sdf_grad")
        return y, g

    def sdf_all_hip(self, x, lod):
        xf = x.detach().reshape(-1, 3).float().contiguous()
        n = xf.shape[0]
        out = torch.empty(lod + 1, n, 1, device=x.device, dtype=torch.float32)
        m = self.descriptor(_lib.FP32)
        lib = _lib.load()
        _lib.check(lib.nglod_sdf_fwd_all(m, _lib.ptr(xf), n, lod, _lib.ptr(out), _lib.stream_of(x)),
                   "nglod_sdf_fwd_all")
        return [out[i] for i in range(lod + 1)]
