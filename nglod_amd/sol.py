"""Sparse ("SOL") model: export from a dense OctreeSDF, the npz on-disk format,
and the sparse sphere march.

Mirrors SOL_NGLOD (sdf-net/lib/models/SOL_NGLOD.py:31-100: occupied cells,
unique corners, FP16 corner features and decoders, save to npz) and the C++
SDF class that consumes the export (sol-renderer/SDF.cu: loadWeights :65-139,
initTrinkets :141-216, getNormal :218-295, sphereTrace :297-472).  Every
compute step runs in libnglod_hip.so; torch holds the device buffers.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .spc import SPC, _call_with_workspace
from .tracer import RenderBuffer, SphereTracer

NPZ_KEYS = ("octree", "cc", "cf", "w0", "b0", "w1", "b1", "pyramid")


def voxel_sparsify(n, net, lod, seed=0, max_batches=1000, device=None):
    """Level-(lod+2) cells that hold traced surface points
    (renderutils.voxel_sparsify / sample_surface, sdf-net/lib/renderutils.py:30-73):
    batches of n random lines through the volume (a point uniform in [-1,1]^3
    and a direction uniform on the sphere), traced at `lod`, until at least n
    rays hit.  Each ray starts 4 units back along its line, outside the box:
    the tracer reports an origin inside the box as a hit at depth 0 (the box
    entry skips such rays, SphereTracer.py:53 with sol_nglod_kernel.cu:109-112),
    which would fill every cell.  A hit counts when |sdf(x)| is below half a
    cell (rays stopped by the step cap are not surface points).  Every such hit
    of the traced batches is kept (the reference keeps the first n).  Rays come
    from seeded counter streams, so the set is reproducible."""
    dev = torch.device(device) if device is not None else net.features[0].fm.device
    lib = _lib.load()
    level = lod + 2
    tracer = SphereTracer()
    saved = net.lod
    net.lod = lod
    xs, masks, hits = [], [], 0
    try:
        for b in range(max_batches):
            o = torch.empty(n, 3, device=dev, dtype=torch.float32)
            d = torch.empty(n, 3, device=dev, dtype=torch.float32)
            s = _lib.stream_of(o)
            _lib.check(lib.nglod_sample_uniform(n, 2 * (seed * max_batches + b), _lib.ptr(o), s),
                       "nglod_sample_uniform")
            _lib.check(lib.nglod_sample_directions(n, 2 * (seed * max_batches + b) + 1, _lib.ptr(d), s),
                       "nglod_sample_directions")
            o.sub_(4.0 * d)
            rb = tracer(net, o, d)
            near = net.sdf_hip(rb.x, lod)[:, 0].abs() < 1.0 / (1 << level)
            keep = rb.hit & near
            xs.append(rb.x)
            masks.append(keep)
            hits += int(keep.sum().item())
            if hits >= n:
                break
    finally:
        net.lod = saved
    x = torch.cat(xs).contiguous()
    m = torch.cat(masks).to(torch.uint8).contiguous()
    count = ctypes.c_int64(0)
    codes = torch.empty(max(hits, 1), dtype=torch.int32, device=dev)
    s = _lib.stream_of(x)
    rc = _call_with_workspace(lambda *a: lib.nglod_spc_codes_from_points(*a, s), dev, _lib.ptr(x), _lib.ptr(m),
                              x.shape[0], level, _lib.ptr(codes), codes.numel(), ctypes.byref(count))
    _lib.check(rc, "nglod_spc_codes_from_points")
    if count.value == 0:
        raise RuntimeError("voxel_sparsify: no ray hit the surface")
    return SPC.from_codes(codes[:count.value], level, dev)


class SolNGLOD:
    """Sparse model on the GPU.

    spc: the octree (its levels 2..2+L-1 carry LODs 0..L-1); cc: int32
    [Nc, 3] corner coordinates, one (x, y, z)-sorted block per level; counts:
    corners per level; cf: [Nc, 32] float32; w0/b0/w1/b1: per-LOD decoder
    lists (float32 device tensors)."""

    def __init__(self, spc, cc, counts, cf, w0, b0, w1, b1):
        self.spc = spc
        self.device = spc.device
        self.cc = cc.to(self.device, torch.int32).contiguous()
        self.counts = [int(c) for c in counts]
        self.cf = cf.to(self.device, torch.float32).contiguous()
        self.w0 = [w.to(self.device, torch.float32).contiguous() for w in w0]
        self.b0 = [b.to(self.device, torch.float32).reshape(-1).contiguous() for b in b0]
        self.w1 = [w.to(self.device, torch.float32).reshape(-1).contiguous() for w in w1]
        self.b1 = [b.to(self.device, torch.float32).reshape(-1).contiguous() for b in b1]
        self.num_lods = len(self.counts)
        self.lod = self.num_lods - 1
        if self.num_lods + 1 > spc.level:
            raise ValueError(f"octree level {spc.level} cannot carry {self.num_lods} LODs (needs {self.num_lods + 1})")
        if len(self.w0) < self.num_lods:
            raise ValueError("one decoder per LOD is required")
        if sum(self.counts) != self.cc.shape[0] or self.cf.shape != (self.cc.shape[0], 32):
            raise ValueError("cc / cf / counts disagree")
        self._check_corner_blocks()
        P = spc.points.shape[0]
        self.trinket_v = torch.zeros(P, 8, dtype=torch.int32, device=self.device)
        self.parent = torch.full((P,), -1, dtype=torch.int32, device=self.device)
        ws = torch.empty(256, dtype=torch.uint8, device=self.device)
        cnt = (ctypes.c_int64 * len(self.counts))(*self.counts)
        lib = _lib.load()
        _lib.check(lib.nglod_sol_trinkets(_lib.ptr(spc.points), ctypes.cast(spc._pyr_c, ctypes.c_void_p), spc.level,
                                          self.num_lods + 1, _lib.ptr(self.cc), ctypes.cast(cnt, ctypes.c_void_p),
                                          _lib.ptr(self.trinket_v), _lib.ptr(self.parent), _lib.ptr(ws), ws.numel(),
                                          _lib.stream_of(ws)), "nglod_sol_trinkets")
        self._ws = torch.empty(_lib.load().nglod_trace_sparse_workspace_size(0), dtype=torch.uint8, device=self.device)
        self.last_evals = None

    def _check_corner_blocks(self):
        """Every level block must be strictly (x, y, z)-increasing (the
        trinket search relies on it)."""
        cc = self.cc.cpu().numpy().astype(np.int64)
        off = 0
        for c in self.counts:
            blk = cc[off:off + c]
            if c > 1:
                key = (blk[:, 0] << 42) | (blk[:, 1] << 21) | blk[:, 2]
                if not np.all(np.diff(key) > 0):
                    raise ValueError("corner block not sorted by (x, y, z) / not unique")
            off += c

    # -- export -----------------------------------------------------------
    @classmethod
    def from_net(cls, net, spc=None, num_samples=2_000_000, seed=0, fp16=True):
        """SOL_NGLOD(net) (SOL_NGLOD.py:31-76).  spc: an octree of level
        num_lods+1 or more (default: voxel_sparsify of the net).  fp16: round
        features and decoders to FP16 as the export stores them."""
        if spc is None:
            spc = voxel_sparsify(num_samples, net, net.num_lods - 1, seed)
        dev = spc.device
        L = min(net.num_lods, spc.level - 1)
        target = L + 1
        lib = _lib.load()
        counts = (ctypes.c_int64 * (target - 1))()
        cap = 8 * int(sum(int(spc.pyramid[0, l]) for l in range(2, target + 1)))
        cc = torch.empty(max(cap, 1), 3, dtype=torch.int32, device=dev)
        s = _lib.stream_of(cc)
        rc = _call_with_workspace(lambda *a: lib.nglod_sol_corners(*a, s), dev, _lib.ptr(spc.points),
                                  ctypes.cast(spc._pyr_c, ctypes.c_void_p), spc.level, target, _lib.ptr(cc), cap,
                                  ctypes.cast(counts, ctypes.c_void_p))
        _lib.check(rc, "nglod_sol_corners")
        nc = sum(counts)
        cc = cc[:nc].contiguous()
        cf = torch.empty(nc, 32, dtype=torch.float32, device=dev)
        m = net.descriptor(_lib.FP32)
        _lib.check(lib.nglod_sol_features(m, target, _lib.ptr(cc), ctypes.cast(counts, ctypes.c_void_p),
                                          _lib.ptr(cf), s), "nglod_sol_features")
        rnd = (lambda t: t.detach().half().float()) if fp16 else (lambda t: t.detach().float())
        dec = [net.louts[net.decoder_index(i)] for i in range(L)]
        return cls(spc, cc, list(counts), rnd(cf), [rnd(d[0].weight) for d in dec], [rnd(d[0].bias) for d in dec],
                   [rnd(d[2].weight) for d in dec], [rnd(d[2].bias) for d in dec])

    def save(self, path):
        """npz with the keys SOL_NGLOD.save writes (SOL_NGLOD.py:78-100): octree
        u8, cc u8 [Nc, 3], cf f16 [Nc, 32], w0 f16 [L, 128, 35], b0 f16 [L, 128],
        w1 f16 [L, 1, 128], b1 f16 [L, 1], pyramid int64 [L] corners per LOD.
        A pos_invariant model stores w0 [L, 128, 32], which only this loader
        reads (the reference's sparse sampler always prepends xyz)."""
        if self.spc.level > 7:
            raise ValueError("cc is stored as uint8: octree levels above 7 do not fit")
        np.savez_compressed(
            path,
            octree=self.spc.octree.cpu().numpy().astype(np.uint8),
            cc=self.cc.cpu().numpy().astype(np.uint8),
            cf=self.cf.cpu().numpy().astype(np.float16),
            w0=torch.stack(self.w0).cpu().numpy().astype(np.float16),
            b0=torch.stack(self.b0).cpu().numpy().astype(np.float16),
            w1=torch.stack([w.reshape(1, -1) for w in self.w1]).cpu().numpy().astype(np.float16),
            b1=torch.stack(self.b1).cpu().numpy().astype(np.float16),
            pyramid=np.asarray(self.counts, dtype=np.int64),
        )

    @classmethod
    def load(cls, path, device="cuda"):
        """SDF::loadWeights + SPC::LoadNPZ + SDF::initTrinkets
        (SDF.cu:65-216, SPC.cu:54-77).  Plain npz arrays, no pickle."""
        with np.load(path, allow_pickle=False) as f:
            missing = [k for k in NPZ_KEYS if k not in f.files]
            if missing:
                raise ValueError(f"{path}: missing keys {missing}")
            arr = {k: f[k] for k in NPZ_KEYS}
        spc = SPC(arr["octree"], device)
        L = arr["w0"].shape[0]
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a).astype(np.float32))
        return cls(spc, torch.from_numpy(arr["cc"].astype(np.int32)), arr["pyramid"].tolist(), t(arr["cf"]),
                   [t(arr["w0"][i]) for i in range(L)], [t(arr["b0"][i]) for i in range(L)],
                   [t(arr["w1"][i]) for i in range(L)], [t(arr["b1"][i]) for i in range(L)])

    # -- queries ----------------------------------------------------------
    def descriptor(self, lod=None):
        lod = self.lod if lod is None else lod
        if not 0 <= lod < self.num_lods:
            raise ValueError(f"lod {lod} out of range 0..{self.num_lods - 1}")
        s = _lib.NglodSol()
        s.target = lod + 2
        s.lod = lod
        s.din = self.w0[lod].shape[1]
        s.psum_target = int(self.spc.pyramid[1, lod + 2])
        s.points = self.spc.points.data_ptr()
        s.trinket_v = self.trinket_v.data_ptr()
        s.parent = self.parent.data_ptr()
        s.cc = self.cc.data_ptr()
        s.cf = self.cf.data_ptr()
        s.w0 = self.w0[lod].data_ptr()
        s.b0 = self.b0[lod].data_ptr()
        s.w1 = self.w1[lod].data_ptr()
        s.b1 = self.b1[lod].data_ptr()
        return s

    def sdf(self, x, pidx, lod=None):
        """Sparse SDF at x [N, 3] in level-(lod+2) cells pidx [N] (index within
        the level)."""
        _lib.require_gpu(x, pidx)
        xf = x.detach().reshape(-1, 3).float().contiguous()
        pi = pidx.to(torch.int32).contiguous()
        out = torch.empty(xf.shape[0], 1, dtype=torch.float32, device=xf.device)
        _lib.check(_lib.load().nglod_sol_sdf(self.descriptor(lod), _lib.ptr(xf), _lib.ptr(pi), xf.shape[0],
                                             _lib.ptr(out), _lib.stream_of(xf)), "nglod_sol_sdf")
        return out

    def trace(self, ray_o, ray_d, lod=None, nuggets=None, iters=None):
        """SDF::sphereTrace + getNormal on rays [N, 3]: RenderBuffer(x, depth,
        hit, normal).  nuggets: optional (nuggets, offsets) from
        spc.raytrace(ray_o, ray_d, lod + 2); iters: optional int32 [N] output."""
        _lib.require_gpu(ray_o, ray_d)
        lod = self.lod if lod is None else lod
        o = ray_o.detach().float().contiguous()
        d = ray_d.detach().float().contiguous()
        n = o.shape[0]
        if nuggets is None:
            nuggets = self.spc.raytrace(o, d, lod + 2)
        nug, offs = nuggets
        dev = o.device
        x = torch.empty(n, 3, device=dev, dtype=torch.float32)
        t = torch.empty(n, 1, device=dev, dtype=torch.float32)
        hit = torch.empty(n, device=dev, dtype=torch.bool)
        normal = torch.empty(n, 3, device=dev, dtype=torch.float32)
        evals = torch.zeros(1, device=dev, dtype=torch.int64)
        _lib.check(_lib.load().nglod_trace_sparse(self.descriptor(lod), _lib.ptr(o), _lib.ptr(d), n, _lib.ptr(offs),
                                                  _lib.ptr(nug), _lib.ptr(x), _lib.ptr(t), _lib.ptr(hit),
                                                  _lib.ptr(normal), _lib.ptr(iters), _lib.ptr(self._ws),
                                                  self._ws.numel(), _lib.ptr(evals), _lib.stream_of(o)),
                   "nglod_trace_sparse")
        self.last_evals = evals
        return RenderBuffer(x=x, depth=t, hit=hit, normal=normal)
