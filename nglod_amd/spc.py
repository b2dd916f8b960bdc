"""Structured point cloud (sparse octree) on the GPU: build from a mesh or
from octree bytes, and ray-octree traversal to (ray, voxel) nuggets.

Mirrors the reference's SPC class (sol-renderer/include/spc/spc/SPC.h:101-140,
SPC.cu) and spc_raytrace (spc_raytrace_cuda.cpp:141-199).  octree: uint8 [osize]
(level 0 first); points: int16 [P, 4]; pyramid: int32 [2, L+2] on the host
(row 0 points per level, row 1 exclusive sums).
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _call_with_workspace(fn, device, *args):
    """Runs an nglod_spc_* setup call, growing its workspace on request."""
    need = ctypes.c_size_t(0)
    ws = torch.empty(1 << 20, dtype=torch.uint8, device=device)
    for _ in range(4):
        rc = fn(*args, _lib.ptr(ws), ws.numel(), ctypes.byref(need))
        if rc != _lib.ERR_WORKSPACE:
            return rc
        ws = torch.empty(int(need.value) + 256, dtype=torch.uint8, device=device)
    return rc


class SPC:
    def __init__(self, octree, device="cuda"):
        """From octree bytes (uint8 tensor or array), as SPC::LoadNPZ + SetGeometry."""
        self.device = torch.device(device)
        self.octree = torch.as_tensor(np.asarray(octree.cpu() if torch.is_tensor(octree) else octree,
                                                 dtype=np.uint8)).to(self.device)
        osize = self.octree.numel()
        self.prefix = torch.empty(osize + 1, dtype=torch.int32, device=self.device)
        pyr = (ctypes.c_int32 * 64)()
        L = ctypes.c_int32(0)
        ps = ctypes.c_int64(0)
        lib = _lib.load()
        # first call sizes the points
        cap = 1
        pts = torch.empty(cap, 4, dtype=torch.int16, device=self.device)
        s = _lib.stream_of(self.octree)
        for _ in range(2):
            rc = _call_with_workspace(lambda *a: lib.nglod_spc_points(*a, s), self.device,
                                      _lib.ptr(self.octree), osize, _lib.ptr(pts), pts.shape[0],
                                      ctypes.cast(pyr, ctypes.c_void_p), 64,
                                      ctypes.byref(L), ctypes.byref(ps), _lib.ptr(self.prefix))
            if rc == _lib.ERR_CAPACITY and ps.value > pts.shape[0]:
                pts = torch.empty(ps.value, 4, dtype=torch.int16, device=self.device)
                continue
            break
        _lib.check(rc, "nglod_spc_points")
        self.points = pts
        self.level = L.value
        self.pyramid = torch.tensor([list(pyr[0:L.value + 2]), list(pyr[L.value + 2:2 * (L.value + 2)])],
                                    dtype=torch.int32)
        self._pyr_c = (ctypes.c_int32 * (2 * (self.level + 2)))(*self.pyramid.reshape(-1).tolist())
        self._ws_async = {}

    @classmethod
    def from_mesh(cls, tris, level, device="cuda"):
        """Octree of the level-`level` cells the triangles [F, 3, 3] overlap."""
        dev = torch.device(device)
        t = torch.as_tensor(tris, dtype=torch.float32).to(dev).contiguous()
        nf = t.shape[0]
        lib = _lib.load()
        s = _lib.stream_of(t)
        count = ctypes.c_int64(0)
        codes = torch.empty(max(nf * 64, 1024), dtype=torch.int32, device=dev)
        for _ in range(2):
            rc = _call_with_workspace(lambda *a: lib.nglod_spc_occupancy(*a, s), dev, _lib.ptr(t), nf, level,
                                      _lib.ptr(codes), codes.numel(), ctypes.byref(count))
            if rc == _lib.ERR_CAPACITY:
                codes = torch.empty(count.value, dtype=torch.int32, device=dev)
                continue
            break
        _lib.check(rc, "nglod_spc_occupancy")
        codes = codes[:count.value].contiguous()
        return cls.from_codes(codes, level, dev)

    @classmethod
    def from_codes(cls, codes, level, device="cuda"):
        dev = torch.device(device)
        c = torch.as_tensor(codes).to(dev).to(torch.int32).contiguous()
        lib = _lib.load()
        s = _lib.stream_of(c)
        osize = ctypes.c_int64(0)
        counts = (ctypes.c_int32 * 16)()
        oct_ = torch.empty(c.numel() * 2 + 16, dtype=torch.uint8, device=dev)
        for _ in range(2):
            rc = _call_with_workspace(lambda *a: lib.nglod_spc_from_codes(*a, s), dev, _lib.ptr(c), c.numel(), level,
                                      _lib.ptr(oct_), oct_.numel(), ctypes.byref(osize),
                                      ctypes.cast(counts, ctypes.c_void_p))
            if rc == _lib.ERR_CAPACITY:
                oct_ = torch.empty(osize.value, dtype=torch.uint8, device=dev)
                continue
            break
        _lib.check(rc, "nglod_spc_from_codes")
        return cls(oct_[:osize.value].contiguous(), dev)

    def points_at(self, level):
        """Points of one level (SPC::GetPoints(l) without the levels below)."""
        beg = int(self.pyramid[1, level])
        return self.points[beg:beg + int(self.pyramid[0, level])]

    def raytrace(self, ray_o, ray_d, target):
        """Nuggets [M, 2] int32 and per-ray CSR offsets [N+1] (int32)."""
        _lib.require_gpu(ray_o, ray_d)
        o = ray_o.detach().float().contiguous()
        d = ray_d.detach().float().contiguous()
        n = o.shape[0]
        lib = _lib.load()
        s = _lib.stream_of(o)
        offs = torch.empty(n + 1, dtype=torch.int32, device=o.device)
        count = ctypes.c_int64(0)
        nug = torch.empty(max(4 * n, 1024), 2, dtype=torch.int32, device=o.device)
        for _ in range(2):
            rc = _call_with_workspace(lambda *a: lib.nglod_spc_raytrace(*a, s), o.device, _lib.ptr(self.octree),
                                      _lib.ptr(self.prefix), _lib.ptr(self.points),
                                      ctypes.cast(self._pyr_c, ctypes.c_void_p), self.level,
                                      _lib.ptr(o), _lib.ptr(d), n, target, _lib.ptr(nug), nug.shape[0],
                                      ctypes.byref(count), _lib.ptr(offs))
            if rc == _lib.ERR_CAPACITY:
                nug = torch.empty(count.value, 2, dtype=torch.int32, device=o.device)
                continue
            break
        _lib.check(rc, "nglod_spc_raytrace")
        return nug[:count.value], offs

    def raytrace_async(self, ray_o, ray_d, target, capacity=None, out=None):
        """The single-pass traversal without a host read-back (targets up to
        9): returns (nuggets [capacity, 2] int32, offsets [N+1] int32,
        status [2] int32 on the device: the nugget count, and flags, bit 0 =
        more nuggets than `capacity` (the list holds the first `capacity`;
        offsets are complete), bit 1 = traversal error).  Rows past the count
        are not written.  capacity defaults to 8 per ray; `out` reuses a
        (nuggets, offsets, status) triple of the right sizes."""
        _lib.require_gpu(ray_o, ray_d)
        o = ray_o.detach().float().contiguous()
        d = ray_d.detach().float().contiguous()
        n = o.shape[0]
        dev = o.device
        lib = _lib.load()
        if out is None:
            cap = 8 * n if capacity is None else int(capacity)
            nug = torch.empty(max(cap, 1), 2, dtype=torch.int32, device=dev)
            offs = torch.empty(n + 1, dtype=torch.int32, device=dev)
            status = torch.empty(2, dtype=torch.int32, device=dev)
        else:
            nug, offs, status = out
            cap = nug.shape[0] if capacity is None else int(capacity)
            if offs.numel() < n + 1 or status.numel() < 2 or cap > nug.shape[0]:
                raise ValueError("raytrace_async: `out` buffers too small")
        need = lib.nglod_spc_raytrace_async_workspace_size(n)
        ws = self._ws_async.get(dev)
        if ws is None or ws.numel() < need:
            ws = torch.empty(max(need, 256), dtype=torch.uint8, device=dev)
            self._ws_async[dev] = ws
        _lib.check(lib.nglod_spc_raytrace_async(_lib.ptr(self.octree), _lib.ptr(self.prefix), _lib.ptr(self.points),
                                                ctypes.cast(self._pyr_c, ctypes.c_void_p), self.level, _lib.ptr(o),
                                                _lib.ptr(d), n, target, _lib.ptr(nug), cap, _lib.ptr(offs),
                                                _lib.ptr(status), _lib.ptr(ws), ws.numel(), _lib.stream_of(o)),
                   "nglod_spc_raytrace_async")
        return nug, offs, status

    @staticmethod
    def mark_unique_rays(nuggets):
        info = torch.empty(nuggets.shape[0], dtype=torch.int32, device=nuggets.device)
        _lib.check(_lib.load().nglod_mark_unique_rays(_lib.ptr(nuggets), nuggets.shape[0], _lib.ptr(info),
                                                      _lib.stream_of(nuggets)), "nglod_mark_unique_rays")
        return info
