"""Python binding of the CPU oracle (liboracle.so, built from oracle/*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAX_LODS = 16

_f = ctypes.c_void_p


class OrcModel(ctypes.Structure):
    _fields_ = [
        ("num_lods", ctypes.c_int32),
        ("fdim", ctypes.c_int32),
        ("hidden", ctypes.c_int32),
        ("pos_invariant", ctypes.c_int32),
        ("num_decoders", ctypes.c_int32),
        ("res", ctypes.c_int32 * MAX_LODS),
        ("grid", _f * MAX_LODS),
        ("w0", _f * MAX_LODS),
        ("b0", _f * MAX_LODS),
        ("w1", _f * MAX_LODS),
        ("b1", _f * MAX_LODS),
    ]


class OrcTraceParams(ctypes.Structure):
    _fields_ = [
        ("num_steps", ctypes.c_int32),
        ("min_dis", ctypes.c_double),
        ("step_size", ctypes.c_double),
        ("far_clip", ctypes.c_double),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class ModelArrays:
    """Host copy of an OctreeSDF: corner-major FP32 grids + decoder weights."""

    def __init__(self, grids, res, w0, b0, w1, b1, pos_invariant):
        self.grids = [np.ascontiguousarray(g, dtype=np.float32).reshape(-1) for g in grids]
        self.res = list(res)
        self.w0 = [np.ascontiguousarray(w, dtype=np.float32) for w in w0]
        self.b0 = [np.ascontiguousarray(b, dtype=np.float32) for b in b0]
        self.w1 = [np.ascontiguousarray(w, dtype=np.float32).reshape(-1) for w in w1]
        self.b1 = [np.ascontiguousarray(b, dtype=np.float32).reshape(-1) for b in b1]
        self.pos_invariant = bool(pos_invariant)
        self.fdim = self.grids[0].size // (self.res[0] + 1) ** 3
        self.hidden = self.w0[0].shape[0]
        self._s = None

    @classmethod
    def from_module(cls, net):
        """From an OctreeSDF-like module (reference parameter names)."""
        grids, res = [], []
        for f in net.features:
            fm = f.fm.detach().cpu().float()  # [1, C, D, H, W]
            grids.append(fm[0].permute(1, 2, 3, 0).contiguous().numpy())
            res.append(int(fm.shape[-1]) - 1)
        w0 = [l[0].weight.detach().cpu().numpy() for l in net.louts]
        b0 = [l[0].bias.detach().cpu().numpy() for l in net.louts]
        w1 = [l[2].weight.detach().cpu().numpy() for l in net.louts]
        b1 = [l[2].bias.detach().cpu().numpy() for l in net.louts]
        return cls(grids, res, w0, b0, w1, b1, net.pos_invariant)

    def struct(self):
        if self._s is None:
            s = OrcModel()
            s.num_lods = len(self.grids)
            s.fdim = self.fdim
            s.hidden = self.hidden
            s.pos_invariant = 1 if self.pos_invariant else 0
            s.num_decoders = len(self.w0)
            for i, g in enumerate(self.grids):
                s.res[i] = self.res[i]
                s.grid[i] = g.ctypes.data
            for i in range(len(self.w0)):
                s.w0[i] = self.w0[i].ctypes.data
                s.b0[i] = self.b0[i].ctypes.data
                s.w1[i] = self.w1[i].ctypes.data
                s.b1[i] = self.b1[i].ctypes.data
            self._s = s
        return self._s


def _dt(prec):
    return np.float32 if prec == "f32" else np.float64


def sdf(model, x, lod, prec="f32"):
    x = np.ascontiguousarray(x, dtype=_dt(prec)).reshape(-1, 3)
    out = np.empty(x.shape[0], dtype=_dt(prec))
    fn = getattr(lib(), f"orc_sdf_{prec}")
    fn(ctypes.byref(model.struct()), _p(x), ctypes.c_int64(x.shape[0]), ctypes.c_int(lod), _p(out))
    return out


def sdf_all(model, x, lod, prec="f32"):
    x = np.ascontiguousarray(x, dtype=_dt(prec)).reshape(-1, 3)
    out = np.empty((x.shape[0], lod + 1), dtype=_dt(prec))
    fn = getattr(lib(), f"orc_sdf_all_{prec}")
    fn(ctypes.byref(model.struct()), _p(x), ctypes.c_int64(x.shape[0]), ctypes.c_int(lod), _p(out))
    return out


def aabb(o, d, prec="f32"):
    o = np.ascontiguousarray(o, dtype=_dt(prec)).reshape(-1, 3)
    d = np.ascontiguousarray(d, dtype=_dt(prec)).reshape(-1, 3)
    n = o.shape[0]
    x = np.empty((n, 3), dtype=_dt(prec))
    t = np.empty(n, dtype=_dt(prec))
    hit = np.empty(n, dtype=np.uint8)
    getattr(lib(), f"orc_aabb_{prec}")(_p(o), _p(d), _p(x), _p(t), _p(hit), ctypes.c_int64(n))
    return x, t, hit.astype(bool)


def trace(model, o, d, lod, num_steps=128, min_dis=3e-4, step_size=1.0, far_clip=10.0, prec="f32"):
    """SphereTracer.forward per ray; returns dict x, t, hit, normal, evals
    (all SDF evaluations), steps (in-loop evaluations) and margin (closest
    approach of any termination input to its threshold)."""
    o = np.ascontiguousarray(o, dtype=_dt(prec)).reshape(-1, 3)
    d = np.ascontiguousarray(d, dtype=_dt(prec)).reshape(-1, 3)
    n = o.shape[0]
    P = OrcTraceParams(num_steps, min_dis, step_size, far_clip)
    x = np.empty((n, 3), dtype=_dt(prec))
    t = np.empty(n, dtype=_dt(prec))
    hit = np.empty(n, dtype=np.uint8)
    nrm = np.empty((n, 3), dtype=_dt(prec))
    evals = np.empty(n, dtype=np.int32)
    margin = np.empty(n, dtype=_dt(prec))
    getattr(lib(), f"orc_trace_{prec}")(ctypes.byref(model.struct()), _p(o), _p(d), ctypes.c_int64(n),
                                        ctypes.c_int(lod), ctypes.byref(P), _p(x), _p(t), _p(hit), _p(nrm),
                                        _p(evals), _p(margin))
    hit = hit.astype(bool)
    _, _, entered = aabb(o, d, prec)
    steps = evals - entered.astype(np.int32) - 6 * hit.astype(np.int32)
    return dict(x=x, t=t, hit=hit, normal=nrm, evals=evals, margin=margin, steps=steps)


def normals(model, x, lod, prec="f32"):
    x = np.ascontiguousarray(x, dtype=_dt(prec)).reshape(-1, 3)
    out = np.empty_like(x)
    getattr(lib(), f"orc_normals_{prec}")(ctypes.byref(model.struct()), _p(x), ctypes.c_int64(x.shape[0]),
                                          ctypes.c_int(lod), _p(out))
    return out


# ---- training targets ------------------------------------------------------
def mesh2sdf(pts, tris):
    """FP64 signed distance (restated mesh2sdf) of pts [N,3] to tris [F,3,3]."""
    p = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
    out = np.empty(p.shape[0], dtype=np.float64)
    lib().orc_mesh2sdf_f64(_p(p), ctypes.c_int64(p.shape[0]), _p(t), ctypes.c_int64(t.shape[0]), _p(out))
    return out


def point_tri_dist2(p, a, b, c):
    f = lib().orc_point_tri_dist2_f64
    f.restype = ctypes.c_double
    arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (p, a, b, c)]
    return f(*[_p(v) for v in arr])


def winding(p, tris):
    f = lib().orc_winding_f64
    f.restype = ctypes.c_double
    pp = np.ascontiguousarray(p, dtype=np.float64)
    t = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 9)
    return f(_p(pp), _p(t), ctypes.c_int64(t.shape[0]))


def splitmix64(x):
    f = lib().orc_splitmix64
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint64]
    return f(x)


def sample_uniform(seed, n):
    out = np.empty((n, 3), dtype=np.float32)
    lib().orc_sample_uniform(ctypes.c_uint64(seed), ctypes.c_int64(n), _p(out))
    return out


def sample_surface(tris, cdf, seed, n, sigma):
    t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    c = np.ascontiguousarray(cdf, dtype=np.float64)
    out = np.empty((n, 3), dtype=np.float32)
    face = np.empty(n, dtype=np.int32)
    lib().orc_sample_surface(_p(t), _p(c), ctypes.c_int64(t.shape[0]), ctypes.c_uint64(seed), ctypes.c_int64(n),
                             ctypes.c_double(sigma), _p(out), _p(face))
    return out, face


# ---- sparse octree ------------------------------------------------------------
def morton(x, y, z):
    f = lib().orc_morton
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_uint32] * 3
    return f(x, y, z)


def occupancy(tris, level):
    t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    f = lib().orc_occupancy
    f.restype = ctypes.c_int64
    cap = 1 << 22
    out = np.empty(cap, dtype=np.uint64)
    n = f(_p(t), ctypes.c_int64(t.shape[0]), ctypes.c_int(level), _p(out), ctypes.c_int64(cap))
    assert n <= cap
    return out[:n].copy()


def octree_from_codes(codes, level):
    c = np.ascontiguousarray(codes, dtype=np.uint64)
    f = lib().orc_octree_from_codes
    f.restype = ctypes.c_int64
    cap = max(1, 2 * c.size + 16)
    out = np.empty(cap, dtype=np.uint8)
    counts = np.zeros(level + 1, dtype=np.int64)
    n = f(_p(c), ctypes.c_int64(c.size), ctypes.c_int(level), _p(out), ctypes.c_int64(cap), _p(counts))
    return out[:n].copy(), counts


def octree_points(octree):
    o = np.ascontiguousarray(octree, dtype=np.uint8)
    psize = 1 + int(np.unpackbits(o).sum())
    pts = np.empty((psize, 4), dtype=np.int16)
    pyr = np.zeros(64, dtype=np.int32)
    f = lib().orc_octree_points
    L = f(_p(o), ctypes.c_int64(o.size), _p(pts), ctypes.c_int64(psize), _p(pyr), ctypes.c_int(64))
    return pts, pyr[:2 * (L + 2)].reshape(2, L + 2).copy(), L


def spc_raytrace(octree, points, pyramid, level, o, d, target):
    oc = np.ascontiguousarray(octree, dtype=np.uint8)
    pts = np.ascontiguousarray(points, dtype=np.int16)
    pyr = np.ascontiguousarray(pyramid, dtype=np.int32).reshape(-1)
    oo = np.ascontiguousarray(o, dtype=np.float32).reshape(-1, 3)
    dd = np.ascontiguousarray(d, dtype=np.float32).reshape(-1, 3)
    f = lib().orc_spc_raytrace
    f.restype = ctypes.c_int64
    cap = 1 << 24
    out = np.empty((cap, 2), dtype=np.int32)
    m = f(_p(oc), ctypes.c_int64(oc.size), _p(pts), _p(pyr), ctypes.c_int(level), _p(oo), _p(dd),
          ctypes.c_int64(oo.shape[0]), ctypes.c_int(target), _p(out), ctypes.c_int64(cap))
    assert m <= cap
    return out[:m].copy()


def mark_unique_rays(nuggets):
    n = np.ascontiguousarray(nuggets, dtype=np.int32)
    info = np.empty(n.shape[0], dtype=np.int32)
    lib().orc_mark_unique_rays(_p(n), ctypes.c_int64(n.shape[0]), _p(info))
    return info


def sample_directions(seed, n):
    out = np.empty((n, 3), dtype=np.float32)
    lib().orc_sample_directions(ctypes.c_uint64(seed), ctypes.c_int64(n), _p(out))
    return out


# ---- sparse (SOL) model -------------------------------------------------------
def unique_corners(level_points):
    """geoutils.unique_corners (sdf-net/lib/geoutils.py:60-77): the 8 corners
    of every cell, deduplicated and sorted by (x, y, z) as torch.unique(dim=0)."""
    p = np.asarray(level_points)[:, :3].astype(np.int64)
    offs = np.array([[(j & 4) >> 2, (j & 2) >> 1, j & 1] for j in range(8)], dtype=np.int64)
    allc = (p[None, :, :] + offs[:, None, :]).reshape(-1, 3)
    return np.unique(allc, axis=0).astype(np.int32)


def codes_from_points(x, level):
    """Leaf codes of the cells holding x: floor((x+1)/2*res) per axis in FP32
    (renderutils.py:67-70), clamped to the grid, Morton-coded, sorted, unique."""
    xf = np.asarray(x, dtype=np.float32).reshape(-1, 3)
    res = np.float32(1 << level)
    v = np.floor(((xf + np.float32(1.0)) / np.float32(2.0)) * res)
    v = np.clip(v, 0, res - 1).astype(np.uint32)
    codes = np.array([morton(int(a), int(b), int(c)) for a, b, c in v], dtype=np.uint64)
    return np.unique(codes)


class OrcSol(ctypes.Structure):
    _fields_ = [
        ("target", ctypes.c_int32),
        ("lod", ctypes.c_int32),
        ("din", ctypes.c_int32),
        ("points", _f),
        ("psum", _f),
        ("tv", _f),
        ("parent", _f),
        ("cc", _f),
        ("cf", _f),
        ("w0", _f),
        ("b0", _f),
        ("w1", _f),
        ("b1", _f),
    ]


def sol_trinkets(points, pyramid, target, cc_blocks):
    """Per point of levels 2..target: 8 corner indices into the concatenated
    corner list and the parent point (index_trinket.cuh:31-98)."""
    pts = np.ascontiguousarray(points, dtype=np.int16)
    pyr = np.asarray(pyramid, dtype=np.int32)
    cnt = np.ascontiguousarray(pyr[0], dtype=np.int32)
    psum = np.ascontiguousarray(pyr[1], dtype=np.int32)
    cc = np.ascontiguousarray(np.concatenate(cc_blocks, axis=0), dtype=np.int32)
    sizes = np.array([b.shape[0] for b in cc_blocks], dtype=np.int64)
    off = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]), dtype=np.int64)
    sizes = np.ascontiguousarray(sizes)
    P = pts.shape[0]
    tv = np.full((P, 8), 0xFFFFFFFF, dtype=np.uint32)
    parent = np.full(P, -1, dtype=np.int32)
    lib().orc_sol_trinkets(_p(pts), _p(cnt), _p(psum), ctypes.c_int(target), _p(cc), _p(off), _p(sizes), _p(tv),
                           _p(parent))
    return tv, parent, cc


class SolArrays:
    """Host arrays of a sparse model; .struct() keeps them alive."""

    def __init__(self, points, pyramid, target, tv, parent, cc, cf, w0, b0, w1, b1):
        self.target = int(target)
        self.lod = self.target - 2
        self.points = np.ascontiguousarray(points, dtype=np.int16)
        self.psum = np.ascontiguousarray(np.asarray(pyramid)[1], dtype=np.int32)
        self.tv = np.ascontiguousarray(tv, dtype=np.uint32)
        self.parent = np.ascontiguousarray(parent, dtype=np.int32)
        self.cc = np.ascontiguousarray(cc, dtype=np.int32)
        self.cf = np.ascontiguousarray(cf, dtype=np.float32)
        self.w0 = np.ascontiguousarray(w0, dtype=np.float32)
        self.b0 = np.ascontiguousarray(b0, dtype=np.float32).reshape(-1)
        self.w1 = np.ascontiguousarray(w1, dtype=np.float32).reshape(-1)
        self.b1 = np.ascontiguousarray(b1, dtype=np.float32).reshape(-1)
        self.din = self.w0.shape[1]

    def struct(self):
        return OrcSol(self.target, self.lod, self.din, _p(self.points), _p(self.psum), _p(self.tv), _p(self.parent),
                      _p(self.cc), _p(self.cf), _p(self.w0), _p(self.b0), _p(self.w1), _p(self.b1))


def sol_eval(sol, x, pidx):
    xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
    pi = np.ascontiguousarray(pidx, dtype=np.int64)
    out = np.empty(xx.shape[0], dtype=np.float32)
    s = sol.struct()
    lib().orc_sol_eval(ctypes.byref(s), _p(xx), _p(pi), ctypes.c_int64(xx.shape[0]), _p(out))
    return out


def sol_trace(sol, o, d, offsets, nuggets):
    oo = np.ascontiguousarray(o, dtype=np.float32).reshape(-1, 3)
    dd = np.ascontiguousarray(d, dtype=np.float32).reshape(-1, 3)
    offs = np.ascontiguousarray(offsets, dtype=np.int32)
    nug = np.ascontiguousarray(nuggets, dtype=np.int32).reshape(-1, 2)
    if nug.size == 0:
        nug = np.zeros((1, 2), dtype=np.int32)
    n = oo.shape[0]
    x = np.empty((n, 3), dtype=np.float32)
    t = np.empty(n, dtype=np.float32)
    hit = np.empty(n, dtype=np.uint8)
    nrm = np.empty((n, 3), dtype=np.float32)
    it = np.empty(n, dtype=np.int32)
    s = sol.struct()
    lib().orc_sol_trace(ctypes.byref(s), _p(oo), _p(dd), ctypes.c_int64(n), _p(offs), _p(nug), _p(x), _p(t), _p(hit),
                        _p(nrm), _p(it))
    return {"x": x, "t": t, "hit": hit.astype(bool), "normal": nrm, "iters": it}


def nugget_offsets(nuggets, n):
    """CSR offsets [n+1] of ray-sorted nuggets."""
    r = np.asarray(nuggets, dtype=np.int64).reshape(-1, 2)[:, 0]
    counts = np.bincount(r, minlength=n)[:n]
    return np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)


def cumulative_grids(model, prec="f32"):
    """Per LOD i, the cumulative grid: at each corner of grid i, the sum over
    levels l <= i of grid l's trilinear sample at that corner (the level sum of
    sdf-net/lib/models/OctreeSDF.py:102-110 evaluated on grid i's corners).
    Requires res[l] to divide res[i].  The corner c of grid i lies at c / q,
    q = res[i] / res[l], in grid l's corner units: lower corner c // q (kept
    <= res[l] - 1, weights (0, 1) at the last corner) and weights
    (q - c % q) / q, (c % q) / q.  Summed in FP64, levels in order, each
    level's 8 corners in x, y, z order with weight (wx * wy) * wz.
    prec 'f64': the FP64 sums; 'f32': (FP32 rounding, FP16 rounding of that).
    Corner-major arrays [(res+1)^3 * C]."""
    C = model.fdim
    out = []
    for li, R in enumerate(model.res):
        S = R + 1
        acc = np.zeros((S, S, S, C), dtype=np.float64)  # [z, y, x, c]
        c = np.arange(S)
        for l in range(li + 1):
            r = model.res[l]
            if R % r:
                raise ValueError("grid resolutions do not nest")
            q = R // r
            i0 = c // q
            rem = c - i0 * q
            over = i0 > r - 1
            i0 = np.where(over, r - 1, i0)
            rem = np.where(over, q, rem)
            w = ((q - rem).astype(np.float64) / q, rem.astype(np.float64) / q)
            g = model.grids[l].reshape(r + 1, r + 1, r + 1, C).astype(np.float64)
            for k in range(8):
                bx, by, bz = k & 1, (k >> 1) & 1, k >> 2
                wk = (w[bx][None, None, :] * w[by][None, :, None]) * w[bz][:, None, None]
                gk = g[np.ix_(i0 + bz, i0 + by, i0 + bx)]
                acc = acc + wk[..., None] * gk
        if prec == "f64":
            out.append(acc.reshape(-1))
        else:
            f32 = acc.astype(np.float32).reshape(-1)
            out.append((f32, f32.astype(np.float16)))
    return out


def tetra_gradient(model, x, lod, prec="f32"):
    """diffutils.gradient(..., 'tetrahedron') (sdf-net/lib/diffutils.py:38-60):
    taps at x + h k_j, k_j the corners (1,-1,-1), (-1,-1,1), (-1,1,-1),
    (1,1,1) of a tetrahedron, h = 1/192 in the working precision; gradient
    (k_0 f_0 + k_1 f_1 + k_2 f_2 + k_3 f_3) / (4 h), summed left to right."""
    dt = _dt(prec)
    x = np.ascontiguousarray(x, dtype=dt).reshape(-1, 3)
    h = dt(1.0 / (64.0 * 3.0))
    ks = [np.array(k, dtype=dt) for k in ([1, -1, -1], [-1, -1, 1], [-1, 1, -1], [1, 1, 1])]
    terms = [k[None, :] * sdf(model, x + (h * k)[None, :], lod, prec)[:, None] for k in ks]
    return (((terms[0] + terms[1]) + terms[2]) + terms[3]) / dt((1.0 / (64.0 * 3.0)) * 4.0)


def sdf_grad(model, x, lod):
    """OctreeSDF.sdf(x, lod) (sdf-net/lib/models/OctreeSDF.py:94-155) and its
    gradient with respect to x as torch autograd computes it, in FP64:
    grid_sample's backward (align_corners=True, padding_mode='border': d u /
    d x = r/2, zero where the unclamped u is <= 0 or >= r, trilinear weight
    derivatives in the cell [floor(u), floor(u) + 1]), ReLU's (zero where the
    pre-activation is <= 0) and the xyz columns of a position-dependent
    decoder.  Returns (y [n], g [n, 3])."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    C = model.fdim
    s = np.zeros((n, C))
    levels = []
    for l in range(lod + 1):
        r = model.res[l]
        g = model.grids[l].reshape(r + 1, r + 1, r + 1, C).astype(np.float64)
        u_raw = ((x + 1.0) * 0.5) * r
        u = np.clip(u_raw, 0.0, r)
        i0 = np.minimum(np.floor(u), r - 1).astype(np.int64)
        w = ((i0 + 1) - u, u - i0)
        gate = np.where((u_raw > 0) & (u_raw < r), r * 0.5, 0.0)
        corners = []
        for k in range(8):
            b = (k & 1, (k >> 1) & 1, k >> 2)
            ck = g[i0[:, 2] + b[2], i0[:, 1] + b[1], i0[:, 0] + b[0]]
            wk = w[b[0]][:, 0] * w[b[1]][:, 1] * w[b[2]][:, 2]
            s = s + wk[:, None] * ck
            corners.append((b, ck))
        levels.append((w, gate, corners))
    dec = 0 if len(model.w0) == 1 else lod
    W0 = model.w0[dec].astype(np.float64)
    b0 = model.b0[dec].astype(np.float64)
    w1 = model.w1[dec].astype(np.float64)
    e = s if model.pos_invariant else np.concatenate([x, s], axis=1)
    pre = e @ W0.T + b0
    y = np.maximum(pre, 0.0) @ w1 + float(model.b1[dec][0])
    dy_de = (w1[None, :] * (pre > 0)) @ W0
    off = 0 if model.pos_invariant else 3
    v = dy_de[:, off:]
    grad = np.zeros((n, 3)) if model.pos_invariant else dy_de[:, :3].copy()
    for w, gate, corners in levels:
        d = np.zeros((n, 3))
        for b, ck in corners:
            pk = (ck * v).sum(axis=1)
            for a in range(3):
                o1, o2 = [q for q in range(3) if q != a]
                sign = 1.0 if b[a] else -1.0
                d[:, a] += sign * pk * w[b[o1]][:, o1] * w[b[o2]][:, o2]
        grad += gate * d
    return y, grad


def get_min(model, o, d, lod, num_steps=128, min_dis=3e-4, step_size=1.0, far_clip=10.0):
    """SphereTracer.get_min (sdf-net/lib/tracer/SphereTracer.py:134-218) for a
    batch of rays, FP32, the reference's batch loop restated with numpy masks
    and the oracle's FP32 sdf and box entry: returns x, t, hit and min_x."""
    f32 = np.float32
    o = np.ascontiguousarray(o, dtype=f32).reshape(-1, 3)
    d = np.ascontiguousarray(d, dtype=f32).reshape(-1, 3)
    x, t, _ = aabb(o, d, "f32")
    x = x.copy()
    t = t.reshape(-1).astype(f32)
    D = sdf(model, x, lod, "f32").astype(f32)
    Dp = D.copy()
    mind = D.copy()
    minx = x.copy()
    hit = np.zeros(o.shape[0], dtype=bool)
    md, md3, far, step = f32(min_dis), f32(min_dis * 3.0), f32(far_clip), f32(step_size)
    for _ in range(num_steps):
        hit = np.abs(t) < far
        cond = (np.abs(D) > md) & (np.abs((D + Dp) / f32(2.0)) > md3) & hit
        if not cond.any():
            break
        x = np.where(cond[:, None], o + d * t[:, None], x).astype(f32)
        new = D < mind
        mind[new] = D[new]
        minx[new] = x[new]
        Dp = np.where(cond, D, Dp)
        if cond.any():
            D[cond] = sdf(model, x[cond], lod, "f32").astype(f32) * step
        t = np.where(cond, t + D, t).astype(f32)
    hit = hit & ~(np.abs(x) > 1.0).any(axis=1)
    return dict(x=x, t=t, hit=hit, min_x=minx)
