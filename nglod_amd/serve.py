"""Render serving: continuous batching of camera requests over resident models.

The reference has no serving layer; its nearest analogues are the batch loop
of Renderer.render (sdf-net/lib/renderer.py:118-123) and the per-frame driver
of the SOL viewer (sol-renderer/sdfRenderer.cpp:217-242).

Design:
- Models stay resident with their mixed-precision buffers packed once.
- One device byte arena (the ray pool) holds every live request's rays,
  outputs and trace workspace; a request is admitted when a slot (a HIP
  stream) is free and the pool has room for its bytes, in arrival order.
- A request's work is one camera-ray launch and one persistent trace call
  (nglod_trace_dense) on its slot's stream, so requests that arrive while
  others run start as soon as a slot and bytes free up, and their kernels
  share the device (a small frame's waves fill the tail of a large one).
- Completion is seen by polling each request's event; its pool range and
  slot are released at once.  Per-ray arithmetic does not depend on which
  other rays are in flight, so a result is the same alone and in the mix.
"""
import ctypes
import math
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from . import _lib

_ALIGN = 256


def _up(n):
    return (n + _ALIGN - 1) // _ALIGN * _ALIGN


@dataclass
class RenderRequest:
    model: int
    width: int
    height: int
    lod: int
    eye: tuple
    at: tuple = (0.0, 0.0, 0.0)
    fov: float = 30.0
    arrival: float = 0.0  # seconds after the stream starts
    rid: int = 0

    @property
    def rays(self):
        return self.width * self.height


@dataclass
class RequestResult:
    rid: int
    arrival: float
    start: float
    finish: float
    rays: int
    outputs: Optional[dict] = None

    @property
    def latency(self):
        return self.finish - self.arrival


class RayPool:
    """First-fit allocator over one device byte buffer (256-byte aligned)."""

    def __init__(self, nbytes, device):
        self.nbytes = _up(int(nbytes))
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        self.free_list = [(0, self.nbytes)]  # sorted (offset, size)
        self.used = {}

    def alloc(self, nbytes):
        nbytes = _up(int(nbytes))
        for k, (off, size) in enumerate(self.free_list):
            if size >= nbytes:
                if size == nbytes:
                    self.free_list.pop(k)
                else:
                    self.free_list[k] = (off + nbytes, size - nbytes)
                self.used[off] = nbytes
                return off
        return None

    def free(self, off):
        size = self.used.pop(off)
        fl = self.free_list
        fl.append((off, size))
        fl.sort()
        merged = [fl[0]]
        for o, s in fl[1:]:
            po, ps = merged[-1]
            if po + ps == o:
                merged[-1] = (po, ps + s)
            else:
                merged.append((o, s))
        self.free_list = merged

    @property
    def bytes_in_use(self):
        return sum(self.used.values())

    def view(self, off, nbytes, dtype, shape):
        return self.buf[off:off + nbytes].view(dtype).view(*shape)


@dataclass
class ServeStats:
    results: List[RequestResult] = field(default_factory=list)
    wall_s: float = 0.0
    peak_pool_bytes: int = 0
    peak_inflight: int = 0

    @property
    def rays(self):
        return sum(r.rays for r in self.results)

    def latency_ms(self, q):
        lat = np.array([r.latency for r in self.results]) * 1e3
        return float(np.percentile(lat, q)) if lat.size else float("nan")

    def summary(self):
        return {
            "requests": len(self.results),
            "rays": self.rays,
            "wall_s": self.wall_s,
            "rays_per_s": self.rays / self.wall_s if self.wall_s > 0 else float("nan"),
            "p50_ms": self.latency_ms(50),
            "p99_ms": self.latency_ms(99),
            "peak_pool_bytes": self.peak_pool_bytes,
            "peak_inflight": self.peak_inflight,
        }


class TraceServer:
    """Continuous-batching render server over resident OctreeSDF models.

    slots: concurrent requests on the device (one HIP stream each, default
    priority); pool_bytes: the ray-pool capacity that caps admission."""

    def __init__(self, models, pool_bytes=1 << 30, slots=4, form="mixed", num_steps=128, min_dis=3e-4,
                 far_clip=10.0, step_size=1.0, refine_steps=4, return_dist=0.0, near_dist=1e-2,
                 keep_outputs=False):
        self.models = list(models)
        self.device = self.models[0].features[0].fm.device
        self.form = _lib.MIXED if form == "mixed" else _lib.FP32
        self.descs = [m.descriptor(self.form) for m in self.models]  # parameters fixed while serving
        self.pool = RayPool(pool_bytes, self.device)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)]
        self.keep_outputs = keep_outputs
        p = _lib.NglodTraceParams()
        p.num_steps = int(num_steps)
        p.min_dis = float(min_dis)
        p.step_size = float(step_size)
        p.far_clip = float(far_clip)
        p.normals = 1
        p.refine_steps = int(refine_steps)
        p.return_dist = float(return_dist)
        p.near_dist = float(near_dist)
        self.params = p
        self.lib = _lib.load()

    def request_bytes(self, req):
        n = req.rays
        ws = int(self.lib.nglod_trace_dense_workspace_size(n, self.form))
        return sum(_up(b) for b in (12 * n, 12 * n, 12 * n, 4 * n, n, 12 * n, ws))

    def _launch(self, req, slot, off):
        n = req.rays
        lay, cur = {}, off
        for name, nb in (("o", 12 * n), ("d", 12 * n), ("x", 12 * n), ("t", 4 * n), ("hit", n), ("normal", 12 * n)):
            lay[name] = (cur, nb)
            cur += _up(nb)
        ws_bytes = int(self.lib.nglod_trace_dense_workspace_size(n, self.form))
        pool = self.pool
        o = pool.view(*lay["o"], torch.float32, (n, 3))
        d = pool.view(*lay["d"], torch.float32, (n, 3))
        x = pool.view(*lay["x"], torch.float32, (n, 3))
        t = pool.view(*lay["t"], torch.float32, (n, 1))
        hit = pool.view(*lay["hit"], torch.bool, (n,))
        normal = pool.view(*lay["normal"], torch.float32, (n, 3))
        ws_off = cur
        stream = self.streams[slot]
        sp = ctypes.c_void_p(stream.cuda_stream)
        eye = (ctypes.c_float * 3)(*req.eye)
        at = (ctypes.c_float * 3)(*req.at)
        lib = self.lib
        _lib.check(lib.nglod_camera_rays(eye, at, float(req.fov), req.width, req.height, 0, _lib.ptr(o), _lib.ptr(d),
                                         sp), "nglod_camera_rays")
        _lib.check(lib.nglod_trace_dense(self.descs[req.model], _lib.ptr(o), _lib.ptr(d), n, req.lod, self.params,
                                         self.form, _lib.ptr(x), _lib.ptr(t), _lib.ptr(hit), _lib.ptr(normal),
                                         ctypes.c_void_p(pool.buf.data_ptr() + ws_off), ws_bytes, None, None, sp),
                   "nglod_trace_dense")
        outputs = None
        if self.keep_outputs:
            with torch.cuda.stream(stream):
                outputs = {"x": x.clone(), "depth": t.clone(), "hit": hit.clone(), "normal": normal.clone(),
                           "ray_o": o.clone(), "ray_d": d.clone()}
        ev = torch.cuda.Event()
        ev.record(stream)
        return ev, outputs

    def run(self, requests):
        """Serves `requests` (arrival times in seconds from now) to completion."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in reqs:
            if self.request_bytes(r) > self.pool.nbytes:
                raise ValueError(f"request {r.rid} ({r.width}x{r.height}) needs more than the whole ray pool")
            if not 0 <= r.model < len(self.models):
                raise ValueError(f"request {r.rid}: no model {r.model}")
        stats = ServeStats()
        free_slots = list(range(len(self.streams)))
        waiting = []
        inflight = []  # (req, slot, off, event, start, outputs)
        nxt = 0
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        while nxt < len(reqs) or waiting or inflight:
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                waiting.append(reqs[nxt])
                nxt += 1
            still = []
            for item in inflight:
                req, slot, off, ev, start, outs = item
                if ev.query():
                    fin = time.perf_counter() - t0
                    self.pool.free(off)
                    free_slots.append(slot)
                    stats.results.append(RequestResult(req.rid, req.arrival, start, fin, req.rays, outs))
                else:
                    still.append(item)
            inflight = still
            # admit in arrival order while a slot and pool bytes are free
            while waiting and free_slots:
                req = waiting[0]
                off = self.pool.alloc(self.request_bytes(req))
                if off is None:
                    break
                waiting.pop(0)
                slot = free_slots.pop(0)
                start = time.perf_counter() - t0
                ev, outs = self._launch(req, slot, off)
                inflight.append((req, slot, off, ev, start, outs))
            stats.peak_pool_bytes = max(stats.peak_pool_bytes, self.pool.bytes_in_use)
            stats.peak_inflight = max(stats.peak_inflight, len(inflight))
            if not inflight and not waiting and nxt < len(reqs):
                # idle until the next arrival
                delay = reqs[nxt].arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(min(delay, 1e-3))
        stats.wall_s = time.perf_counter() - t0 - (reqs[0].arrival if reqs else 0.0)
        stats.results.sort(key=lambda r: r.rid)
        return stats


RESOLUTIONS = ((64, 64), (256, 256), (640, 480), (1280, 720), (1920, 1080))


def orbit_eye(rng):
    """Camera on a sphere of radius U[2.5, 4] around the origin, direction
    uniform on S^2."""
    v = rng.normal(size=3)
    v /= np.linalg.norm(v)
    return tuple(float(c) for c in v * rng.uniform(2.5, 4.0))


def request_stream(n, rate, num_models, seed=0, resolutions=RESOLUTIONS, lods=(0, 1, 2, 3, 4)):
    """n requests with Poisson arrivals at `rate` per second (rate = inf: all
    at t = 0), uniform resolution, LOD and model, orbit cameras."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        if math.isfinite(rate):
            t += rng.exponential(1.0 / rate)
        w, h = resolutions[rng.integers(len(resolutions))]
        out.append(RenderRequest(model=int(rng.integers(num_models)), width=w, height=h,
                                 lod=int(lods[rng.integers(len(lods))]), eye=orbit_eye(rng),
                                 arrival=t if math.isfinite(rate) else 0.0, rid=i))
    return out
