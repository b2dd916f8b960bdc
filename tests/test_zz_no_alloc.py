"""nglod_trace_dense_noalloc (scratch in a static device array, no
allocation) against nglod_trace_dense (caller workspace) through the C ABI:
the same outputs bit for bit, in both forms, on two calls in a row with the
second smaller than the first; the limits are rejected with an error before
anything runs."""
import ctypes

import numpy as np
import pytest
import torch

from nglod_amd import _lib, workloads

pytestmark = pytest.mark.gpu


def _params(return_dist=0.0):
    p = _lib.NglodTraceParams()
    p.num_steps = 128
    p.min_dis = 3e-4
    p.step_size = 1.0
    p.far_clip = 10.0
    p.normals = 1
    p.refine_steps = 4
    p.return_dist = return_dist
    p.near_dist = 1e-2
    p.pass_ms = None
    return p


def _outputs(n, dev):
    return (torch.full((n, 3), -7.0, device=dev), torch.full((n,), -7.0, device=dev),
            torch.full((n,), 7, dtype=torch.uint8, device=dev), torch.full((n, 3), -7.0, device=dev),
            torch.zeros(2, dtype=torch.int64, device=dev), torch.full((n,), -1, dtype=torch.int32, device=dev))


def _trace(lib, m, o, d, lod, form, noalloc):
    n = o.shape[0]
    x, t, hit, nrm, ev, steps = _outputs(n, o.device)
    p = _params()
    s = _lib.stream_of(o)
    if noalloc:
        rc = lib.nglod_trace_dense_noalloc(m, _lib.ptr(o), _lib.ptr(d), n, lod, p, form, _lib.ptr(x), _lib.ptr(t),
                                           _lib.ptr(hit), _lib.ptr(nrm), _lib.ptr(ev), _lib.ptr(steps), s)
        _lib.check(rc, "nglod_trace_dense_noalloc")
    else:
        ws = torch.empty(lib.nglod_trace_dense_workspace_size(n, form), dtype=torch.uint8, device=o.device)
        rc = lib.nglod_trace_dense(m, _lib.ptr(o), _lib.ptr(d), n, lod, p, form, _lib.ptr(x), _lib.ptr(t),
                                   _lib.ptr(hit), _lib.ptr(nrm), _lib.ptr(ws), ws.numel(), _lib.ptr(ev),
                                   _lib.ptr(steps), s)
        _lib.check(rc, "nglod_trace_dense")
    torch.cuda.synchronize()
    return {"x": x, "t": t, "hit": hit, "normal": nrm, "evals": ev, "steps": steps}


@pytest.mark.parametrize("form", ["fp32", "mixed"])
def test_noalloc_matches_workspace_entry(dev, form):
    lib = _lib.load()
    code = _lib.MIXED if form == "mixed" else _lib.FP32
    net = workloads.make_sphere_model(radius=0.6, num_lods=5, device=dev, seed=2, noise=2e-3)
    m = net.descriptor(code)
    lod = 4
    # first call: a 192x160 frame; second: a smaller, different frame
    for (w, h), eye in (((192, 160), (0.0, 0.4, 3.0)), ((64, 48), (1.2, -0.5, 2.6))):
        o, d = workloads.camera_rays(w, h, eye=eye, device=dev)
        print(f"noalloc {form} {w}x{h}", flush=True)
        want = _trace(lib, m, o, d, lod, code, noalloc=False)
        got = _trace(lib, m, o, d, lod, code, noalloc=True)
        for k in want:
            assert torch.equal(got[k], want[k]), (form, w, h, k)
        assert int(want["hit"].sum().item()) > 100


def test_noalloc_limits_are_rejected(dev):
    lib = _lib.load()
    net = workloads.make_sphere_model(radius=0.6, num_lods=2, device=dev)
    m = net.descriptor(_lib.MIXED)
    o = torch.zeros(4, 3, device=dev)
    x, t, hit, nrm, ev, steps = _outputs(4, dev)
    big = (1 << 20) + 1  # beyond NGLOD_NOALLOC_MAX_MIXED_RAYS: rejected before any pointer is read
    rc = lib.nglod_trace_dense_noalloc(m, _lib.ptr(o), _lib.ptr(o), big, 1, _params(), _lib.MIXED, _lib.ptr(x),
                                       _lib.ptr(t), _lib.ptr(hit), _lib.ptr(nrm), None, None, _lib.stream_of(o))
    assert rc == -4  # NGLOD_ERR_CAPACITY
    rc = lib.nglod_trace_dense_noalloc(m, _lib.ptr(o), _lib.ptr(o), 4, 1, _params(return_dist=1e-2), _lib.MIXED,
                                       _lib.ptr(x), _lib.ptr(t), _lib.ptr(hit), _lib.ptr(nrm), None, None,
                                       _lib.stream_of(o))
    assert rc == -2  # NGLOD_ERR_UNSUPPORTED
    assert b"return_dist" in lib.nglod_last_error()
