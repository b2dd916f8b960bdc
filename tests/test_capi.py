"""The C-ABI library loads and exports every symbol include/*.h declares;
host-side argument checks return their status without touching a GPU."""
import ctypes
import glob
import os
import re

import pytest

from nglod_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(nglod_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "nglod_trace_dense" in names and "nglod_sdf_fwd" in names and "nglod_aabb" in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_header_constants_match_the_binding():
    text = open(os.path.join(ROOT, "include", "nglod_hip.h")).read()
    defines = dict(re.findall(r"^#define\s+(NGLOD_\w+)\s+(-?\d+)\s*$", text, flags=re.M))
    assert int(defines["NGLOD_MAX_LODS"]) == _lib.MAX_LODS
    assert int(defines["NGLOD_POOL_COUNTER_BYTES"]) == _lib.POOL_COUNTER_BYTES


def test_abi_version():
    assert _lib.load().nglod_abi_version() == 1


def test_host_argument_checks():
    lib = _lib.load()
    assert lib.nglod_aabb(None, None, -1, None, None, None, None) == _lib.ERR_ARG
    assert lib.nglod_aabb(None, None, 0, None, None, None, None) == _lib.OK  # empty input
    assert lib.nglod_aabb(None, None, 5, None, None, None, None) == _lib.ERR_ARG
    assert b"NULL" in lib.nglod_last_error()
    m = _lib.NglodModel()
    m.num_lods = 5
    m.feature_dim = 16
    m.hidden_dim = 128
    m.num_decoders = 5
    assert lib.nglod_sdf_fwd(ctypes.byref(m), None, 10, 4, 0, None, None) == _lib.ERR_UNSUPPORTED
    m.feature_dim = 32
    assert lib.nglod_sdf_fwd(ctypes.byref(m), None, 10, 7, 0, None, None) == _lib.ERR_ARG  # lod out of range
    p = _lib.NglodTraceParams()
    p.num_steps = 0
    for i in range(5):
        m.res[i] = 4 << i
        m.grid[i] = 4096 * (i + 1)
        m.w0[i] = m.b0[i] = m.w1[i] = m.b1[i] = 4096
    assert lib.nglod_trace_dense(ctypes.byref(m), None, None, 10, 4, ctypes.byref(p), 0, None, None, None, None,
                                 None, 0, None, None, None) == _lib.ERR_ARG
    p.num_steps = 8
    assert lib.nglod_trace_dense(ctypes.byref(m), 16, 16, 10, 4, ctypes.byref(p), 0, 16, 16, 16, 16,
                                 None, 0, None, None, None) == _lib.ERR_WORKSPACE


def test_shading_argument_checks():
    """The shading entry points reject bad arguments before any launch, and
    accept empty frames."""
    lib = _lib.load()
    light = (ctypes.c_float * 3)(-1.5, 4.5, -1.5)
    assert lib.nglod_shadow_rays(None, None, 0, 0.0, light, None, None, None, None, None, None, None, None, None,
                                 None) == _lib.OK
    assert lib.nglod_shadow_rays(None, None, -1, 0.0, light, None, None, None, None, None, None, None, None, None,
                                 None) == _lib.ERR_ARG
    assert lib.nglod_shadow_rays(None, None, 4, 0.0, light, None, None, None, None, None, None, None, None, None,
                                 None) == _lib.ERR_ARG
    assert lib.nglod_shadow_rays(16, 16, 4, 0.0, None, None, 16, 16, 16, 16, 16, 16, 16, 16, None) == _lib.ERR_ARG
    assert lib.nglod_ao_workspace_size() >= 4
    m = _lib.NglodModel()
    m.num_lods, m.feature_dim, m.hidden_dim, m.num_decoders = 3, 32, 128, 1
    for i in range(3):
        m.res[i] = 4 << i
        m.grid[i] = 4096 * (i + 1)
    m.w0[0] = m.b0[0] = m.w1[0] = m.b1[0] = 4096
    assert lib.nglod_ao(ctypes.byref(m), 16, 16, 16, None, 8, 2, 7, 16, 16, 256, None) == _lib.ERR_ARG  # form
    assert lib.nglod_ao(ctypes.byref(m), 16, 16, 16, None, 8, 2, 0, 16, 16, 8, None) == _lib.ERR_WORKSPACE
    assert lib.nglod_ao(ctypes.byref(m), None, 16, 16, None, 8, 2, 0, 16, 16, 256, None) == _lib.ERR_ARG
    assert lib.nglod_ao(ctypes.byref(m), None, None, None, None, 0, 2, 0, None, None, 0, None) == _lib.OK
    assert lib.nglod_shadow_map(None, None, 0, 7, None, None, None) == _lib.OK
    assert lib.nglod_shadow_map(None, None, -1, 7, None, None, None) == _lib.ERR_ARG
    assert lib.nglod_shadow_map(None, None, 3, 7, None, None, None) == _lib.ERR_ARG
    assert lib.nglod_matcap_shade(16, 16, 16, 8, None, 16, 1, 8, 3, None, None, 16, None) == _lib.ERR_ARG  # tw < 2
    assert lib.nglod_matcap_shade(16, 16, 16, 8, None, 16, 8, 8, 2, None, None, 16, None) == _lib.ERR_ARG  # 2 ch
    assert lib.nglod_matcap_shade(None, 16, 16, 8, None, 16, 8, 8, 3, None, None, 16, None) == _lib.ERR_ARG
    assert lib.nglod_matcap_shade(None, None, None, 0, None, None, 8, 8, 3, None, None, None, None) == _lib.OK


def test_gradient_and_pack_argument_checks():
    """nglod_gradient, nglod_sdf_grad, nglod_trace_min and nglod_pack_mixed
    reject bad arguments before any launch (fake device pointers are never
    dereferenced)."""
    lib = _lib.load()
    m = _lib.NglodModel()
    m.num_lods, m.feature_dim, m.hidden_dim, m.num_decoders = 3, 32, 128, 1
    for i in range(3):
        m.res[i] = 4 << i
        m.grid[i] = 4096 * (i + 1)
    m.w0[0] = m.b0[0] = m.w1[0] = m.b1[0] = 4096
    assert lib.nglod_gradient(ctypes.byref(m), 16, 8, 2, 0, 7, None, 0, 16, None) == _lib.ERR_ARG  # method
    assert lib.nglod_gradient(ctypes.byref(m), 16, 8, 5, 0, 2, None, 0, 16, None) == _lib.ERR_ARG  # lod
    assert lib.nglod_gradient(ctypes.byref(m), None, 8, 2, 0, 2, None, 0, 16, None) == _lib.ERR_ARG  # NULL x
    assert lib.nglod_gradient(ctypes.byref(m), None, 0, 2, 0, 2, None, 0, None, None) == _lib.OK  # empty
    # the mixed form needs the cumulative grids nglod_pack_mixed builds
    assert lib.nglod_gradient(ctypes.byref(m), 16, 8, 2, 1, 2, None, 0, 16, None) == _lib.ERR_ARG
    assert b"cumulative" in lib.nglod_last_error()
    assert lib.nglod_sdf_grad(ctypes.byref(m), 16, 8, 2, 0, None, None, None) == _lib.ERR_ARG  # NULL grad
    p = _lib.NglodTraceParams()
    p.num_steps = 0
    assert lib.nglod_trace_min(ctypes.byref(m), 16, 16, 8, 2, ctypes.byref(p), 16, 16, 16, 16, 16, 4096,
                               None) == _lib.ERR_ARG
    p.num_steps = 8
    need = lib.nglod_trace_min_workspace_size(8)
    assert need >= 8 * 4
    assert lib.nglod_trace_min(ctypes.byref(m), 16, 16, 8, 2, ctypes.byref(p), 16, 16, 16, 16, 16, need - 1,
                               None) == _lib.ERR_WORKSPACE
    # resolutions that do not nest (3 does not divide 4) are refused before any launch
    for i in range(3):
        m.grid_cum[i] = 8192 * (i + 1)
        m.grid_h[i] = 8192 * (i + 1)
    m.w0_bf16[0] = 8192
    m.res[0], m.res[1], m.res[2] = 3, 4, 8
    assert lib.nglod_pack_mixed(ctypes.byref(m), None) == _lib.ERR_UNSUPPORTED
    m.res[0] = 0
    assert lib.nglod_pack_mixed(ctypes.byref(m), None) == _lib.ERR_ARG
