"""ctypes binding of libnglod_hip.so (the C ABI declared in include/nglod_hip.h).

The library is the only compute path of this package: when it is missing the
import of any op raises, there is no CPU fallback.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnglod_hip.so")

MAX_LODS = 8
POOL_COUNTER_BYTES = 4096  # NGLOD_POOL_COUNTER_BYTES

OK = 0
ERR_ARG = -1
ERR_UNSUPPORTED = -2
ERR_CAPACITY = -4
ERR_HIP = -5
ERR_WORKSPACE = -6

FP32 = 0
MIXED = 1
GRAD_METHODS = {"finitediff": 0, "tetrahedron": 1, "autodiff": 2}

_STATUS = {
    ERR_ARG: "bad argument",
    ERR_UNSUPPORTED: "unsupported shape",
    ERR_CAPACITY: "capacity exceeded",
    ERR_HIP: "HIP runtime error",
    ERR_WORKSPACE: "workspace too small",
}

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t


class NglodModel(ctypes.Structure):
    """Mirror of nglod_model_t."""

    _fields_ = [
        ("num_lods", _i32),
        ("feature_dim", _i32),
        ("hidden_dim", _i32),
        ("pos_invariant", _i32),
        ("num_decoders", _i32),
        ("res", _i32 * MAX_LODS),
        ("grid", _p * MAX_LODS),
        ("w0", _p * MAX_LODS),
        ("b0", _p * MAX_LODS),
        ("w1", _p * MAX_LODS),
        ("b1", _p * MAX_LODS),
        ("grid_h", _p * MAX_LODS),
        ("w0_bf16", _p * MAX_LODS),
        ("grid_cum", _p * MAX_LODS),
    ]


class NglodTraceParams(ctypes.Structure):
    """Mirror of nglod_trace_params_t."""

    _fields_ = [
        ("num_steps", _i32),
        ("min_dis", ctypes.c_double),
        ("step_size", ctypes.c_double),
        ("far_clip", ctypes.c_double),
        ("normals", _i32),
        ("refine_steps", _i32),
        ("return_dist", ctypes.c_double),
        ("near_dist", ctypes.c_double),
        ("pass_ms", _p),
    ]


class NglodPoolRequest(ctypes.Structure):
    """Mirror of nglod_pool_request_t."""

    _fields_ = [("model", _i32), ("lod", _i32), ("offset", _i64), ("count", _i64)]


class NglodModelGrads(ctypes.Structure):
    """Mirror of nglod_model_grads_t."""

    _fields_ = [(name, _p * MAX_LODS) for name in ("grid", "w0", "b0", "w1", "b1")]


class NglodSol(ctypes.Structure):
    """Mirror of nglod_sol_t."""

    _fields_ = [
        ("target", _i32),
        ("lod", _i32),
        ("din", _i32),
        ("reserved", _i32),
        ("psum_target", _i64),
        ("points", _p),
        ("trinket_v", _p),
        ("parent", _p),
        ("cc", _p),
        ("cf", _p),
        ("w0", _p),
        ("b0", _p),
        ("w1", _p),
        ("b1", _p),
    ]


# name -> (restype, argtypes); every symbol include/nglod_hip.h declares.
SIGNATURES = {
    "nglod_abi_version": (ctypes.c_int, []),
    "nglod_last_error": (ctypes.c_char_p, []),
    "nglod_aabb": (ctypes.c_int, [_p, _p, _i64, _p, _p, _p, _p]),
    "nglod_grid_sample": (ctypes.c_int, [_p, _i32, _i32, _p, _i64, _p, _p]),
    "nglod_sdf_fwd": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _i64, _i32, _i32, _p, _p]),
    "nglod_sdf_fwd_all": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _i64, _i32, _p, _p]),
    "nglod_sdf_fwd_ordered_workspace_size": (_sz, [_i64]),
    "nglod_sdf_fwd_ordered": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _i64, _i32, _i32, _p, _p, _sz, _p]),
    "nglod_trace_dense_workspace_size": (_sz, [_i64, _i32]),
    "nglod_trace_dense": (
        ctypes.c_int,
        [ctypes.POINTER(NglodModel), _p, _p, _i64, _i32, ctypes.POINTER(NglodTraceParams), _i32,
         _p, _p, _p, _p, _p, _sz, _p, _p, _p],
    ),
    "nglod_trace_dense_noalloc": (
        ctypes.c_int,
        [ctypes.POINTER(NglodModel), _p, _p, _i64, _i32, ctypes.POINTER(NglodTraceParams), _i32,
         _p, _p, _p, _p, _p, _p, _p],
    ),
    "nglod_normals_fd": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _i64, _i32, _i32, _p, _p]),
    "nglod_pack_mixed": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p]),
    "nglod_gradient": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _i64, _i32, _i32, _i32, _p, _i32, _p, _p]),
    "nglod_sdf_grad": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _i64, _i32, _i32, _p, _p, _p]),
    "nglod_trace_min_workspace_size": (_sz, [_i64]),
    "nglod_trace_min": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _p, _i64, _i32, ctypes.POINTER(NglodTraceParams),
                                       _p, _p, _p, _p, _p, _sz, _p]),
    "nglod_sdf_bwd_workspace_size": (_sz, [ctypes.POINTER(NglodModel), _i64, _i32]),
    "nglod_sdf_bwd": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _i64, _i32, _p, _i32,
                                      ctypes.POINTER(NglodModelGrads), _p, _sz, _p]),
    "nglod_spc_occupancy": (ctypes.c_int, [_p, _i64, _i32, _p, _i64, ctypes.POINTER(_i64), _p, _sz,
                                            ctypes.POINTER(_sz), _p]),
    "nglod_spc_from_codes": (ctypes.c_int, [_p, _i64, _i32, _p, _i64, ctypes.POINTER(_i64), _p, _p, _sz,
                                             ctypes.POINTER(_sz), _p]),
    "nglod_spc_points": (ctypes.c_int, [_p, _i64, _p, _i64, _p, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i64), _p,
                                         _p, _sz, ctypes.POINTER(_sz), _p]),
    "nglod_spc_raytrace": (ctypes.c_int, [_p, _p, _p, _p, _i32, _p, _p, _i64, _i32, _p, _i64, ctypes.POINTER(_i64),
                                           _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "nglod_mark_unique_rays": (ctypes.c_int, [_p, _i64, _p, _p]),
    "nglod_mesh2sdf_workspace_size": (_sz, [_i64]),
    "nglod_mesh2sdf_f64": (ctypes.c_int, [_p, _i64, _p, _i64, _p, _p, _p, _sz, _p]),
    "nglod_sample_uniform": (ctypes.c_int, [_i64, ctypes.c_uint64, _p, _p]),
    "nglod_sample_surface": (ctypes.c_int, [_p, _p, _i64, _i64, ctypes.c_uint64, ctypes.c_double, _p, _p, _p]),
    "nglod_sample_directions": (ctypes.c_int, [_i64, ctypes.c_uint64, _p, _p]),
    "nglod_camera_rays": (ctypes.c_int, [_p, _p, ctypes.c_float, _i32, _i32, _i32, _p, _p, _p]),
    "nglod_sol_corners": (ctypes.c_int, [_p, _p, _i32, _i32, _p, _i64, _p, _p, _sz, ctypes.POINTER(_sz), _p]),
    "nglod_sol_features": (ctypes.c_int, [ctypes.POINTER(NglodModel), _i32, _p, _p, _p, _p]),
    "nglod_sol_trinkets": (ctypes.c_int, [_p, _p, _i32, _i32, _p, _p, _p, _p, _p, _sz, _p]),
    "nglod_sol_sdf": (ctypes.c_int, [ctypes.POINTER(NglodSol), _p, _p, _i64, _p, _p]),
    "nglod_trace_sparse_workspace_size": (_sz, [_i64]),
    "nglod_trace_sparse": (ctypes.c_int, [ctypes.POINTER(NglodSol), _p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p,
                                          _sz, _p, _p]),
    "nglod_spc_codes_from_points": (ctypes.c_int, [_p, _p, _i64, _i32, _p, _i64, ctypes.POINTER(_i64), _p, _sz,
                                                   ctypes.POINTER(_sz), _p]),
    "nglod_shadow_rays": (ctypes.c_int, [_p, _p, _i64, ctypes.c_float, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "nglod_ao_workspace_size": (_sz, []),
    "nglod_ao": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _p, _p, _p, _i64, _i32, _i32, _p, _p, _sz, _p]),
    "nglod_shadow_map": (ctypes.c_int, [_p, _p, _i32, _i32, _p, _p, _p]),
    "nglod_matcap_shade": (ctypes.c_int, [_p, _p, _p, _i64, _p, _p, _i32, _i32, _i32, _p, _p, _p, _p]),
    "nglod_vl1d_probe": (ctypes.c_int, [_p, _i64, _i32, _i32, _p, _i64, _p]),
    "nglod_model_table_bytes": (_sz, [_i32]),
    "nglod_model_table_fill": (ctypes.c_int, [ctypes.POINTER(NglodModel), _i32, _p, _sz, _p]),
    "nglod_trace_pool_workspace_size": (_sz, [_i64]),
    "nglod_trace_pool": (ctypes.c_int, [ctypes.POINTER(NglodModel), _p, _i32, ctypes.POINTER(NglodPoolRequest), _i32,
                                        ctypes.POINTER(NglodTraceParams), _p, _p, _p, _p, _p, _p, _i64, _p, _sz, _p,
                                        _p, _p]),
    "nglod_spc_raytrace_async_workspace_size": (_sz, [_i64]),
    "nglod_spc_raytrace_async": (ctypes.c_int, [_p, _p, _p, _p, _i32, _p, _p, _i64, _i32, _p, _i64, _p, _p, _p, _sz,
                                                _p]),
}

_lib = None


def load():
    """Loads libnglod_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m nglod_amd.build` "
            "(the HIP library is the only compute path)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NglodError(RuntimeError):
    pass


def check(rc, what):
    if rc != OK:
        msg = load().nglod_last_error().decode(errors="replace")
        raise NglodError(f"{what} failed ({rc}: {_STATUS.get(rc, 'error')}): {msg}")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def stream_of(t):
    """Raw hipStream_t of torch's current stream on t's device."""
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def require_gpu(*tensors):
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise ValueError("nglod_amd ops run on the GPU: pass tensors on a cuda (HIP) device")
