"""ctypes binding of ``libneuralangelo_amd.so`` (the C ABI in include/neuralangelo_amd.h).

The product path has no fallback: if the library is missing or a call fails,
a ``RuntimeError`` carrying the library's own message is raised.
"""
import ctypes
import os

import torch

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libneuralangelo_amd.so")

NA_OK = 0
NA_ERR_ARGUMENT = -1
NA_ERR_HIP = -2
NA_ERR_UNSUPPORTED = -3

NA_F32 = 0
NA_BF16 = 1
NA_F64 = 3
NA_MAX_LEVELS = 32

NA_ACT_NONE = 0
NA_ACT_RELU = 1
NA_ACT_SOFTPLUS = 2
NA_ACT_SIGMOID = 3
NA_EPI_STORE = 0
NA_EPI_ATOMIC = 1

_DTYPE_CODE = {torch.float32: NA_F32, torch.bfloat16: NA_BF16, torch.float64: NA_F64}


class GridDesc(ctypes.Structure):
    _fields_ = [
        ("n_levels", ctypes.c_int32),
        ("n_features", ctypes.c_int32),
        ("log2_hashmap_size", ctypes.c_int32),
        ("base_resolution", ctypes.c_int32),
        ("per_level_scale", ctypes.c_float),
        ("log2_per_level_scale", ctypes.c_float),
        ("scale", ctypes.c_float * NA_MAX_LEVELS),
        ("resolution", ctypes.c_uint32 * NA_MAX_LEVELS),
        ("offset", ctypes.c_uint32 * (NA_MAX_LEVELS + 1)),
    ]


class GemmArgs(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_int64), ("N", ctypes.c_int64), ("K", ctypes.c_int64),
        ("A", ctypes.c_void_p), ("lda", ctypes.c_int64), ("trans_a", ctypes.c_int32),
        ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("trans_b", ctypes.c_int32),
        ("in_dtype", ctypes.c_int32), ("epilogue", ctypes.c_int32),
        ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64), ("c_dtype", ctypes.c_int32),
        ("bias", ctypes.c_void_p), ("act", ctypes.c_int32), ("act_beta", ctypes.c_float),
        ("dact_y", ctypes.c_void_p), ("ld_dact", ctypes.c_int64), ("dact_dtype", ctypes.c_int32),
        ("dact", ctypes.c_int32),
        ("r1_a", ctypes.c_void_p), ("r1_b", ctypes.c_void_p), ("colsum", ctypes.c_void_p),
        ("split_k", ctypes.c_int32),
        ("z_out", ctypes.c_void_p), ("ld_z", ctypes.c_int64),
        ("zc", ctypes.c_void_p), ("ld_zc", ctypes.c_int64), ("zc_rows", ctypes.c_int64),
        ("head_w", ctypes.c_void_p), ("head_out", ctypes.c_void_p),
        ("wg_partials", ctypes.c_void_p), ("wg_partials_len", ctypes.c_int64),
    ]


NA_ADAMW_MULTI_MAX = 32
NA_WGRAD_STATIC_MAX_FLOATS = 20971520


class AdamwTensor(ctypes.Structure):
    _fields_ = [("param", ctypes.c_void_p), ("grad", ctypes.c_void_p),
                ("exp_avg", ctypes.c_void_p), ("exp_avg_sq", ctypes.c_void_p),
                ("n", ctypes.c_int64)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_F32 = ctypes.c_float
_F64 = ctypes.c_double
_PGRID = ctypes.POINTER(GridDesc)
_PGEMM = ctypes.POINTER(GemmArgs)

# name -> (restype, argtypes); mirrors include/neuralangelo_amd.h one to one
_SIGNATURES = {
    "na_last_error": (ctypes.c_char_p, []),
    "na_version": (ctypes.c_int, []),
    "na_tune_set": (ctypes.c_int64, [ctypes.c_char_p, _I64]),
    "na_tune_get": (ctypes.c_int64, [ctypes.c_char_p]),
    "na_tune_count": (ctypes.c_int, []),
    "na_tune_name": (ctypes.c_char_p, [_I32]),
    "na_device_synchronize": (ctypes.c_int, []),
    "na_hashgrid_layout": (_I64, [_I32, _I32, _I32, _I32, _F32, _PGRID]),
    "na_hashgrid_corners": (ctypes.c_int, [_PGRID, _P, _I64, _P, _P, _P]),
    "na_hashgrid_fwd": (ctypes.c_int, [_PGRID, _P, _I64, _P, _I32, _P, _I32, _I32, _P]),
    "na_hashgrid_bwd": (ctypes.c_int, [_PGRID, _P, _I64, _P, _I32, _P, _I32, _P]),
    "na_gemm": (ctypes.c_int, [_PGEMM, _P]),
    "na_gemm_wgrad_workspace": (ctypes.c_int64, [_PGEMM]),
    "na_gemm_wgrad_static": (ctypes.c_int, [_PGEMM, _P]),
    "na_taps_increment_fwd": (ctypes.c_int, [_P, _I64, _I64, _I64, _P, _I64, _I64, _P, _I64, _P,
                                             _F32, _P, _P, _P]),
    "na_taps_increment_bwd": (ctypes.c_int, [_P, _I64, _I64, _I64, _P, _I64, _I64, _P, _I64, _P,
                                             _F32, _P, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P,
                                             _I32, _P]),
    "na_colsum": (ctypes.c_int, [_P, _I64, _I32, _I64, _I64, _P, _P]),
    "na_head_sum": (ctypes.c_int, [_P, _I32, _I64, _P, _P, _P]),
    "na_linear_narrow_fwd": (ctypes.c_int, [_P, _I64, _I32, _I64, _I64, _P, _P, _I32, _I32, _P, _P]),
    "na_linear_narrow_bwd": (ctypes.c_int, [_P, _I64, _I32, _I64, _I64, _P, _I32, _P, _P, _P, _P, _I64, _I32, _I32, _P, _I64, _I32, _I32, _F32, _P, _P]),
    "na_act_grad": (ctypes.c_int, [_P, _I64, _I32, _P, _I64, _I32, _I64, _I64, _I32, _F32, _P, _I64, _I32, _P, _P]),
    "na_weight_norm_fwd": (ctypes.c_int, [_P, _P, _I64, _I64, _P, _P, _I32, _I64, _P]),
    "na_weight_norm_bwd": (ctypes.c_int, [_P, _P, _P, _I64, _I64, _P, _P, _P]),
    "na_sdf_input_fwd": (ctypes.c_int, [_PGRID, _P, _I64, _F32, _F32, _P, _I32, _I32, _P, _I32, _I64, _P]),
    "na_sdf_input_bwd": (ctypes.c_int, [_PGRID, _P, _I64, _F32, _F32, _P, _I32, _I64, _P, _I32, _P]),
    "na_sdf_taps_bwd": (ctypes.c_int, [_PGRID, _P, _I64, _F32, _F32, _F32, _P, _P, _P, _I32, _I64, _P, _I32, _I32, _P]),
    "na_hashgrid_sorted_workspace": (ctypes.c_int64, [_PGRID, _I64, _I32]),
    "na_hashgrid_bwd_sorted": (ctypes.c_int, [_PGRID, _P, _I64, _P, _I32, _P, _I32, _P, _I64, _P]),
    "na_sdf_input_bwd_sorted": (ctypes.c_int, [_PGRID, _P, _I64, _F32, _F32, _P, _I32, _I64, _P, _I32, _P, _I64, _P]),
    "na_sdf_taps_bwd_sorted": (ctypes.c_int, [_PGRID, _P, _I64, _F32, _F32, _F32, _P, _P, _P, _I32, _I64, _P, _I32, _I32, _P, _I64, _P]),
    "na_sdf_taps_increment_input": (ctypes.c_int, [_PGRID, _P, _I64, _F32, _F32, _F32, _P, _I32, _I32, _P, _I64, _P, _I64, _P, _P]),
    "na_ray_bounds_sample": (ctypes.c_int, [_P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P]),
    "na_sphere_crossings": (ctypes.c_int, [_P, _P, _I64, _F32, _P, _P, _P]),
    "na_camera_rays": (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _I64, _P, _P, _P, _P]),
    "na_sample_dists": (ctypes.c_int, [_P, _P, _F32, _F32, _P, _I64, _I32, _P, _P]),
    "na_sh_encode": (ctypes.c_int, [_P, _I64, _I32, _P, _P]),
    "na_sample_pdf": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P]),
    "na_hier_resample": (ctypes.c_int, [_P, _P, _I64, _I32, _F32, _I32, _P, _P, _P, _P, _P]),
    "na_scatter_merge": (ctypes.c_int, [_P, _P, _I32, _P, _P, _I32, _I64, _P, _P]),
    "na_neus_alpha_fwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _F32, _P, _F32, _F32, _I64, _I32, _P, _P]),
    "na_neus_alpha_bwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _F32, _P, _F32, _F32, _I64, _I32, _P, _P, _P, _P, _P]),
    "na_composite_fwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P, _P]),
    "na_composite_bwd": (ctypes.c_int, [_P, _P, _P, _I64, _I32, _I32, _P, _P, _P, _P, _P]),
    "na_weighted_sum_fwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P]),
    "na_weighted_sum_bwd": (ctypes.c_int, [_P, _P, _I64, _I32, _I32, _P, _P, _P, _P]),
    "na_packed_ray_index": (ctypes.c_int, [_P, _I64, _P, _P]),
    "na_composite_packed_fwd": (ctypes.c_int, [_P, _P, _P, _I64, _I32, _P, _P, _P]),
    "na_composite_packed_bwd": (ctypes.c_int, [_P, _P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P]),
    "na_weighted_sum_packed_fwd": (ctypes.c_int, [_P, _P, _P, _I64, _I32, _P, _P]),
    "na_weighted_sum_packed_bwd": (ctypes.c_int, [_P, _P, _P, _I64, _I32, _P, _P, _P, _P]),
    "na_neus_alpha_packed_fwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _F32, _P, _F32, _F32, _P, _P, _I64, _P, _P]),
    "na_neus_alpha_packed_bwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _F32, _P, _F32, _F32, _P, _P, _I64, _P, _P, _P, _P, _P]),
    "na_neus_composite_packed_fwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _F32, _P, _F32, _F32,
                                                    _P, _I64, _P, _I32, _P, _P, _P, _P, _P]),
    "na_neus_composite_packed_bwd": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _F32, _P, _F32, _F32,
                                                    _P, _I64, _P, _I32, _P, _P, _P, _P, _P, _P,
                                                    _P, _P, _P, _P, _P]),
    "na_bg_encode": (ctypes.c_int, [_P, _I64, _I32, _P, _I32, _I64, _P]),
    "na_bg_alpha_fwd": (ctypes.c_int, [_P, _I64, _P, _I64, _I32, _P, _P, _P]),
    "na_bg_alpha_bwd": (ctypes.c_int, [_P, _I64, _P, _I64, _I32, _P, _P, _P, _I64, _P]),
    "na_taps_points": (ctypes.c_int, [_P, _I64, _F32, _P, _P]),
    "na_taps_combine_fwd": (ctypes.c_int, [_P, _P, _I64, _F32, _F32, _P, _P, _P]),
    "na_taps_combine_bwd": (ctypes.c_int, [_I64, _F32, _F32, _P, _P, _P, _P, _I32, _P]),
    "na_taps_delta_fwd": (ctypes.c_int, [_P, _P, _P, _I64, _F32, _F32, _P, _P, _P]),
    "na_taps_delta_bwd": (ctypes.c_int, [_I64, _F32, _F32, _P, _P, _P, _P, _P, _P]),
    "na_rgb_input_fwd": (ctypes.c_int, [_P, _P, _I64, _P, _I32, _I64, _I32, _P, _I32, _I64, _P]),
    "na_rgb_input_bwd": (ctypes.c_int, [_P, _I32, _I64, _I32, _P, _I32, _I64, _P, _I32, _P]),
    "na_sdf_losses_fwd": (ctypes.c_int, [_P, _P, _P, _I64, _I64, _P, _P]),
    "na_sdf_losses_bwd": (ctypes.c_int, [_P, _P, _P, _I64, _I64, _P, _P, _I32, _P, _P]),
    "na_adamw_step": (ctypes.c_int, [_P, _P, _P, _P, _I64, _F32, _F32, _F32, _F32, _F32, _I64, _P, _I32, _I32, _P]),
    "na_adamw_step_multi": (ctypes.c_int, [ctypes.POINTER(AdamwTensor), _I32, _F32, _F32, _F32,
                                            _F32, _F32, _I64, _P]),
    "na_mc_workspace_size": (_I64, [_I64, _I64, _I64, _I64]),
    "na_mc_count": (ctypes.c_int, [_P, _I64, _I64, _I64, _I64, _F64, _P, _I64, _P, _P]),
    "na_mc_emit": (ctypes.c_int, [_P, _I64, _I64, _I64, _I64, _F64, _P, _I64, _P, _P, _P, _F64, _P, _P, _P]),
    "na_mc_num_blocks": (_I64, [_I64, _I64, _I64, _I64]),
    "na_mc_block_offsets": (ctypes.c_int, [_I64, _I64, _I64, _I64, _P, _I64, _P, _P, _P]),
}

_lib = None


def exported_symbols():
    """Names the Python side binds (every function declared in the public header)."""
    return list(_SIGNATURES)


def load():
    """Load the shared library (once).  Raises RuntimeError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"neuralangelo_amd: HIP library not built ({LIB_PATH} missing); "
            "run `python -m neuralangelo_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    # launch shapes chosen by the autotuner for the architecture the library targets
    from .autotune import apply_cache
    apply_cache(lib)
    _lib = lib
    return lib


def last_error():
    return load().na_last_error().decode(errors="replace")


def check(status, what):
    if status < 0:
        raise RuntimeError(f"{what} failed (status {status}): {last_error()}")
    return status


def call(name, *args):
    """Call a C-ABI entry point and raise on a negative status."""
    fn = getattr(load(), name)
    return check(fn(*args), name)


def dtype_code(dtype):
    try:
        return _DTYPE_CODE[dtype]
    except KeyError:
        raise RuntimeError(f"neuralangelo_amd: unsupported dtype {dtype}") from None


class _TensorPtr(ctypes.c_void_p):
    """A c_void_p that keeps its tensor alive while the argument list exists.

    ``call(..., ptr(x.contiguous()), ...)`` would otherwise drop the temporary
    before the kernel is enqueued, letting the caching allocator hand its block
    to a later argument.  Holding the reference until the C call returns is
    enough: allocations made after the launch are stream-ordered behind it."""


def ptr(t):
    """Raw device pointer of a tensor (None -> NULL); keeps the tensor alive."""
    if t is None:
        return None
    p = _TensorPtr(t.data_ptr())
    p.tensor = t
    return p


def stream_of(t=None):
    """hipStream_t of torch's current stream on the tensor's device."""
    dev = t.device if t is not None else torch.device("cuda", torch.cuda.current_device())
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def require_cuda(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise RuntimeError("neuralangelo_amd: tensors must live on the GPU "
                               f"(got device {t.device}); the HIP path has no CPU fallback")
