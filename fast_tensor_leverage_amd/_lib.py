"""ctypes binding of libftl_hip.so (declared in include/ftl_hip.h).

The library is loaded from the package directory.  There is no fallback: when
the library is missing or no HIP device is visible the calls raise.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libftl_hip.so")

FTL_OK, FTL_EINVAL, FTL_EHIP, FTL_ENOMEM, FTL_EUNSUPPORTED = 0, 1, 2, 3, 4
FTL_FP64, FTL_FP32 = 0, 1
FTL_SAMPLER_EFFICIENT, FTL_SAMPLER_LARSEN_KOLDA, FTL_SAMPLER_LARSEN_KOLDA_HYBRID = 0, 1, 2

_u32, _u64, _i32, _vp, _dp = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p, ctypes.c_double

# (name, argtypes, restype) of every exported entry point; the test suite
# checks that include/ftl_hip.h declares exactly these.
SIGNATURES = {
    "ftl_last_error": ([], ctypes.c_char_p),
    "ftl_version": ([], ctypes.c_char_p),
    "ftl_device_count": ([], _i32),
    "ftl_eigensolver_capped": ([_vp], _i32),
    "ftl_device_pool_trim": ([], _i32),
    "ftl_device_pool_held": ([_vp], _i32),
    "ftl_device_pool_set_cap": ([_u64], _i32),
    "ftl_device_pool_cap": ([_vp], _i32),
    "ftl_tuning_lookup": ([ctypes.c_char_p, _i32, _u64, _u64, _u64, ctypes.c_char_p, _u64], _i32),
    "ftl_sampler_create": ([_u32, _vp, _u64, _u64, _vp, _i32, _vp], _i32),
    "ftl_sampler_create_device": ([_u32, _vp, _u64, _u64, _vp, _i32, _vp], _i32),
    "ftl_sampler_destroy": ([_vp], _i32),
    "ftl_sampler_update": ([_vp, _u32, _vp], _i32),
    "ftl_sampler_update_device": ([_vp, _u32, _vp], _i32),
    "ftl_sampler_draw": ([_vp, _u32, _vp, _vp], _i32),
    "ftl_sampler_draw_materialize": ([_vp, _u32, _vp, _vp, _vp, _vp], _i32),
    "ftl_sampler_draw_device": ([_vp, _u32, _vp, _vp, _vp, _vp, _vp], _i32),
    "ftl_sampler_set_seed": ([_vp, _u64, _u64], _i32),
    "ftl_sampler_get_tree": ([_vp, _u32, _vp, _u64], _i32),
    "ftl_sampler_compute_m": ([_vp, _u32, _vp, _vp], _i32),
    "ftl_sampler_info": ([_vp, _vp], _i32),
    "ftl_sampler_stream": ([_vp], _vp),
    "ftl_sampler_set_option": ([_vp, ctypes.c_char_p, ctypes.c_int64], _i32),
    "ftl_sampler_get_option": ([_vp, ctypes.c_char_p, _vp], _i32),
    "ftl_sampler_enable_timing": ([_vp, _i32], _i32),
    "ftl_sampler_last_timing": ([_vp, _vp], _i32),
    # sparse tensor and sampled CP-ALS
    "ftl_sparse_create": ([_u32, _u64, _vp, _vp, _vp, _vp], _i32),
    "ftl_sparse_create_device": ([_u32, _u64, _vp, _vp, _vp, _vp], _i32),
    "ftl_sparse_destroy": ([_vp], _i32),
    "ftl_sparse_set_lookup": ([_vp, _i32], _i32),
    "ftl_sparse_info": ([_vp, _vp], _i32),
    "ftl_sparse_normsq": ([_vp, _vp], _i32),
    "ftl_sparse_mttkrp": ([_vp, _u32, _u64, _vp, _vp, _u64, _vp, _u64, _vp], _i32),
    "ftl_sparse_lookup": ([_vp, _u32, _u64, _vp, _vp, _vp, _u64, _vp], _i32),
    "ftl_sparse_residual_normsq": ([_vp, _u64, _vp, _vp, _vp, _vp], _i32),
    "ftl_sparse_residual_normsq_noalloc": ([_vp, _u64, _vp, _vp, _vp, _vp, _vp], _i32),
    "ftl_als_create": ([_vp, _u32, _vp, _u64, _vp, _vp], _i32),
    "ftl_als_destroy": ([_vp], _i32),
    "ftl_als_initialize_ds": ([_vp, _u64, _i32], _i32),
    "ftl_als_set_seed": ([_vp, _u64, _u64], _i32),
    "ftl_als_ds_update": ([_vp, _u32, _i32, _i32, _vp], _i32),
    "ftl_als_ds_update_device": ([_vp, _u32, _i32, _i32, _vp], _i32),
    "ftl_als_exact_update": ([_vp, _u32, _i32, _i32], _i32),
    "ftl_als_initialize_ds_sampler": ([_vp, _u64, _i32, _i32], _i32),
    "ftl_lk_sampler_create": ([_u32, _vp, _u64, _u64, _vp, _vp], _i32),
    "ftl_lk_sampler_destroy": ([_vp], _i32),
    "ftl_lkh_sampler_create": ([_u32, _vp, _u64, _u64, _vp, _vp], _i32),
    "ftl_lk_sampler_info": ([_vp, _vp], _i32),
    "ftl_lk_sampler_draw": ([_vp, _u32, _vp, _vp, _vp, _vp], _i32),
    "ftl_lk_sampler_draw_device": ([_vp, _u32, _vp, _vp, _vp, _vp, _vp], _i32),
    "ftl_lk_sampler_update": ([_vp, _u32, _vp], _i32),
    "ftl_lk_sampler_factor_device": ([_vp, _u32, _vp], _i32),
    "ftl_lk_sampler_stream": ([_vp], _vp),
    "ftl_lk_sampler_set_seed": ([_vp, _u64, _u64], _i32),
    "ftl_lk_sampler_leverage": ([_vp, _u32, _vp, _vp], _i32),
    "ftl_als_get_factor": ([_vp, _u32, _vp], _i32),
    "ftl_als_set_factor": ([_vp, _u32, _vp], _i32),
    "ftl_als_update_sampler": ([_vp, _u32], _i32),
    "ftl_als_get_sigma": ([_vp, _vp], _i32),
    "ftl_als_set_sigma": ([_vp, _vp], _i32),
    "ftl_als_renormalize_columns": ([_vp, _i32], _i32),
    "ftl_als_residual_normsq": ([_vp, _vp], _i32),
    "ftl_als_lowrank_normsq": ([_vp, _vp], _i32),
    "ftl_als_factor_device": ([_vp, _u32, _vp], _i32),
    "ftl_als_last_stats": ([_vp, _vp], _i32),
    "ftl_als_last_draw": ([_vp, _vp, _vp, _vp], _i32),
    "ftl_als_enable_timing": ([_vp, _i32], _i32),
    "ftl_als_stream": ([_vp], _vp),
    # draw-request server
    "ftl_server_create": ([_u64, _u32, _vp], _i32),
    "ftl_server_destroy": ([_vp], _i32),
    "ftl_server_submit": ([_vp, _u32, _vp, _u64, _u64, _u32, _u64, _i32, _vp], _i32),
    "ftl_server_progress": ([_vp, _vp, _u64, _vp], _i32),
    "ftl_server_result": ([_vp, _u64, _vp], _i32),
    "ftl_server_factor": ([_vp, _u64, _u32, _vp], _i32),
    "ftl_server_cancel": ([_vp, _u64], _i32),
    "ftl_server_stats": ([_vp, _vp], _i32),
    "ftl_server_add_factors": ([_vp, _u32, _vp, _u64, _vp, _i32, _i32, _u64, _vp], _i32),
    "ftl_server_add_sampler": ([_vp, _vp, _vp], _i32),
    "ftl_server_update_factor": ([_vp, _u64, _u32, _vp, _i32], _i32),
    "ftl_server_remove_factors": ([_vp, _u64], _i32),
    "ftl_server_submit_on": ([_vp, _u64, _u64, _u32, _u64, _vp], _i32),
    "ftl_server_wait": ([_vp, _u64, ctypes.c_double, _vp], _i32),
}

_lib = None


class FTLError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m fast_tensor_leverage_amd.build`")
    # torch ships a HIP runtime with the same soname as ROCm's; loading torch
    # first makes libftl_hip.so bind to that one, so the process has a single
    # HIP runtime (two of them hide the device from the second)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def last_error() -> str:
    msg = load().ftl_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = ""):
    if rc == FTL_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if rc == FTL_EINVAL:
        raise ValueError(msg)
    if rc == FTL_ENOMEM:
        raise MemoryError(msg)
    if rc == FTL_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise FTLError(msg)


def eigensolver_capped() -> int:
    """Eigensolves on the current device that stopped at the Jacobi sweep cap
    before converging, since the library was loaded (waits for the device)."""
    out = ctypes.c_uint64()
    check(load().ftl_eigensolver_capped(ctypes.byref(out)), "eigensolver_capped")
    return int(out.value)


def release_cached_memory() -> None:
    """Release the device blocks the library keeps from destroyed samplers
    for reuse by later ones (hipFree of each)."""
    check(load().ftl_device_pool_trim(), "release_cached_memory")


def set_cached_memory_limit(nbytes: int) -> None:
    """Most bytes of device blocks kept from destroyed samplers (0: none);
    blocks beyond a lowered limit are released, the oldest first."""
    check(load().ftl_device_pool_set_cap(int(nbytes)), "set_cached_memory_limit")


def cached_memory_limit() -> int:
    out = ctypes.c_uint64()
    check(load().ftl_device_pool_cap(ctypes.byref(out)), "cached_memory_limit")
    return int(out.value)


def cached_memory_bytes() -> int:
    """Bytes of device blocks kept from destroyed samplers."""
    out = ctypes.c_uint64()
    check(load().ftl_device_pool_held(ctypes.byref(out)), "cached_memory_bytes")
    return int(out.value)


def tuning_lookup(precision: int, R: int, depth: int, J: int, table_text: str = None):
    """(table file, row index or -1, {option: value}) of the tuning-table row
    for a shape; ``table_text`` parses that text instead of the loaded
    table.  Host-only."""
    buf = ctypes.create_string_buffer(4096)
    txt = None if table_text is None else table_text.encode()
    check(load().ftl_tuning_lookup(txt, int(precision), int(R), int(depth), int(J), buf, len(buf)), "tuning_lookup")
    path, row, opts = buf.value.decode().split("\t")
    kv = {}
    for item in filter(None, opts.split(",")):
        k, v = item.split("=")
        kv[k] = int(v)
    return path, int(row), kv


def device_count() -> int:
    return int(load().ftl_device_count())


def require_device():
    n = device_count()
    if n < 1:
        # a second opinion and the visibility settings, for the report
        try:
            import torch
            tn = torch.cuda.device_count()
        except Exception as e:  # noqa: BLE001
            tn = f"unavailable ({e})"
        env = {k: os.environ.get(k) for k in ("HIP_VISIBLE_DEVICES", "ROCR_VISIBLE_DEVICES", "CUDA_VISIBLE_DEVICES")}
        raise FTLError(f"no HIP device visible: the MI355X path needs a GPU (hipGetDeviceCount gave {n}; "
                       f"torch.cuda.device_count() = {tn}; {env})")
