"""ctypes binding of libtnsgpu.so (C ABI declared in include/tnsgpu.h).

The library is built in-tree (``tnsp_amd/libtnsgpu.so``) by
``tnsp_amd.build``.  Loading it needs no GPU; running plans does.  There is no
fallback: if the library is missing, every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# TNSGPU_LIB selects an alternative build of the same library (e.g. a variant
# compiled with different flags); the default is the in-tree build.
LIB_PATH = os.environ.get("TNSGPU_LIB") or os.path.join(HERE, "libtnsgpu.so")

_lib = None

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_dblp = ctypes.POINTER(ctypes.c_double)
c_vp = ctypes.c_void_p
c_vpp = ctypes.POINTER(ctypes.c_void_p)

# (name, argtypes); every function returns int status unless noted
_SIGNATURES = {
    "tns_device_count": [ctypes.POINTER(ctypes.c_int)],
    "tns_set_svd_max_sweeps": [ctypes.c_int],
    "tns_shape_create": [ctypes.c_int, ctypes.c_int, c_i32p, c_i32p, c_i32p, c_i64p, c_u8p, c_vpp],
    "tns_shape_free": [c_vp],
    "tns_shape_counts": [c_vp, c_i32p, c_i32p, c_i32p, c_i64p, c_i64p],
    "tns_shape_export": [c_vp, c_i32p, c_i32p, c_i32p, c_i64p, c_u8p],
    "tns_shape_blocks": [c_vp, c_i64p, c_i64p],
    "tns_shape_rename": [c_vp, ctypes.c_int, c_i32p, c_i32p, c_vpp],
    "tns_plan_edge_operator": [c_vp, ctypes.c_int, c_i32p, c_i32p, c_i32p, c_vp, ctypes.c_int, c_i32p,
                               ctypes.c_int, c_i32p, c_i32p, c_i32p, ctypes.c_int, c_i32p, ctypes.c_int,
                               ctypes.c_int, c_i32p, ctypes.c_int, c_i32p, ctypes.c_int, c_i32p, ctypes.c_int,
                               c_i32p, c_vpp],
    "tns_plan_contract": [c_vp, c_vp, ctypes.c_int, c_i32p, c_vpp],
    "tns_plan_svd": [c_vp, ctypes.c_int, c_i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                     c_vpp],
    "tns_plan_qr": [c_vp, ctypes.c_char, ctypes.c_int, c_i32p, ctypes.c_int32, ctypes.c_int32, c_vpp],
    "tns_plan_conjugate": [c_vp, ctypes.c_int, c_vpp],
    "tns_plan_free": [c_vp],
    "tns_plan_num_outputs": [c_vp, ctypes.POINTER(ctypes.c_int)],
    "tns_plan_output_shape": [c_vp, ctypes.c_int, c_vpp],
    "tns_plan_workspace": [c_vp, ctypes.c_int, c_i64p],
    "tns_plan_flops": [c_vp, c_i64p],
    "tns_plan_run": [c_vp, ctypes.c_int, c_vpp, c_vpp, c_vp, c_vp],
    "tns_plan_run_noalloc": [c_vp, ctypes.c_int, c_vpp, c_vpp, c_vp],
    "tns_svd_factor": [c_vp, ctypes.c_int, c_vp, c_vp, ctypes.c_int64, ctypes.c_double, c_vp],
    "tns_plan_buffers": [c_vp, c_i64p, ctypes.POINTER(ctypes.c_int)],
    "tns_svd_sectors": [c_vp, c_i64p, c_i64p, c_i32p, ctypes.POINTER(ctypes.c_int)],
    "tns_svd_truncate": [c_vp, c_dblp, ctypes.c_int64, ctypes.c_double],
    "tns_plan_contract_many": [ctypes.c_int, c_vpp, c_vpp, ctypes.c_int, c_i32p, ctypes.c_int, c_vpp],
    "tns_plan_qr_many": [ctypes.c_int, c_vpp, ctypes.c_char, ctypes.c_int, c_i32p, ctypes.c_int32, ctypes.c_int32,
                         ctypes.c_int, c_vpp],
    "tns_plan_svd_many": [ctypes.c_int, c_vpp, ctypes.c_int, c_i32p, ctypes.c_int32, ctypes.c_int32,
                          ctypes.c_int32, ctypes.c_int32, ctypes.c_int, c_vpp],
    "tns_plan_conjugate_many": [ctypes.c_int, c_vpp, ctypes.c_int, ctypes.c_int, c_vpp],
    "tns_plans_run": [ctypes.c_int, c_vpp, ctypes.c_int, c_vpp, c_vpp, c_vpp, c_vp],
    "tns_svd_factor_many": [ctypes.c_int, c_vpp, ctypes.c_int, c_vpp, c_vpp, ctypes.c_int64, ctypes.c_double,
                            ctypes.c_int, c_vp],
    "tns_shapes_free": [ctypes.c_int, c_vpp],
    "tns_plans_free": [ctypes.c_int, c_vpp],
    "tns_shape_rename_many": [ctypes.c_int, c_vpp, ctypes.c_int, c_i32p, c_i32p, c_vpp],
    "tns_plans_output_shapes": [ctypes.c_int, c_vpp, ctypes.c_int, c_vpp, c_i64p],
    "tns_plans_workspace": [ctypes.c_int, c_vpp, ctypes.c_int, c_i64p],
    "tns_engine_create": [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vpp, c_vpp, ctypes.c_int, c_i32p, ctypes.c_int,
                          ctypes.c_int, c_vpp],
    "tns_engine_set_chunk": [c_vp, ctypes.c_int64],
    "tns_engine_set_lanes": [c_vp, ctypes.c_int],
    "tns_engine_set_precision": [c_vp, ctypes.c_int],
    "tns_engine_autotune": [c_vp, ctypes.c_int64, c_i32p, ctypes.c_int64, ctypes.c_double, c_i32p, c_dblp, c_vp],
    "tns_engine_evaluate": [c_vp, ctypes.c_int64, c_i32p, ctypes.c_int64, ctypes.c_double, c_dblp, c_vp],
    "tns_engine_stats": [c_vp, c_dblp, ctypes.c_int],
    "tns_engine_reset_stats": [c_vp],
    "tns_engine_set_profile": [c_vp, ctypes.c_int],
    "tns_engine_refresh_sites": [c_vp],
    "tns_engine_trim": [c_vp],
    "tns_engine_kernel_stats": [c_vp, c_dblp, ctypes.c_int],
    "tns_engine_free": [c_vp],
    "tns_conv3x3_forward": [c_vp, c_vp, c_vp, c_vp] + [ctypes.c_int] * 6 + [c_vp],
    "tns_conv3x3_backward_data": [c_vp, c_vp, c_vp, c_vp] + [ctypes.c_int] * 5 + [c_vp],
    "tns_conv3x3_backward_weight_workspace": [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_i64p],
    "tns_conv3x3_backward_weight": [c_vp] * 5 + [ctypes.c_int] * 5 + [c_vp, c_vp],
    "tns_hop_sample": [c_vp] * 5 + [ctypes.c_int] * 3 + [c_vp],
    "tns_hop_log_prob": [c_vp] * 4 + [ctypes.c_int] * 3 + [c_vp],
    "tns_hop_regression": [c_vp] * 5 + [ctypes.c_int] * 2 + [c_vp],
    "tns_adam_step": [c_vp] * 5 + [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_int, c_vp],
    "tns_to_bf16": [c_vp, c_vp, ctypes.c_int, c_vp],
    "tns_occupations_to_input": [c_vp, c_vp, ctypes.c_int64, c_vp],
    "tns_plan_stage_count": [c_vp, ctypes.c_int, c_i64p],
    "tns_plan_stage_export": [c_vp, ctypes.c_int, c_i64p],
}
_EXTRA = {"tns_last_error": ([], ctypes.c_char_p), "tns_version": ([], ctypes.c_char_p)}


def exported_symbols():
    return list(_SIGNATURES) + list(_EXTRA)


def lib():
    """Load libtnsgpu.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libtnsgpu.so not built ({LIB_PATH}); run tnsp_amd.build.build()")
    # torch ships its own HIP runtime under the same soname as the system
    # one; loading torch first makes the library bind to that copy, so the
    # process holds one runtime whichever module the caller imported first
    import torch  # noqa: F401

    L = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    for name, (args, res) in _EXTRA.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError(lib().tns_last_error().decode())


def arr32(values):
    values = [int(v) for v in values]
    return (ctypes.c_int32 * max(1, len(values)))(*values)


def arr64(values):
    values = [int(v) for v in values]
    return (ctypes.c_int64 * max(1, len(values)))(*values)
