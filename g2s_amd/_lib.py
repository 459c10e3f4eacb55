"""ctypes bindings of libqsmi.so (include/qsmi.h).

The HIP library is the only compute path: loading fails loudly when the
library was not built (run ``python -m g2s_amd.build``), and there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqsmi.so")

_LIB = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
vp = C.c_void_p

# every symbol include/qsmi.h declares: name -> (restype, argtypes)
SIGNATURES = {
    "qsmi_last_error": (C.c_char_p, []),
    "qsmi_get_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "qsmi_fft_size": (C.c_uint, [C.c_uint]),
    "qsmi_ti_create": (C.c_int, [C.c_int, C.c_int, u32p, vp, C.c_int, f32p, C.POINTER(vp)]),
    "qsmi_subimages_from_reflected": (C.c_int, [C.c_int, u32p, u32p, C.c_int, vp, f32p]),
    "qsmi_ti_create_from_reflected": (C.c_int, [C.c_int, C.c_int, u32p, u32p, C.c_int, vp, C.POINTER(vp)]),
    "qsmi_ti_destroy": (C.c_int, [vp]),
    "qsmi_ti_set_tuning": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "qsmi_ti_get_tuning": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_int)]),
    "qsmi_set_default_tuning": (C.c_int, [C.c_char_p, C.c_int]),
    "qsmi_device_create": (C.c_int, [vp, f32p, C.c_int, C.POINTER(vp)]),
    "qsmi_device_destroy": (C.c_int, [vp]),
    "qsmi_set_true_mismatch": (C.c_int, [vp, C.c_int]),
    "qsmi_set_tie_quantum": (C.c_int, [vp, C.c_float]),
    "qsmi_candidate_for_pattern": (C.c_int, [vp, C.c_int, i32p, f32p, f32p, C.POINTER(C.c_int)]),
    "qsmi_search_k_biggest": (C.c_int, [vp, f32p, u32p, C.c_uint, C.c_float]),
    "qsmi_search_k_biggest_noalloc": (C.c_int, [vp, f32p, u32p, C.c_uint, C.c_float]),
    "qsmi_cvt_index_to_position": (C.c_int, [vp, C.c_uint, C.POINTER(C.c_uint)]),
    "qsmi_cvt_position_to_index": (C.c_int, [vp, C.c_uint, C.POINTER(C.c_uint)]),
    "qsmi_get_array_size": (C.c_int, [vp, C.POINTER(C.c_uint)]),
    "qsmi_get_error_map": (C.c_int, [vp, f32p]),
    "qsmi_set_value_in_error_array": (C.c_int, [vp, C.c_uint, C.c_float]),
    "qsmi_device_set_cross_layer": (C.c_int, [vp, f32p]),
    "qsmi_mask_layer_with_variable": (C.c_int, [vp, C.c_uint, C.c_uint]),
    "qsmi_compensate_missing_data": (C.c_int, [vp]),
    "qsmi_get_cross_map": (C.c_int, [vp, f32p]),
    "qsmi_scan_wave": (C.c_int, [vp, C.c_int, i32p, i32p, f32p, f32p, f32p, C.c_float, f32p, u32p, vp]),
    "qsmi_distribution_wave": (C.c_int, [vp, C.c_int, i32p, i32p, f32p, f32p, f32p, C.c_float, u32p, f32p, f32p,
                                         u32p]),
    "qsmi_set_value_in_error_array_with_radius": (C.c_int, [vp, C.c_uint, C.c_float, C.c_float]),
    "qsmi_calibration": (C.c_int, [vp, vp]),
    "qsmi_wave_create": (C.c_int, [vp, C.c_int, i32p, i32p, f32p, f32p, f32p, C.c_float, C.POINTER(vp)]),
    "qsmi_wave_run": (C.c_int, [vp]),
    "qsmi_wave_read": (C.c_int, [vp, vp, vp, vp]),
    "qsmi_wave_destroy": (C.c_int, [vp]),
    "qsmi_synchronize": (C.c_int, [vp]),
    "qsmi_last_stage_ms": (C.c_int, [vp, f32p]),
    "qsmi_set_stage_timing": (C.c_int, [vp, C.c_int]),
    "qsmi_categories": (C.c_int, [C.c_size_t, C.c_int, f32p, C.c_int, f32p, C.c_int, C.POINTER(C.c_int)]),
    "qsmi_coef_matrix": (C.c_int, [C.c_int, i32p, i32p, f32p, i32p, i32p, C.POINTER(C.c_int)]),
    "qsmi_ti_subimages": (C.c_int, [C.c_int, u32p, C.c_int, i32p, f32p, i32p, f32p, f32p]),
    "qsmi_offsets": (C.c_int, [C.c_int, u32p, C.c_int, vp, i32p, C.c_int, C.POINTER(C.c_int)]),
    "qsmi_kernel_generate": (C.c_int, [C.c_int, u32p, C.c_int, i32p, f32p, f32p, f32p]),
    "qsmi_path": (C.c_int, [C.c_size_t, u8p, C.c_uint, u32p, f32p, C.POINTER(C.c_uint)]),
    "qsmi_qs_run_in_memory": (C.c_int, [vp, vp]),
    "qsmi_lp_create": (C.c_int, [C.c_int, C.c_uint, C.POINTER(vp)]),
    "qsmi_lp_destroy": (C.c_int, [vp]),
    "qsmi_lp_param_count": (C.c_int, [C.POINTER(C.c_longlong)]),
    "qsmi_lp_get_params": (C.c_int, [vp, f32p]),
    "qsmi_lp_set_params": (C.c_int, [vp, f32p]),
    "qsmi_lp_gemm": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, f32p, f32p, f32p]),
    "qsmi_lp_encode": (C.c_int, [vp, C.c_int, C.c_int, f32p, f32p]),
    "qsmi_lp_loss_grad": (C.c_int, [vp, C.c_int, C.c_int, f32p, f32p, C.c_float, C.POINTER(C.c_float), vp]),
    "qsmi_lp_train": (C.c_int, [vp, vp, vp, vp]),
    "qsmi_lp_build_keys": (C.c_int, [vp, u32p, f32p, i32p]),
    "qsmi_coef_matrix_x": (C.c_int, [C.c_int, i32p, f32p, f32p, i32p, i32p, C.POINTER(C.c_int)]),
    "qsmi_ti_subimages_x": (C.c_int, [C.c_int, u32p, C.c_int, i32p, f32p, f32p]),
    "qsmi_bgrid_bytes": (C.c_int, [C.c_int, u32p, C.c_int, C.POINTER(C.c_size_t)]),
    "qsmi_bgrid_serialize": (C.c_int, [C.c_int, u32p, C.c_int, i32p, C.c_int, f32p, vp, C.c_size_t]),
    "qsmi_bgrid_valid": (C.c_int, [vp, C.c_size_t]),
    "qsmi_bgrid_header": (C.c_int, [vp, C.c_size_t, C.POINTER(C.c_int), u32p, C.POINTER(C.c_int), i32p, C.c_int,
                                    C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "qsmi_bgrid_write_file": (C.c_int, [C.c_char_p, vp, C.c_size_t, C.c_int]),
    "qsmi_bgrid_read_file": (C.c_int, [C.c_char_p, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
}


class CalibRequest(C.Structure):
    _fields_ = [
        ("gpu", C.c_int),
        ("ndim", C.c_int),
        ("nTI", C.c_int),
        ("tiSize", C.c_void_p),
        ("ti", C.c_void_p),
        ("nVar", C.c_int),
        ("varTypes", C.c_void_p),
        ("nKernels", C.c_int),
        ("kernelSize", C.c_void_p),
        ("kernels", C.c_void_p),
        ("nConfigs", C.c_int),
        ("nbWidth", C.c_int),
        ("nbList", C.c_void_p),
        ("nDensity", C.c_int),
        ("density", C.c_void_p),
        ("maxK", C.c_int),
        ("power", C.c_float),
        ("maxIter", C.c_uint),
        ("minIter", C.c_uint),
        ("maxSeconds", C.c_double),
        ("noise", C.c_float),
        ("seed", C.c_uint),
        ("waveSize", C.c_int),
        ("logCap", C.c_int),
        ("logSetting", C.c_void_p),
        ("logCell", C.c_void_p),
        ("logPos", C.c_void_p),
        ("stageTiming", C.c_int),
    ]


class CalibResult(C.Structure):
    _fields_ = [
        ("meanError", C.c_void_p),
        ("devError", C.c_void_p),
        ("samples", C.c_void_p),
        ("distributions", C.c_longlong),
        ("iterations", C.c_int),
        ("waves", C.c_int),
        ("ms", C.c_double),
        ("scanMs", C.c_double),
        ("drawMs", C.c_double),
        ("setupMs", C.c_double),
        ("stageMs", C.c_double * 4),
    ]


class QsmiError(RuntimeError):
    pass


def lib():
    """Load libqsmi.so; raises when it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise QsmiError(f"{LIB_PATH} is missing: build it with `python -m g2s_amd.build` "
                        "(the HIP library is the only implementation)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().qsmi_last_error().decode(errors="replace")
        raise QsmiError(f"qsmi error {rc}: {msg}")
    return rc


class QsRequest(C.Structure):
    _fields_ = [
        ("gpu", C.c_int),
        ("ndim", C.c_int),
        ("tiSize", C.c_uint * 3),
        ("nVar", C.c_int),
        ("varTypes", vp),
        ("ti", vp),
        ("simSize", C.c_uint * 3),
        ("di", vp),
        ("kernelSize", C.c_uint * 3),
        ("kernel", vp),
        ("n", C.c_int),
        ("k", C.c_float),
        ("seed", C.c_uint),
        ("waveSize", C.c_int),
        ("tieQuantum", C.c_float),
        ("stageTiming", C.c_int),
        ("recordCap", C.c_int),
        ("recOffsets", vp),
        ("recValues", vp),
        ("recBest", vp),
        ("lp", vp),
        ("lpShortlist", C.c_int),
        ("lpFallbackEvery", C.c_int),
        ("lpFallbackMse", C.c_float),
        ("lpAudit", C.c_int),
        ("lpMinCover", C.c_float),
        ("noVerbatim", C.c_int),
        ("nTI", C.c_int),
        ("tis", C.c_void_p),
        ("nImage", C.c_void_p),
        ("kImage", C.c_void_p),
        ("nKernels", C.c_int),
        ("kernelSizes", C.c_void_p),
        ("kernels", C.c_void_p),
        ("kernelIndexImage", C.c_void_p),
        ("circularSim", C.c_int),
        ("tiIndexImage", C.c_void_p),
        ("circularTI", C.c_int),
        ("tiSizes", C.c_void_p),
        ("progressSeconds", C.c_int),
    ]


class QsResult(C.Structure):
    _fields_ = [
        ("simulation", vp),
        ("index", vp),
        ("setupMs", C.c_double),
        ("pathLoopMs", C.c_double),
        ("scanMs", C.c_double),
        ("stageMs", C.c_double * 4),
        ("nodes", C.c_longlong),
        ("scannedNodes", C.c_longlong),
        ("waves", C.c_int),
        ("phaseMs", C.c_double * 4),
        ("lpProposed", C.c_longlong),
        ("lpFallbacks", C.c_longlong),
        ("lpOverflows", C.c_longlong),
        ("lpShortlistSum", C.c_longlong),
        ("lpAuditNodes", C.c_longlong),
        ("lpAuditHits", C.c_longlong),
        ("lpAuditBestHits", C.c_longlong),
        ("exactNodes", C.c_longlong),
        ("bandPositions", C.c_longlong),
        ("secondPassBlocks", C.c_longlong),
        ("fullRescoreNodes", C.c_longlong),
    ]


class LpTrainSet(C.Structure):
    _fields_ = [
        ("tiSize", C.c_uint * 2),
        ("ti", vp),
        ("canon", vp),
        ("nPairs", C.c_int),
        ("offsets", vp),
        ("values", vp),
        ("best", vp),
    ]


class LpTrainOpts(C.Structure):
    _fields_ = [
        ("steps", C.c_int),
        ("batch", C.c_int),
        ("negatives", C.c_int),
        ("lr", C.c_float),
        ("invTau", C.c_float),
        ("seed", C.c_uint),
    ]
