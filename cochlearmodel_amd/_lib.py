"""ctypes binding of libcochlea_amd.so (the C-ABI declared in include/cochlea_amd.h).

The library is built in-tree by ``cochlearmodel_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcochlea_amd.so")

SECTIONS = 256
LAMBDA_COUNT = 3


class BMScalars(C.Structure):
    """cm_bm_scalars (cm_types.h)."""
    _fields_ = [(n, C.c_float) for n in (
        "w_ohc", "time_step", "time_step_out", "delta_x", "alpha_r", "Ts", "alpha_l", "alpha_s",
        "Gme", "a0", "a1", "a2", "sigma_ow", "eta1", "eta2", "dt_min", "dt_max")] + [
        ("j_fast", C.c_int32), ("j_slow", C.c_int32), ("max_passes", C.c_int32),
        ("decision_mode", C.c_int32), ("jacobi_mode", C.c_int32), ("check_every", C.c_int32),
        ("max_sweeps", C.c_int32), ("tol_rel", C.c_float), ("warm_start", C.c_int32)]


ITEM_DTYPE = np.dtype([("start", np.int32), ("mosi", np.int32), ("mesi", np.int32),
                       ("write_start", np.int32), ("tol", np.float32), ("m1tol", np.float32),
                       ("set", np.int32), ("pad", np.int32)])
STATS_DTYPE = np.dtype([("steps", np.int64), ("passes", np.int64), ("sweeps", np.int64)])
TUNING_DTYPE = np.dtype([("waves_per_block", np.int32), ("order_lpt", np.int32), ("n_trial_items", np.int32),
                         ("jacobi_unroll", np.int32), ("ms", np.float32, 18), ("sections_per_lane", np.int32),
                         ("ms_split", np.float32, 12)])
TUNING_UNROLLS = (2, 4, 1)   # ms index 6 u + 2 w + o; ms_split index 6 v + 2 u + o
TUNING_SPLITS = (2, 1)       # sections per lane of ms_split's v
IHC_INTERVAL_DTYPE = np.dtype([("row0", np.int64), ("len", np.int32), ("kbase", np.int32), ("L", np.int32),
                               ("tla", np.int32), ("frdp", np.int32), ("set", np.int32)])
assert IHC_INTERVAL_DTYPE.itemsize == 32
PROFILE_DTYPE = np.dtype([("dA", np.float64), ("recipdA", np.float64), ("dBSPLSignal", np.float64),
                          ("dBSPLNoise", np.float64), ("Wn", np.float64), ("isReference", np.int8),
                          ("pad0", np.int8, 3), ("calculateIndexPosition", np.int32),
                          ("referenceIndex", np.int32), ("pad1", np.int32), ("frequency", np.float64)])
assert PROFILE_DTYPE.itemsize == 64

_lib = None

vp = C.c_void_p
i32, i64, f32, f64, sz = C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_size_t

# name -> (restype, argtypes)
_SIGS = {
    "cm_last_error": (C.c_char_p, []),
    "cm_set_stream": (C.c_int, [vp]),
    "cm_device_synchronize": (C.c_int, []),
    "viewGPUStatus": (C.c_int, [C.c_int, C.c_char_p]),
    "cudaEventsCreate": (C.c_int, [vp, vp, C.c_int]),
    "cudaEventsStartTimer": (C.c_int, [vp, vp, C.c_int]),
    "cudaEventsStopTimer": (C.c_int, [vp, vp, C.c_int]),
    "cudaEventsStopQueryTimer": (f32, [vp, vp, C.c_int, C.c_char_p]),
    "cudaEventsDestroy": (C.c_int, [vp, vp]),
    "cm_last_kernel_ms": (C.c_float, []),
    "cm_set_bm_launch": (C.c_int, [C.c_int, C.c_int]),
    "cm_set_bm_jacobi_unroll": (C.c_int, [C.c_int]),
    "cm_set_bm_sections_per_lane": (C.c_int, [C.c_int]),
    "cm_set_bm_order": (C.c_int, [vp, C.c_int]),
    "cm_set_bm_solver": (C.c_int, [C.c_int, f32, C.c_int, C.c_int, C.c_int]),
    "cm_bm_autotune": (C.c_int, [vp, vp, i64, vp, vp, vp, C.c_int, vp, i64, C.c_int, vp]),
    "BMOHCKernel_Init": (C.c_int, [vp] * 11 + [f32, f32, sz, sz, sz, C.c_int, C.c_int, C.c_int]),
    "BMOHCNewKernel": (C.c_int, [vp, C.c_int, f32, f32, f32, f32, f32, C.c_int, C.c_int, C.c_int, f32, f32, f32,
                                 f32, f32, f32, f32, f32, f32, f32, C.c_int, C.c_int, C.c_int, C.c_long, C.c_int,
                                 f32, f32, C.c_int, f32, f32, C.c_int, C.c_int, f32, f32, f32, C.c_int, C.c_int,
                                 C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cm_bm_run": (C.c_int, [vp, vp, i64, vp, vp, vp, C.c_int, vp, i64] + [vp] * 9 + [C.c_int, C.c_int, C.c_int]),
    "cm_get_last_blocks": (C.c_int, [vp, C.c_int]),
    "cm_get_last_stats": (C.c_int, [vp, C.c_int]),
    "BMOHCKernel_Wait_Threads": (C.c_int, []),
    "BMOHCKernel_Copy_Results": (C.c_int, [vp, sz, sz]),
    "ReverseKernel_Copy_Results": (C.c_int, [vp, sz]),
    "extractConvergenceTimes": (C.c_int, [vp, sz]),
    "getCudaFailedTimeNodes": (vp, []),
    "getCudaFailedBlocks": (vp, []),
    "getCudaConvergeSpeedBlocks": (vp, []),
    "getCudaConvergedTimeNodes": (vp, []),
    "getCudaConvergedBlocks": (vp, []),
    "getCudaConvergedJacobyLoopsPerIteration": (vp, []),
    "getCudaConvergedJacobyLoopsPerIterationBlocks": (vp, []),
    "getCudaBMVelocity": (vp, []),
    "getCudaInputSamples": (vp, []),
    "GeneralKernel_Copy_Results_Bytes": (C.c_int, [vp, vp, sz, sz]),
    "GeneralKernel_Copy_Results": (C.c_int, [vp, vp, sz]),
    "GeneralKernel_Copy_Results_Double": (C.c_int, [vp, vp, sz]),
    "setupToleranceProfile": (C.c_int, [vp, C.c_int, f32, f32, C.c_int, f32, f32, C.c_int, C.c_int, C.c_int]),
    "BMOHCKernel_Free": (C.c_int, []),
    "IHCNewKernel": (C.c_int, [vp, vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, f64, vp, C.c_int,
                               C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, f32, f32, f32, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_int, C.c_int, f32, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int]),
    "loadAihc": (C.c_int, [vp]),
    "enableloadAihc": (C.c_int, []),
    "allocateBuffer": (C.c_int, [C.c_int, C.c_int]),
    "releaseBuffer": (C.c_int, [C.c_int]),
    "RunIHCKernel": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "BMOHCKernel_Copy_Lambda": (C.c_int, [vp, sz, C.c_int]),
    "ReverseKernel_Copy_Results_Float": (C.c_int, [vp, vp, sz, sz, C.c_int]),
    "updateCUDALambdaArray": (C.c_int, [vp, vp, sz, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cm_fix_lambda_run": (C.c_int, [vp, vp, C.c_int64, vp]),
    "getCudaLambda": (vp, []),
    "getCudaBuffer": (vp, []),
    "getCudaMeanRate": (vp, []),
    "cm_ihc_lambda_run": (C.c_int, [vp, i64, vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, vp, vp, f32,
                                    f32, f32, f32, vp, vp, vp]),
    "cm_ihc_lambda_run_noalloc": (C.c_int, [vp, i64, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, vp, f32, f32, f32,
                                            f32, vp, vp]),
    "CudaCalculateJND": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, f64, f64, f64, vp, vp, C.c_int, vp, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp,
                                   C.c_int, C.c_int, f64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cm_jnd_run": (C.c_int, [vp, i64, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp, vp, f64, f64,
                             C.c_int, C.c_int, vp, vp, vp]),
    "cm_complex_jnd_intervals": (C.c_int, [C.c_int, C.c_int, C.c_int, vp]),
    "cm_complex_jnd": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, f32, vp, vp, vp]),
    "cm_jnd_final": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, f64, f64, C.c_int, C.c_int, vp,
                               vp, vp, vp]),
    "cm_remez_lowpass": (C.c_int, [C.c_int, f64, f64, f64, f64, C.c_int, vp, vp, vp]),
    "cm_design_ac_filter": (C.c_int, [f64, C.c_char_p, vp, vp, C.c_int, vp, vp]),
    "cm_post_bm_graph_create": (C.c_int, [vp, C.c_int64, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, vp, f32,
                                          f32, f32, f32, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp,
                                          C.c_int, vp, vp, f64, f64, C.c_int, C.c_int, vp, vp, vp, vp]),
    "cm_post_bm_graph_launch": (C.c_int, [vp]),
    "cm_post_bm_graph_destroy": (C.c_int, [vp]),
    "cm_jnd_failed_signals": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, C.c_int, vp]),
    "InputProfilesArrayInitialization": (C.c_int, [C.c_int] * 5),
    "InputProfilesArrayTermination": (C.c_int, []),
    "uploadProfiles": (C.c_int, [vp, C.c_int, C.c_int]),
    "cm_input_normalization": (C.c_int, [vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, C.c_int, f64, f64, C.c_int,
                                         C.c_int, C.c_int, C.c_int, f32, vp]),
    "generateInputFromProfile": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, C.c_int, C.c_int, C.c_int, f64, f64,
                                           C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.c_int, C.c_int, f32, C.c_int, vp, C.c_int, C.c_int,
                                           f32, f32, C.c_int, f32, f32, C.c_int, C.c_int, vp, C.c_int, vp,
                                           C.c_int]),
    "cm_generate_input_run": (C.c_int, [vp, vp, vp, vp, f64, f64, C.c_int, f64, f64, C.c_int, C.c_int, C.c_int,
                                        C.c_int, f32]),
    "IHCKernel_Free": (C.c_int, []),
    "cm_model_vectors": (C.c_int, [vp]),
    "cm_tri_matrix": (C.c_int, [vp, vp, vp]),
    "cm_pack_section_params": (C.c_int, [vp]),
    "cm_default_bm_scalars": (C.c_int, [f32, vp]),
    "cm_gemm_bf16": (C.c_int, [C.c_int] * 5 + [vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, f32]),
    "cm_gelu_fwd": (C.c_int, [vp, vp, vp, i64]),
    "cm_gelu_bwd": (C.c_int, [vp, vp, vp, vp, i64]),
    "cm_to_bf16": (C.c_int, [vp, vp, i64]),
    "cm_colsum": (C.c_int, [vp, vp, C.c_int, C.c_int]),
    "cm_mse_grad": (C.c_int, [vp, vp, vp, vp, vp, C.c_int, C.c_int]),
    "cm_adam": (C.c_int, [vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, C.c_int, f32]),
    "cm_ws_load": (C.c_int, [vp, vp, vp, vp, vp, vp, f32]),
    "cm_ws_set_every": (C.c_int, [C.c_int]),
    "cm_ws_set_group": (C.c_int, [C.c_int]),
    "cm_ws_set_capture": (C.c_int, [vp, i64, C.c_int, C.c_int]),
    "cm_ws_load_file": (C.c_int, [C.c_char_p]),
    "cm_ws_targets": (C.c_int, [vp, vp, i64, f32, vp]),
    "cm_ws_release": (C.c_int, []),
}
WS_RECORD = 1544      # floats per captured training record
WS_REC_Y, WS_REC_PLIN, WS_REC_OUT = 772, 1028, 1288


class CochleaError(RuntimeError):
    pass


def lib():
    """Load libcochlea_amd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CochleaError(f"{LIB_PATH} is missing: run `python -m cochlearmodel_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGS)


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().cm_last_error().decode(errors="replace")
        raise CochleaError(f"{what} failed ({rc}): {msg}")


def ptr(a):
    """Address of a numpy array, a torch tensor or None."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    if isinstance(a, int):
        return C.c_void_p(a)
    raise TypeError(type(a))
