"""ctypes binding of libchaoslab_hip.so (the C ABI declared in include/chaoslab_hip.h).

The library is built in-tree by chaos_amd/build.py. Loading fails loudly when it is missing:
there is no CPU fallback in the product path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchaoslab_hip.so")

CL_OK, CL_ERR_NO_DEVICE, CL_ERR_UNSUPPORTED, CL_ERR_OOM, CL_ERR_DEVICE_LOST, CL_ERR_ARGUMENT = 0, 2, 3, 4, 9, 10
CL_ELT_F64, CL_ELT_F64_F32TAN, CL_ELT_DD = 1, 2, 3
CL_SYS_STANDARD_MAP, CL_SYS_WRAPPED_STANDARD_MAP, CL_SYS_STANDARD_MAP_CHAIN, CL_SYS_HENON_HEILES_GAUSS = 1, 2, 3, 4
CL_SYS_STANDARD_MAP_K = 5
CL_SYS_PAIR_STANDARD_MAP = 6
CL_SYS_PAIR_WRAPPED_STANDARD_MAP = 7
CL_RED_QR, CL_RED_BENETTIN, CL_RED_MEGNO, CL_RED_SALI, CL_RED_FLI, CL_RED_OFLI, CL_RED_GALI = 1, 2, 3, 4, 5, 6, 7
CL_RED_MONITOR = 9
CL_RED_RUNMONITOR = 10
CL_RED_STRETCHING = 11
CL_QR_CGS, CL_QR_MGS, CL_QR_RGS, CL_QR_HOUSEHOLDER, CL_QR_AUGMENTED, CL_QR_GIVENS = 1, 2, 3, 4, 5, 6
CL_MAX_REDUCERS = 8
CL_CHAIN_RESET_FRAME = 1

# FALLBACK_REASONS of the reference (src/Foundation/Foundation.jl:71-75), indexed from 1.
FALLBACK_REASONS = [
    "CUDA.jl not loadable", "no functional device", "kernel does not compile for the element type",
    "device out of memory", "NumericalHealthError raised only on the GPU", "reducer is CPU-only",
    "GPU not faster for the element type", "launch exceeds the display watchdog", "device lost",
]


class cl_system(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("N", ctypes.c_int32), ("ring", ctypes.c_int32),
                ("stages", ctypes.c_int32), ("maxiter", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("K", ctypes.c_double), ("eps", ctypes.c_double), ("h", ctypes.c_double), ("tol", ctypes.c_double)]


CL_RED_BIRKHOFF = 12
CL_RED_RLI = 13
CL_RED_PHASE_ADVANCE = 14
CL_OBS_COS, CL_OBS_SIN, CL_OBS_LIFTED = 1, 2, 3
CL_OBS_PLANAR, CL_OBS_LIFTED_ANGLE = 4, 5
CL_WB_MAX_WINDOWS, CL_WB_MAX_OBSERVABLES = 16, 8


class cl_birkhoff(ctypes.Structure):
    _fields_ = [("H", ctypes.c_int32), ("K", ctypes.c_int32), ("order", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("window", ctypes.c_int64 * CL_WB_MAX_WINDOWS), ("kind", ctypes.c_int32 * CL_WB_MAX_OBSERVABLES),
                ("a", ctypes.c_int32 * CL_WB_MAX_OBSERVABLES), ("b", ctypes.c_int32 * CL_WB_MAX_OBSERVABLES)]


class cl_reducer(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("alg", ctypes.c_int32), ("replace", ctypes.c_int32),
                ("frame", ctypes.c_int32), ("tol", ctypes.c_double), ("tol2", ctypes.c_double),
                ("lo", ctypes.c_double), ("hi", ctypes.c_double), ("bins", ctypes.c_int32), ("pad2_", ctypes.c_int32),
                ("birkhoff", ctypes.POINTER(cl_birkhoff))]


class cl_chain(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("flags", ctypes.c_int32), ("r", cl_reducer * CL_MAX_REDUCERS)]


class cl_lattice(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int32), ("pad_", ctypes.c_int32), ("m", ctypes.c_int64), ("Ns", ctypes.c_void_p),
                ("dx", ctypes.c_void_p), ("dW", ctypes.c_void_p), ("dstates", ctypes.c_void_p),
                ("dcode", ctypes.c_void_p), ("dhstep", ctypes.c_void_p)]


class cl_tuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("split", ctypes.c_int32), ("pipe", ctypes.c_int32),
                ("chain_qr", ctypes.c_int32)]


CL_MLP_MAX_LAYERS = 8


class cl_mlp_net(ctypes.Structure):
    _fields_ = [("nlayers", ctypes.c_int32), ("sizes", ctypes.c_int32 * (CL_MLP_MAX_LAYERS + 1))] + [
        (nm, ctypes.c_void_p * CL_MLP_MAX_LAYERS) for nm in ("W", "Wb", "b", "mW", "vW", "mb", "vb")]


class cl_probe_t(ctypes.Structure):
    _fields_ = [("loadable", ctypes.c_int32), ("functional", ctypes.c_int32), ("compiles", ctypes.c_int32),
                ("gpu_faster", ctypes.c_int32), ("device", ctypes.c_char * 64)]


P = ctypes.c_void_p
I64 = ctypes.c_int64
I = ctypes.c_int
DP = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes), exactly the declarations of include/chaoslab_hip.h
SIGNATURES = {
    "cl_probe": (I, [I, ctypes.POINTER(cl_probe_t)]),
    "cl_free_memory": (I64, []),
    "cl_create": (P, [I]),
    "cl_destroy": (None, [P]),
    "cl_last_error": (ctypes.c_char_p, [P]),
    "cl_set_tuning": (I, [P, ctypes.POINTER(cl_tuning)]),
    "cl_get_tuning": (I, [P, ctypes.POINTER(cl_tuning)]),
    "cl_state_bytes": (I64, [ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I, I, I]),
    "cl_propagate_chunk": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I, I64, I, I, I, I64, I64,
                               I64, I, P, P, P, P, P, P, P]),
    "cl_propagate_device": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I, I64, I, I, I, I64, I64,
                                I64, I, P, P, P, P, P, P, P]),
    "cl_propagate_device_ws": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I, I64, I, I, I, I64, I64,
                                   I64, I, P, P, P, P, P, P, P, P]),
    "cl_workspace_entries": (I64, [ctypes.POINTER(cl_chain), I, I]),
    "cl_birkhoff_weights": (I, [I64, I, P]),
    "cl_propagate_device_traced": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I, I64, I, I, I, I64,
                                       I64, I64, I, P, P, P, P, P, P, P, I64, I, P]),
    "cl_propagate_ragged": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I, I64, I, I, I, P, P, I,
                                P, P, P, P, P, P]),
    "cl_propagate_mixed": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I, I, ctypes.POINTER(cl_lattice),
                               I, I, I64, P]),
    "cl_propagate_mixed_noalloc": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I, I,
                                       ctypes.POINTER(cl_lattice), I, I, I64, P]),
    "cl_record_rows": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I64, I, I, I, I, P, P, P, P, P,
                           P]),
    "cl_ind_record_rows": (I, [P, ctypes.POINTER(cl_system), ctypes.POINTER(cl_chain), I64, I, I, I, P, P, P, P, P,
                               P]),
    "cl_ind_row_width": (I, [ctypes.POINTER(cl_chain), I]),
    "cl_shadow_device": (I, [P, ctypes.POINTER(cl_system), I64, I, I64, I64, I64, I, ctypes.c_double, P, P, P, P]),
    "cl_propagate_shadow_chunk": (I, [P, ctypes.POINTER(cl_system), I64, I, I64, I64, I64, I, ctypes.c_double, P, P,
                                      P]),
    "cl_shadow_record_rows": (I, [P, I64, I, I, I, I64, P, P, P, P]),
    "cl_initial_deviations": (I, [ctypes.c_uint64, I64, I64, I, I, P]),
    "cl_initial_deviations_device": (I, [P, ctypes.c_uint64, I64, I64, I, I, P, P]),
    "cl_math_host": (None, [I64, P, P, P, P]),
    "cl_gw_host": (None, [I64, P, P, P, P, P]),
    "cl_math_device": (I, [P, I64, P, P, P, P, P]),
    "cl_sqrt12_device": (I, [P, I64, P, P, P]),
    "cl_mlp_gemm": (I, [P, I, I, I, I, I, P, I, P, I, P, I, I, P, I, P, I, P]),
    "cl_mlp_rowsum": (I, [P, I, I, P, I, P, P]),
    "cl_mlp_loss": (I, [P, I, P, P, P, P, ctypes.c_float, P, P, P]),
    "cl_mlp_adam": (I, [P, I64, P, P, P, P, P, I, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                        ctypes.c_float, P]),
    "cl_mlp_normalize": (I, [P, I, I64, P, P, P, P, P]),
    "cl_mlp_feature_stats": (I, [P, I, I64, P, P, P, P]),
    "cl_f32_to_bf16": (I, [P, I64, P, P, P]),
    "cl_mlp_train_step": (I, [P, P, P, I, I, P, P, P, ctypes.c_float, I, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float, ctypes.c_float, P, P]),
}

_lib = None


def load():
    """Load the library (once) and set the argument types of every exported symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m chaos_amd.build` "
                          "(the product has no CPU fallback)")
    # torch first: the library then binds to the HIP runtime torch has loaded, so the process holds one runtime
    # (loaded the other way round, two runtimes see no device in one of them)
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HIPError(RuntimeError):
    """A failed library call; `code` is the FALLBACK_REASONS index (or 10 for an invalid argument)."""

    def __init__(self, code, msg):
        reason = FALLBACK_REASONS[code - 1] if 1 <= code <= len(FALLBACK_REASONS) else "invalid argument"
        super().__init__(f"chaoslab_hip error {code} ({reason}): {msg}")
        self.code = code
        self.reason = reason
