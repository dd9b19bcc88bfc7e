"""ctypes binding of libphylocsf_amd.so (the C ABI in include/phylocsf_hip.h).

The HIP library is the only implementation: if it is missing, every call
raises; there is no CPU fallback in the package.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libphylocsf_amd.so")

PCSF_OK = 0
PCSF_EINVAL = -1
PCSF_EFAIL = -2
PCSF_EDEVICE = -3
PCSF_ENOMEM = -4
PCSF_MARGINALIZE = 255
PCSF_FP64 = 0
PCSF_FP32 = 1
PCSF_FP16S = 2
PCSF_FP16 = 3
PCSF_BF16 = 4
PCSF_FP64M = 5  # FP64 on matrix cores
PCSF_N_VARIANTS = 7  # launch shapes of the fused kernel (pcsf_autotune)
PCSF_NOALLOC_SCRATCH_BYTES = 32 << 20  # scratch of pcsf_forward_noalloc, per GPU

EXPORTED = [
    "pcsf_ctx_create", "pcsf_ctx_destroy", "pcsf_last_error", "pcsf_ctx_set_timing",
    "pcsf_ctx_stream", "pcsf_ctx_synchronize", "pcsf_ecm_read", "pcsf_newick_tree",
    "pcsf_model_create", "pcsf_model_destroy", "pcsf_model_info", "pcsf_model_transition",
    "pcsf_batch_create", "pcsf_batch_destroy", "pcsf_batch_columns", "pcsf_forward",
    "pcsf_backward", "pcsf_batch_device_results", "pcsf_batch_last_timing",
    "pcsf_lik_columns", "pcsf_simulate", "pcsf_autotune", "pcsf_set_variant", "pcsf_score",
    "pcsf_model_specialized", "pcsf_batch_create_frames", "pcsf_simulate_nucleotides",
    "pcsf_program", "pcsf_batch_codes", "pcsf_forward_noalloc", "pcsf_score_omega",
    "pcsf_lik_posteriors",
]
PCSF_STRATEGY_FIXED = 0
PCSF_STRATEGY_MLE = 1
PCSF_STRATEGY_MLE_GRAD = 2

_lib = None


class PcsfError(RuntimeError):
    """A non-zero status from the C ABI.  code mirrors the reference's
    exception kind: EINVAL = Invalid_argument, EFAIL = Failure."""

    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


class InvalidArgument(PcsfError):
    pass


class Failure(PcsfError):
    pass


def raise_for(code, msg):
    if code == PCSF_EINVAL:
        raise InvalidArgument(code, msg)
    if code == PCSF_EFAIL:
        raise Failure(code, msg)
    raise PcsfError(code, msg)


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libphylocsf_amd.so is not built (run __graft_entry__.build() or "
                          "make -C phylocsf_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    i32p = C.POINTER(C.c_int32)
    i64p = C.POINTER(C.c_int64)
    u8p = C.POINTER(C.c_uint8)
    ip = C.POINTER(C.c_int)
    L.pcsf_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pcsf_ctx_destroy.argtypes = [vp]
    L.pcsf_ctx_destroy.restype = None
    L.pcsf_last_error.argtypes = [vp]
    L.pcsf_last_error.restype = C.c_char_p
    L.pcsf_ctx_set_timing.argtypes = [vp, C.c_int]
    L.pcsf_ctx_stream.argtypes = [vp]
    L.pcsf_ctx_stream.restype = vp
    L.pcsf_ctx_synchronize.argtypes = [vp]
    L.pcsf_ecm_read.argtypes = [C.c_char_p, dp, dp, C.c_char_p, C.c_int]
    L.pcsf_newick_tree.argtypes = [C.c_char_p, C.c_char_p, C.c_int, ip, i32p, dp, C.c_char_p,
                                   C.c_int, C.c_char_p, C.c_int]
    L.pcsf_model_create.argtypes = [vp, C.c_int, i32p, dp, C.c_int, dp, dp, C.POINTER(vp)]
    L.pcsf_model_destroy.argtypes = [vp]
    L.pcsf_model_destroy.restype = None
    L.pcsf_model_info.argtypes = [vp, ip, ip, ip, ip]
    L.pcsf_model_transition.argtypes = [vp, C.c_int, C.c_double, dp]
    L.pcsf_batch_create.argtypes = [vp, u8p, C.c_int64, i64p, C.c_int32, C.POINTER(vp)]
    L.pcsf_batch_create_frames.argtypes = [vp, u8p, C.c_int64, i64p, C.c_int32, C.c_int,
                                           C.POINTER(vp)]
    L.pcsf_batch_destroy.argtypes = [vp]
    L.pcsf_batch_destroy.restype = None
    L.pcsf_batch_columns.argtypes = [vp]
    L.pcsf_batch_columns.restype = C.c_int64
    L.pcsf_batch_codes.argtypes = [vp, u8p, i64p]
    L.pcsf_forward.argtypes = [vp, dp, C.c_int, C.c_int, dp, dp, C.c_int]
    L.pcsf_backward.argtypes = [vp, dp, dp, dp]
    L.pcsf_forward_noalloc.argtypes = [vp, vp, C.c_int64, vp, C.c_int32, dp, C.c_int, vp, vp]
    L.pcsf_batch_device_results.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
    L.pcsf_batch_last_timing.argtypes = [vp, dp, dp, dp]
    L.pcsf_lik_columns.argtypes = [vp, C.c_int, C.c_int, i32p, dp, dp, i32p, C.c_int64, dp, dp]
    L.pcsf_lik_posteriors.argtypes = [vp, C.c_int, C.c_int, i32p, dp, dp, i32p, dp, C.c_int64,
                                      dp, dp, dp, dp]
    L.pcsf_simulate.argtypes = [vp, C.c_int32, i64p, i32p, dp, C.c_int, C.c_double,
                                C.c_double, C.c_double, C.c_uint64, C.c_int, u8p]
    L.pcsf_simulate_nucleotides.argtypes = L.pcsf_simulate.argtypes
    L.pcsf_autotune.argtypes = [vp, C.c_int, C.c_int, dp, ip]
    L.pcsf_set_variant.argtypes = [vp, C.c_int, C.c_int]
    L.pcsf_model_specialized.argtypes = [vp, C.c_int, ip]
    L.pcsf_score.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp, dp, dp, i32p, i32p]
    L.pcsf_score_omega.argtypes = [vp, C.c_double, C.c_double, C.c_int, dp, dp, dp, dp, i32p,
                                   i32p]
    L.pcsf_program.argtypes = [C.c_int, i32p, C.c_int, i32p, C.c_int, ip, ip]
    _lib = L
    return L
