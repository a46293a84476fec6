"""ctypes bindings of include/neighbor_amd.h (libneighbor_amd.so)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libneighbor_amd.so"

NB_OK = 0
NB_ERR_INVALID = -1
NB_ERR_TUNING = -2
NB_ERR_IMAGES = -3
NB_ERR_HIP = -4
NB_ERR_ALLOC = -5
NB_ERR_INTERNAL = -6
NB_SENTINEL = 0x7FFFFFFF
NB_VARIANT_QUANTIZED = 0
NB_VARIANT_FLOAT_ROPE = 1
NB_VARIANT_FLOAT_ROPE_EXACT = 2
NB_VARIANT_WAVE = 3
NB_VARIANT_WAVE_QUANTIZED = 4
NB_BATCH_WAVE = 0
NB_BATCH_LANE = 1

# (name, restype, argtypes) for every symbol the header declares
P = ctypes.c_void_p
U32 = ctypes.c_uint32
I32 = ctypes.c_int32
F32 = ctypes.c_float
SZ = ctypes.c_size_t
PP = ctypes.POINTER(ctypes.c_void_p)
PU32 = ctypes.POINTER(ctypes.c_uint32)
PI32 = ctypes.POINTER(ctypes.c_int32)
INT = ctypes.c_int

SIGNATURES = [
    ("nb_last_error_string", ctypes.c_char_p, []),
    ("nb_version", INT, []),
    ("nb_lbvh_create", INT, [PP]),
    ("nb_lbvh_destroy", INT, [P]),
    ("nb_lbvh_build_points", INT, [P, P, U32, P, P, U32, P]),
    ("nb_lbvh_build_spheres", INT, [P, P, U32, F32, P, P, U32, P]),
    ("nb_lbvh_build_postype", INT, [P, P, P, U32, F32, P, P, U32, P]),
    ("nb_lbvh_refit_postype", INT, [P, P, P, U32, F32, U32, P]),
    ("nb_lbvh_refit_points", INT, [P, P, U32, U32, P]),
    ("nb_lbvh_refit_spheres", INT, [P, P, U32, F32, U32, P]),
    ("nb_max_displacement2", INT, [P, P, U32, P, P, P, P]),
    ("nb_lbvh_info", INT, [P, PU32, PU32, PU32, PI32]),
    ("nb_lbvh_arrays", INT, [P, PP, PP, PP, PP, PP, PP, PP, PP]),
    ("nb_lbvh_tunable_parameters", INT, [P, P, U32, PU32]),
    ("nb_traverser_create", INT, [PP]),
    ("nb_traverser_destroy", INT, [P]),
    ("nb_traverser_set_variant", INT, [P, INT]),
    ("nb_traverser_set_bucket_size", INT, [P, U32]),
    ("nb_traverser_set_tail_split", INT, [P, U32, U32]),
    ("nb_traverser_setup", INT, [P, P, P, U32, P]),
    ("nb_traverser_set_point_stride", INT, [P, U32]),
    ("nb_traverser_set_visit_log", INT, [P, P, ctypes.c_uint64]),
    ("nb_visit_log_words", SZ, [U32]),
    ("nb_traverser_reset", INT, [P]),
    ("nb_traverser_data", INT, [P, PP, PP, PP, PP, PP]),
    ("nb_traverse_count", INT, [P, P, P, U32, P, U32, P, P, P, U32, P]),
    ("nb_traverse_list", INT, [P, P, P, U32, P, U32, P, P, P, P, U32, U32, P]),
    ("nb_traverse_pairs", INT, [P, P, P, U32, P, U32, P, P, P, P, P, P, ctypes.c_uint64, U32, P]),
    ("nb_traverse_exact_count", INT, [P, P, P, U32, P, P, U32, P, P, P, U32, P]),
    ("nb_traverse_exact_pairs", INT, [P, P, P, U32, P, P, U32, P, P, P, P, P, P, ctypes.c_uint64, U32, P]),
    ("nb_scan_scratch_bytes", SZ, [U32]),
    ("nb_exclusive_scan_u32_u64", INT, [P, U32, P, P, P]),
    ("nb_exclusive_scan_permuted_u32_u64", INT, [P, P, U32, P, P, P]),
    ("nb_exclusive_scan_u32_u64_noalloc", INT, [P, P, U32, P, P]),
    ("nb_rows_filter", INT, [P, P, U32, P, P, P, P, INT, U32, P, P, P, P, P, U32, P]),
    ("nb_rows_pack", INT, [P, P, U32, P, P, P, P, P, P, P, P]),
    ("nb_traverse_exact_rows", INT, [P, P, P, U32, P, P, U32, P, P, P, P, P, P, P, ctypes.c_uint64, P, U32, P]),
    ("nb_traverse_exact_rows_placed", INT, [P, P, P, U32, P, P, U32, P, P, P, P, P, P, P, ctypes.c_uint64, P, P]),
    ("nb_capacity_create", INT, [PP, ctypes.c_uint64]),
    ("nb_capacity_destroy", INT, [P]),
    ("nb_capacity_parameters", INT, [P, P, P, PU32]),
    ("nb_capacity_features", INT, [P, P, P, P, F32, PP, P]),
    ("nb_capacity_train", INT, [P, P, P, U32, U32, F32, ctypes.c_uint64, ctypes.POINTER(F32), P]),
    ("nb_capacity_margin", INT, [P, P, F32, ctypes.POINTER(F32), P]),
    ("nb_capacity_predict", INT, [P, P, F32, U32, P, P]),
    ("nb_capacity_predict_direct", INT, [P, P, P, P, F32, F32, U32, P, P]),
    ("nb_capacity_predict_rows", INT, [P, P, P, P, F32, F32, U32, P, P, P, P]),
    ("nb_capacity_forward", INT, [P, P, U32, P, P]),
    ("nb_capacity_gradient", INT, [P, P, P, U32, P, P]),
    ("nb_overflow_rows", INT, [P, P, U32, P, P, P]),
    ("nb_relocate_scratch_bytes", SZ, [U32]),
    ("nb_relocate_rows", INT, [P, U32, P, ctypes.c_uint64, P, P, P, P]),
    ("nb_lbvh_error_word", INT, [P, PP]),
    ("nb_batch_error_word", INT, [P, PP]),
    ("nb_lbvh_check", INT, [P, P]),
    ("nb_lbvh_enable_sorted_event", INT, [P, INT]),
    ("nb_lbvh_wait_sorted", INT, [P, P]),
    ("nb_lbvh_set_span_lds_limit", INT, [P, U32]),
    ("nb_batch_check", INT, [P, P]),
    ("nb_batch_create", INT, [PP]),
    ("nb_batch_destroy", INT, [P]),
    ("nb_batch_build", INT, [P, P, U32, P, P, P, U32, P]),
    ("nb_batch_set_params", INT, [P, U32, U32, INT, U32]),
    ("nb_batch_get_params", INT, [P, PU32, PU32, ctypes.POINTER(INT), PU32]),
    ("nb_batch_setup", INT, [P, P]),
    ("nb_batch_count", INT, [P, P, INT, P, P]),
    ("nb_batch_pairs", INT, [P, P, INT, P, P, P, P, ctypes.c_uint64, P]),
    ("nb_batch_rows", INT, [P, P, INT, P, P, P, P, P, ctypes.c_uint64, P, P]),
    ("nb_batch_capacity", INT, [P, P, F32, U32, P, P]),
    ("nb_batch_arrays", INT, [P, PP, PP, PP, PP, PP, PP, PP, PP, PP, PU32, PU32, PU32]),
    ("nb_batch_tune", INT, [P, P, U32, P, P, P, U32, P, INT, P, PU32, PU32, ctypes.POINTER(INT),
                            ctypes.POINTER(F32), ctypes.POINTER(F32), P]),
    ("nb_batch_plan_create", INT, [PP, P, P, U32, P, P, P, U32, P, INT, P, P, P, P, P, ctypes.c_uint64]),
    ("nb_batch_plan_launch", INT, [P, P]),
    ("nb_batch_plan_destroy", INT, [P]),
]


class NeighborError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where} failed with status {code}: {message}")
        self.code = code


_lib = None


def library_path() -> str:
    return os.path.join(HERE, _LIB_NAME)


def lib() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python neighbor_amd/build.py` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(path)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int, where: str) -> None:
    if code != NB_OK:
        msg = lib().nb_last_error_string()
        raise NeighborError(code, where, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)


# --- minimal HIP runtime access for device-to-device copies of library-owned arrays
_hip = None


def hip():
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMemcpyAsync.restype = INT
        _hip.hipMemcpyAsync.argtypes = [P, P, SZ, INT, P]
    return _hip


def memcpy_d2d(dst: int, src: int, nbytes: int, stream: int) -> None:
    if nbytes == 0:
        return
    rc = hip().hipMemcpyAsync(P(dst), P(src), nbytes, 3, P(stream))  # 3 = hipMemcpyDeviceToDevice
    if rc != 0:
        raise RuntimeError(f"hipMemcpyAsync failed ({rc})")
