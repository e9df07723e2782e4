"""ctypes binding of include/qudrip_hip.h (libqudrip_hip.so).

Vectors passed to the library are device pointers; the Python side gets them
from torch tensors (``tensor.data_ptr()``) -- torch is only the allocator and
stream provider here.  Loading fails loudly if the HIP library is missing:
there is no CPU fallback anywhere on the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqudrip_hip.so")

QD_OK = 0
STATUS = {
    0: "QD_OK", 1: "QD_E_ARG", 2: "QD_E_OOM", 3: "QD_E_HIP", 4: "QD_E_BREAKDOWN",
    5: "QD_E_EIGEN", 6: "QD_E_CAPACITY", 7: "QD_E_UNSUPPORTED", 8: "QD_E_BUDGET", 9: "QD_E_STATE",
    10: "QD_E_CANCELLED",
}
STATUS_CODE = {v: k for k, v in STATUS.items()}
QD_OP_BIT = 0
QD_OP_HOP = 1
QD_OP_MODE = 2
QD_MAX_MODE_DIM = 16
QD_MAX_OPS_PER_TERM = 8


class qd_constraint(ctypes.Structure):
    _fields_ = [("mask", ctypes.c_uint64), ("count", ctypes.c_int64)]


class qd_op(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("site_a", ctypes.c_int32), ("site_b", ctypes.c_int32),
                ("string", ctypes.c_int32), ("U", ctypes.c_double * 8)]


class qd_term(ctypes.Structure):
    _fields_ = [("coeff", ctypes.c_double * 2), ("op_begin", ctypes.c_int32), ("n_ops", ctypes.c_int32)]


class qd_jump(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("mode", ctypes.c_int32), ("rate", ctypes.c_double)]


class qd_traj_stats(ctypes.Structure):
    _fields_ = [("jumps", ctypes.c_int64), ("losses", ctypes.c_int64), ("sectors", ctypes.c_int32),
                ("lanczos", ctypes.c_int32), ("lanczos_fallback_steps", ctypes.c_int32)]


QD_JUMP_DEPHASE = 0
QD_JUMP_LOSS = 1


QD_REQ_GROUND_STATE = 0
QD_REQ_EVOLVE = 1


class qd_request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("model", ctypes.c_int32), ("coeff", ctypes.c_double),
                ("k", ctypes.c_int32), ("chunk", ctypes.c_int32), ("tol", ctypes.c_double),
                ("steps", ctypes.c_int32), ("dt", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("tag", ctypes.c_int64)]


class qd_result(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_int64), ("status", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("energy", ctypes.c_double), ("energy_start", ctypes.c_double), ("observable", ctypes.c_double),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_end", ctypes.c_double),
                ("krylov_mean", ctypes.c_double), ("worker", ctypes.c_int32)]


class qd_server_stats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("finished", ctypes.c_int64), ("rejected", ctypes.c_int64),
                ("queued", ctypes.c_int64), ("running", ctypes.c_int32), ("max_running", ctypes.c_int32),
                ("bytes_in_use", ctypes.c_int64), ("peak_bytes", ctypes.c_int64)]


class qd_krylov_info(ctypes.Structure):
    _fields_ = [("act_dim", ctypes.c_int), ("breakdown", ctypes.c_int), ("error", ctypes.c_double),
                ("beta_last", ctypes.c_double), ("n_spmv", ctypes.c_int64), ("ritz_resid", ctypes.c_double),
                ("ritz_gap", ctypes.c_double), ("full_steps", ctypes.c_int)]


class qd_quantity(ctypes.Structure):
    _fields_ = [("sel", ctypes.c_uint64), ("flip", ctypes.c_uint64), ("shift", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("count", ctypes.c_int64)]


class qd_hlanczos_info(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("error", ctypes.c_double), ("n_spmv", ctypes.c_int64)]


VP = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
DP = ctypes.POINTER(ctypes.c_double)
PP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); every function declared in include/qudrip_hip.h
SIGNATURES = {
    "qd_abi_version": (I, []),
    "qd_ctx_create": (I, [I, PP]),
    "qd_ctx_create_on_stream": (I, [I, VP, PP]),
    "qd_ctx_destroy": (I, [VP]),
    "qd_last_error": (ctypes.c_char_p, [VP]),
    "qd_ctx_sync": (I, [VP]),
    "qd_ctx_stream": (VP, [VP]),
    "qd_malloc": (I, [VP, ctypes.c_size_t, PP]),
    "qd_free": (I, [VP, VP]),
    "qd_memcpy_h2d": (I, [VP, VP, VP, ctypes.c_size_t]),
    "qd_memcpy_d2h": (I, [VP, VP, VP, ctypes.c_size_t]),
    "qd_basis_create": (I, [VP, I, ctypes.POINTER(qd_constraint), I, PP]),
    "qd_basis_create_q": (I, [VP, I, ctypes.POINTER(qd_quantity), I, PP]),
    "qd_basis_create_modes": (I, [VP, I, I, ctypes.POINTER(qd_quantity), I, PP]),
    "qd_basis_destroy": (I, [VP]),
    "qd_basis_size": (I64, [VP]),
    "qd_basis_download": (I, [VP, VP]),
    "qd_basis_device_states": (VP, [VP]),
    "qd_basis_rank": (I, [VP, VP, I64, VP]),
    "qd_termprog_create": (I, [VP, I, ctypes.POINTER(qd_term), I, ctypes.POINTER(qd_op), I, PP]),
    "qd_termprog_create_modes": (I, [VP, I, I, ctypes.POINTER(qd_term), I, ctypes.POINTER(qd_op), I, DP, I, PP]),
    "qd_termprog_destroy": (I, [VP]),
    "qd_csr_build": (I, [VP, VP, VP, PP]),
    "qd_csr_destroy": (I, [VP]),
    "qd_csr_info": (I, [VP, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I)]),
    "qd_csr_download": (I, [VP, VP, VP, VP]),
    "qd_csr_device_arrays": (I, [VP, PP, PP, PP]),
    "qd_csr_zero": (I, [VP, I64, PP]),
    "qd_csr_identity": (I, [VP, I64, PP]),
    "qd_csr_scale": (I, [VP, VP, ctypes.c_double, ctypes.c_double, I, PP]),
    "qd_csr_add": (I, [VP, VP, VP, I, PP]),
    "qd_csr_adjoint": (I, [VP, VP, PP]),
    "qd_csr_multiply": (I, [VP, VP, VP, PP]),
    "qd_linop_from_csr": (I, [VP, VP, PP]),
    "qd_linop_matfree": (I, [VP, VP, VP, PP]),
    "qd_linop_destroy": (I, [VP]),
    "qd_linop_with_diagonal": (I, [VP, VP, VP, PP]),
    "qd_linop_dim": (I64, [VP]),
    "qd_linop_set_lanes": (I, [VP, I]),
    "qd_linop_get_lanes": (I, [VP]),
    "qd_linop_set_structured": (I, [VP, I, ctypes.POINTER(I)]),
    "qd_linop_autotune": (I, [VP, VP, VP, VP, I, ctypes.POINTER(I), ctypes.POINTER(ctypes.c_float)]),
    "qd_spmv": (I, [VP, VP, VP, VP]),
    "qd_spmv_acc": (I, [VP, VP, DP, VP, VP]),
    "qd_expect": (I, [VP, VP, VP, DP]),
    "qd_expect_noalloc": (I, [VP, VP, VP, VP]),
    "qd_norm2": (I, [VP, I64, VP, DP]),
    "qd_seeded_state": (I, [VP, I64, ctypes.c_uint64, VP]),
    "qd_krylov_create": (I, [VP, I64, I, I, PP]),
    "qd_krylov_destroy": (I, [VP]),
    "qd_krylov_ground_state": (I, [VP, VP, VP, I, DP]),
    "qd_krylov_evolve": (I, [VP, VP, VP, VP, ctypes.c_double, I]),
    "qd_krylov_evolve_tol": (I, [VP, VP, VP, VP, ctypes.c_double, I, I, ctypes.c_double, ctypes.POINTER(I),
                                 ctypes.POINTER(I)]),
    "qd_krylov_get_info": (I, [VP, ctypes.POINTER(qd_krylov_info)]),
    "qd_krylov_eigenvalues": (I, [VP, DP, I, ctypes.POINTER(I)]),
    "qd_krylov_converge_ev": (I, [VP, I, ctypes.c_double, ctypes.c_double]),
    "qd_krylov_reset": (I, [VP]),
    "qd_krylov_set_method": (I, [VP, I]),
    "qd_hlanczos_create": (I, [VP, I64, I, PP]),
    "qd_hlanczos_destroy": (I, [VP]),
    "qd_hlanczos_ground_state": (I, [VP, VP, VP, DP]),
    "qd_hlanczos_evolve": (I, [VP, VP, VP, VP, ctypes.c_double]),
    "qd_hlanczos_get_info": (I, [VP, ctypes.POINTER(qd_hlanczos_info)]),
    "qd_hlanczos_eigenvalues": (I, [VP, DP, I, ctypes.POINTER(I)]),
    "qd_hlanczos_coefficients": (I, [VP, DP, DP, I, ctypes.POINTER(I)]),
    "qd_traj_run": (I, [VP, I, ctypes.POINTER(qd_term), I, ctypes.POINTER(qd_op), I, I, I, VP,
                        ctypes.POINTER(qd_jump), I, I, I, ctypes.c_double, I, ctypes.c_uint64, VP, VP, VP,
                        ctypes.POINTER(qd_traj_stats)]),
    "qd_server_create": (I, [I, ctypes.c_size_t, I, PP]),
    "qd_server_add_model": (I, [VP, I, ctypes.POINTER(qd_constraint), I, ctypes.POINTER(qd_term), I,
                                ctypes.POINTER(qd_op), I, ctypes.POINTER(qd_term), I, ctypes.POINTER(qd_op), I, I,
                                ctypes.POINTER(I)]),
    "qd_server_submit": (I, [VP, ctypes.POINTER(qd_request)]),
    "qd_server_set_batch": (I, [VP, I]),
    "qd_server_set_gs_method": (I, [VP, I]),
    "qd_stream_bandwidth": (I, [VP, ctypes.c_size_t, I, DP, DP]),
    "qd_server_cancel": (I, [VP, ctypes.c_int64]),
    "qd_server_poll": (I, [VP, ctypes.POINTER(qd_result), I, ctypes.POINTER(I)]),
    "qd_server_drain": (I, [VP]),
    "qd_server_clock": (ctypes.c_double, [VP]),
    "qd_server_get_stats": (I, [VP, ctypes.POINTER(qd_server_stats)]),
    "qd_server_last_error": (ctypes.c_char_p, [VP]),
    "qd_server_destroy": (I, [VP]),
    "qd_host_tridiag_eig": (I, [I, DP, DP, DP, DP]),
    "qd_host_expm": (I, [I, DP, DP]),
}

_lib = None


def load():
    """Load and bind the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(the GPU path has no fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class QDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def check(ctx_ptr, rc, what=""):
    if rc != QD_OK:
        msg = load().qd_last_error(ctx_ptr) if ctx_ptr else b""
        raise QDError(rc, f"{what} {msg.decode(errors='replace') if msg else ''}".strip())
