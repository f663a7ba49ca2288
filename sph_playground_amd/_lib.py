"""ctypes binding of libsph_hip.so (the C-ABI of include/sph_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsph_hip.so")

# exported entry points, as declared in include/sph_hip.h
EXPORTS = (
    "sph_default_options", "sph_ctx_create", "sph_ctx_destroy", "sph_iterate", "sph_iterate_device",
    "sph_load", "sph_save", "sph_step", "sph_neighbours", "sph_leapfrog", "sph_sts", "sph_sts_coeffs",
    "sph_get_stats", "sph_error_string", "sph_diff_prepare", "sph_diff_apply", "sph_diff_apply_adjoint",
    "sph_diff_forward", "sph_diff_backward", "sph_rad_exchange", "sph_batch_create", "sph_batch_destroy",
    "sph_batch_sts", "sph_batch_get_info", "sph_batch_offsets", "sph_neighbours_device_noalloc",
    "sph_prerun_check", "sph_set_timing",
)

SPH_OK = 0
ERRORS = {
    1: "NEIGH_TRUNCATED", 2: "NEG_H", 3: "Q_NAN_OR_NEG", 4: "FLD_ANISO", 5: "BAD_DDW",
    6: "PRECONDITION", 7: "BAD_ARG", 8: "HIP", 9: "NOMEM", 10: "UNSUPPORTED", 11: "RAD_NEGATIVE",
}


class Options(ct.Structure):
    _fields_ = [("precision", ct.c_int), ("tensor_mode", ct.c_int), ("kappa_perp", ct.c_double),
                ("max_sweeps", ct.c_int), ("newton_tol", ct.c_double), ("neigh_cap", ct.c_int),
                ("device", ct.c_int), ("cell_div", ct.c_int)]


class Stats(ct.Structure):
    _fields_ = [("n_real", ct.c_int64), ("n_fixed", ct.c_int64), ("n_pairs", ct.c_int64),
                ("loop1_pairs", ct.c_int64), ("sweeps", ct.c_int), ("sorted", ct.c_int),
                ("energy", ct.c_double), ("ms_neigh", ct.c_double), ("ms_loop1", ct.c_double),
                ("ms_loop2", ct.c_double), ("ms_total", ct.c_double), ("ms_k_neigh", ct.c_double),
                ("ms_k_loop1", ct.c_double), ("ms_k_loop2", ct.c_double),
                ("ext_records", ct.c_int64)]


class BatchInfo(ct.Structure):
    _fields_ = [("n_problems", ct.c_int), ("n_total", ct.c_int64), ("n_pairs", ct.c_int64),
                ("list_cap", ct.c_int), ("block", ct.c_int), ("tune_block", ct.c_int * 4),
                ("tune_ms", ct.c_float * 4)]


class SphError(RuntimeError):
    """A non-zero SPH_* code; ``code`` mirrors the reference abort it replaces."""

    def __init__(self, code, where=""):
        self.code = code
        self.name = ERRORS.get(code, "UNKNOWN")
        msg = lib().sph_error_string(code).decode() if _lib is not None else self.name
        super().__init__("%s%s (%d: %s)" % (where + ": " if where else "", msg, code, self.name))


_lib = None


def lib():
    """Load libsph_hip.so once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libsph_hip.so is not built (python -m sph_playground_amd.build)")
    L = ct.CDLL(LIB_PATH)
    D = ct.POINTER(ct.c_double)
    P = ct.c_void_p
    L.sph_default_options.argtypes = [ct.POINTER(Options)]
    L.sph_ctx_create.argtypes = [D, ct.c_int, ct.POINTER(Options), ct.POINTER(P)]
    L.sph_ctx_destroy.argtypes = [P]
    L.sph_iterate.argtypes = [P, ct.c_int64, D, ct.c_int64, D, D]
    L.sph_iterate_device.argtypes = [P, ct.c_int64, ct.c_void_p, ct.c_int64, D, D]
    L.sph_load.argtypes = [P, ct.c_int64, ct.c_void_p, ct.c_int64, ct.c_int]
    L.sph_save.argtypes = [P, ct.c_void_p, ct.c_int64, ct.c_int]
    L.sph_step.argtypes = [P, D, D]
    L.sph_neighbours.argtypes = [P, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int8),
                                 ct.c_int64, ct.POINTER(ct.c_int64)]
    L.sph_leapfrog.argtypes = [P, D, ct.c_double, D]
    L.sph_sts.argtypes = [P, D, ct.c_int, ct.c_double]
    L.sph_rad_exchange.argtypes = [P, D, ct.c_double]
    L.sph_sts_coeffs.argtypes = [ct.c_int, ct.c_int, D]
    L.sph_get_stats.argtypes = [P, ct.POINTER(Stats)]
    L.sph_set_timing.argtypes = [P, ct.c_int]
    L.sph_diff_prepare.argtypes = [P, D]
    L.sph_batch_create.argtypes = [ct.POINTER(P), ct.c_int, D, ct.c_int, ct.POINTER(P)]
    L.sph_batch_destroy.argtypes = [P]
    L.sph_batch_sts.argtypes = [P, ct.c_int, D, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int]
    L.sph_batch_get_info.argtypes = [P, ct.POINTER(BatchInfo)]
    L.sph_batch_offsets.argtypes = [P, ct.POINTER(ct.c_int64)]
    L.sph_diff_apply.argtypes = [P, D, ct.c_void_p, ct.c_void_p]
    L.sph_diff_apply_adjoint.argtypes = [P, D, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, D, D]
    L.sph_diff_forward.argtypes = [P, D, ct.c_int, ct.c_double, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p]
    L.sph_diff_backward.argtypes = [P, ct.c_void_p, ct.c_void_p, D]
    L.sph_neighbours_device_noalloc.argtypes = [P, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int64,
                                                ct.POINTER(ct.c_int64)]
    L.sph_prerun_check.argtypes = [P, D, ct.POINTER(ct.c_int)]
    L.sph_error_string.argtypes = [ct.c_int]
    L.sph_error_string.restype = ct.c_char_p
    for name in EXPORTS:
        if name != "sph_error_string":
            getattr(L, name).restype = ct.c_int
    _lib = L
    return L


def check(rc, where=""):
    if rc != SPH_OK:
        raise SphError(rc, where)
