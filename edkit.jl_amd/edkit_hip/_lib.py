"""Loader of libedkit_hip.so (the HIP engine) with its C ABI signatures.

The shared library is built in-tree next to this file.  There is no CPU
fallback: if the library is missing or fails to load, every entry point
raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# EDK_HIP_LIB may name another build of the same library (for A/B timing of
# kernel variants); unset, the in-tree build is loaded
LIB_PATH = os.environ.get("EDK_HIP_LIB") or os.path.join(_HERE, "libedkit_hip.so")

EDK_OK = 0
(EDK_TENSOR, EDK_PROJECTED, EDK_TRANSLATIONAL, EDK_TRANSLATION_PARITY, EDK_ABELIAN, EDK_PARITY, EDK_FLIP,
 EDK_PARITY_FLIP, EDK_TRANSLATION_FLIP) = range(9)
EDK_C64, EDK_C32 = 0, 1
(EDK_VARIANT_AUTO, EDK_VARIANT_GENERIC, EDK_VARIANT_TILE, EDK_VARIANT_LANE2, EDK_VARIANT_LANE4,
 EDK_VARIANT_LANE8) = range(6)
VARIANT_NAMES = {0: "auto", 1: "generic", 2: "tile", 3: "lane2", 4: "lane4", 5: "lane8"}

# every symbol include/edkit_hip.h declares
SYMBOLS = [
    "edk_last_error", "edk_version",
    "edk_ctx_create", "edk_ctx_destroy", "edk_ctx_synchronize", "edk_ctx_stream",
    "edk_malloc", "edk_free", "edk_memcpy_htod", "edk_memcpy_dtoh", "edk_memset", "edk_cache_info",
    "edk_basis_build", "edk_basis_upload", "edk_basis_dim", "edk_basis_order",
    "edk_basis_download", "edk_basis_index", "edk_basis_destroy",
    "edk_operator_create", "edk_operator_destroy", "edk_operator_count_offdiag",
    "edk_matvec", "edk_matvec_host", "edk_matvec_multi",
    "edk_operator_assemble", "edk_operator_download_csr", "edk_operator_release_csr",
    "edk_operator_set_variant", "edk_operator_autotune",
    "edk_batch_create", "edk_batch_create_ex", "edk_batch_launches", "edk_batch_launch", "edk_batch_destroy",
    "edk_server_default_params", "edk_server_create", "edk_server_create_ex", "edk_server_submit",
    "edk_server_wait", "edk_server_clock", "edk_server_stats", "edk_server_batch_stats", "edk_server_destroy",
    "edk_vec_scale", "edk_vec_axpy", "edk_vec_dot", "edk_vec_nrm2", "edk_vec_dot_noalloc",
    "edk_lanczos_ground",
    "edk_lanczos_ground_ex",
    "edk_schmidt", "edk_ent_spec",
    "edk_krylov_default_params", "edk_krylov_create", "edk_krylov_evolve",
    "edk_krylov_diagnostics", "edk_krylov_accepted_intervals", "edk_krylov_destroy",
    "edk_double_operator_create", "edk_double_matvec", "edk_double_operator_destroy", "edk_double_map",
    "edk_liouvillian_create", "edk_liouvillian_create_csr", "edk_liouvillian_apply", "edk_liouvillian_destroy",
    "edk_lindblad_default_params", "edk_lindblad_create", "edk_lindblad_evolve",
    "edk_lindblad_diagnostics", "edk_lindblad_records", "edk_lindblad_dim", "edk_lindblad_destroy",
]


class CacheStats(C.Structure):
    """edk_cache_stats: diagnostics of the device block cache."""
    _fields_ = [(n, C.c_int64) for n in ("cached_bytes", "owned_blocks", "releases", "released_bytes",
                                          "double_frees", "foreign_frees")]


def cache_info(device=0):
    """The device block cache's counters as a dict (edk_cache_info)."""
    st = CacheStats()
    check(lib().edk_cache_info(device, C.byref(st)))
    return {n: getattr(st, n) for n, _ in CacheStats._fields_}


class EdkError(RuntimeError):
    """Error raised from a nonzero status of the C ABI."""

    def __init__(self, code, msg):
        super().__init__(f"[edk status {code}] {msg}")
        self.code = code


class BasisSpec(C.Structure):
    _fields_ = [("kind", C.c_int32), ("L", C.c_int32), ("base", C.c_int32), ("a", C.c_int32),
                ("N", C.c_int32), ("k", C.c_int32), ("p", C.c_int32), ("z", C.c_int32),
                ("has_k", C.c_int32), ("has_p", C.c_int32), ("has_z", C.c_int32),
                ("ncustom", C.c_int32), ("perms", C.c_void_p), ("ks", C.c_void_p),
                ("invs", C.c_void_p)]


class BatchItem(C.Structure):
    _fields_ = [("op", C.c_void_p), ("x", C.c_void_p), ("y", C.c_void_p), ("nb", C.c_int64),
                ("accumulate", C.c_int32)]


EDK_JOB_GROUND, EDK_JOB_EVOLVE = 0, 1
EDK_REORTH_FULL, EDK_REORTH_PARTIAL = 0, 1
REORTH_CODES = {"full": EDK_REORTH_FULL, "partial": EDK_REORTH_PARTIAL}


class Job(C.Structure):
    _fields_ = [("kind", C.c_int32), ("dtype", C.c_int32), ("op", C.c_void_p), ("psi0", C.c_void_p),
                ("psi0_index", C.c_int64), ("m", C.c_int32), ("refine", C.c_int32), ("tol", C.c_double),
                ("t", C.c_double), ("m_init", C.c_int32), ("m_max", C.c_int32),
                ("want_vector", C.c_int32), ("reorth", C.c_int32)]


class JobResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("steps", C.c_int32), ("E0", C.c_double), ("resid", C.c_double),
                ("matvecs", C.c_int64), ("overlap_re", C.c_double), ("overlap_im", C.c_double),
                ("norm", C.c_double), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_end", C.c_double), ("message", C.c_char * 160)]


class ServerParams(C.Structure):
    _fields_ = [("batch", C.c_int32), ("batch_wait_us", C.c_double), ("affinity", C.c_int32)]


class ServerInfo(C.Structure):
    _fields_ = [("completed", C.c_int64), ("queued", C.c_int64), ("running", C.c_int64),
                ("peak_running", C.c_int64), ("inflight_bytes", C.c_int64),
                ("peak_inflight_bytes", C.c_int64), ("budget_bytes", C.c_int64)]


class KrylovParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("m_init", C.c_int32), ("m_max", C.c_int32),
                ("extend_step", C.c_int32), ("nsample", C.c_int32), ("bisect_iters", C.c_int32),
                ("extend_basis", C.c_int32), ("normalize_output", C.c_int32), ("reorth", C.c_int32)]


class KrylovDiag(C.Structure):
    _fields_ = [("basis_builds", C.c_int64), ("basis_extensions", C.c_int64),
                ("restarts", C.c_int64), ("matvecs", C.c_int64),
                ("total_times_served", C.c_int64), ("max_dim_used", C.c_int64),
                ("n_accepted_intervals", C.c_int64)]


class LindbladParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("m_init", C.c_int32), ("m_max", C.c_int32),
                ("extend_step", C.c_int32), ("nsample", C.c_int32), ("bisect_iters", C.c_int32),
                ("extend_basis", C.c_int32), ("normalize_trace", C.c_int32), ("record_min_eig", C.c_int32)]


class LindbladDiag(C.Structure):
    _fields_ = [("basis_builds", C.c_int64), ("basis_extensions", C.c_int64),
                ("restarts", C.c_int64), ("liouvillian_applies", C.c_int64),
                ("total_times_served", C.c_int64), ("max_dim_used", C.c_int64),
                ("n_accepted_intervals", C.c_int64), ("max_trace_drift", C.c_double),
                ("max_hermiticity_drift", C.c_double), ("min_min_hermitian_eig", C.c_double)]


_LIB = None


def lib():
    """Load the library once and declare the signatures."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libedkit_hip.so not found at {LIB_PATH}: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, D, SZ = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
    PP = C.POINTER(C.c_void_p)
    sig = {
        "edk_last_error": (C.c_char_p, []),
        "edk_version": (I32, []),
        "edk_ctx_create": (I32, [I32, P, PP]),
        "edk_ctx_destroy": (I32, [P]),
        "edk_ctx_synchronize": (I32, [P]),
        "edk_ctx_stream": (P, [P]),
        "edk_malloc": (I32, [P, SZ, PP]),
        "edk_free": (I32, [P, P]),
        "edk_memcpy_htod": (I32, [P, P, P, SZ]),
        "edk_memcpy_dtoh": (I32, [P, P, P, SZ]),
        "edk_memset": (I32, [P, P, I32, SZ]),
        "edk_cache_info": (I32, [I32, C.POINTER(CacheStats)]),
        "edk_basis_build": (I32, [P, C.POINTER(BasisSpec), PP]),
        "edk_basis_upload": (I32, [P, C.POINTER(BasisSpec), P, P, I64, PP]),
        "edk_basis_dim": (I64, [P]),
        "edk_basis_order": (I32, [P]),
        "edk_basis_download": (I32, [P, P, P]),
        "edk_basis_index": (I32, [P, P, I64, P, P, P]),
        "edk_basis_destroy": (I32, [P]),
        "edk_operator_create": (I32, [P, P, I64, P, P, P, P, P, P, P, PP]),
        "edk_operator_destroy": (I32, [P]),
        "edk_operator_count_offdiag": (I32, [P, C.POINTER(C.c_int64)]),
        "edk_matvec": (I32, [P, I32, P, P, I64, I32]),
        "edk_matvec_host": (I32, [P, I32, P, P, I64, I32]),
        "edk_matvec_multi": (I32, [P, I32, I32, P, P]),
        "edk_operator_assemble": (I32, [P, C.POINTER(C.c_int64)]),
        "edk_operator_download_csr": (I32, [P, P, P, P]),
        "edk_operator_release_csr": (I32, [P]),
        "edk_operator_set_variant": (I32, [P, I32, I64, I32]),
        "edk_operator_autotune": (I32, [P, I32, I64, I32, C.POINTER(C.c_int), C.POINTER(D)]),
        "edk_batch_create": (I32, [P, I32, C.POINTER(BatchItem), I64, I32, PP]),
        "edk_batch_create_ex": (I32, [P, I32, C.POINTER(BatchItem), I64, I32, I32, PP]),
        "edk_batch_launches": (I64, [P]),
        "edk_batch_launch": (I32, [P]),
        "edk_batch_destroy": (I32, [P]),
        "edk_server_create": (I32, [I32, I32, SZ, PP]),
        "edk_server_default_params": (None, [C.POINTER(ServerParams)]),
        "edk_server_create_ex": (I32, [I32, I32, SZ, C.POINTER(ServerParams), PP]),
        "edk_server_batch_stats": (I32, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(D)]),
        "edk_server_submit": (I32, [P, C.POINTER(Job), C.POINTER(C.c_int64)]),
        "edk_server_wait": (I32, [P, I64, D, C.POINTER(JobResult), P]),
        "edk_server_clock": (D, [P]),
        "edk_server_stats": (I32, [P, C.POINTER(ServerInfo)]),
        "edk_server_destroy": (I32, [P]),
        "edk_vec_scale": (I32, [P, I32, P, P, D, I64]),
        "edk_vec_axpy": (I32, [P, I32, P, P, D, D, I64]),
        "edk_vec_dot": (I32, [P, I32, P, P, I64, C.POINTER(D), C.POINTER(D)]),
        "edk_vec_nrm2": (I32, [P, I32, P, I64, C.POINTER(D)]),
        "edk_vec_dot_noalloc": (I32, [P, I32, P, P, I64, C.POINTER(D), C.POINTER(D)]),
        "edk_lanczos_ground": (I32, [P, I32, P, I32, D, I32, C.POINTER(D), C.POINTER(D),
                                     C.POINTER(C.c_int), P]),
        "edk_lanczos_ground_ex": (I32, [P, I32, P, I32, D, I32, I32, C.POINTER(D), C.POINTER(D),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int), P]),
        "edk_schmidt": (I32, [P, I32, P, I32, P, P]),
        "edk_ent_spec": (I32, [P, I32, P, I32, P, P, C.POINTER(I64)]),
        "edk_krylov_default_params": (None, [C.POINTER(KrylovParams)]),
        "edk_krylov_create": (I32, [P, I32, P, C.POINTER(KrylovParams), PP]),
        "edk_krylov_evolve": (I32, [P, D, P]),
        "edk_krylov_diagnostics": (I32, [P, C.POINTER(KrylovDiag)]),
        "edk_krylov_accepted_intervals": (I32, [P, P, I64, C.POINTER(C.c_int64)]),
        "edk_krylov_destroy": (I32, [P]),
        "edk_double_operator_create": (I32, [P, P, P, I64, P, P, P, P, P, P, P, PP]),
        "edk_double_matvec": (I32, [P, I32, P, P, I64, I32]),
        "edk_double_operator_destroy": (I32, [P]),
        "edk_double_map": (I32, [P, P, I32, P, P, I64]),
        "edk_liouvillian_create": (I32, [P, I64, P, I32, P, PP]),
        "edk_liouvillian_create_csr": (I32, [P, I64, P, P, P, I32, P, P, P, PP]),
        "edk_liouvillian_apply": (I32, [P, P, P]),
        "edk_liouvillian_destroy": (I32, [P]),
        "edk_lindblad_default_params": (None, [C.POINTER(LindbladParams)]),
        "edk_lindblad_create": (I32, [P, P, C.POINTER(LindbladParams), PP]),
        "edk_lindblad_evolve": (I32, [P, D, P]),
        "edk_lindblad_diagnostics": (I32, [P, C.POINTER(LindbladDiag)]),
        "edk_lindblad_records": (I32, [P, P, P, P, P, I64, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64)]),
        "edk_lindblad_dim": (I32, [P, C.POINTER(C.c_int)]),
        "edk_lindblad_destroy": (I32, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(rc):
    """Raise EdkError for a nonzero status."""
    if rc != EDK_OK:
        raise EdkError(rc, lib().edk_last_error().decode())
    return rc
