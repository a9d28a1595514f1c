"""rocblas_amd -- MI355X-native triangular solve with rocBLAS's trsm C ABI.

The product is ``librocblas_amd.so`` (built from ``csrc/`` by ``build.py``);
this module binds its C entry points with ctypes under their rocBLAS names
(``rocblas_dtrsm``, ``rocblas_dtrsm_strided_batched``, ...), with the same
argument order, enum values and status codes as
library/include/internal/rocblas-functions.h, so Python callers and tests read
like the reference's own clients.  Device memory and streams come from
PyTorch (plumbing only); pointers cross the ABI as plain integers.

There is no fallback: if the shared library is missing, loading raises.
"""
import ctypes
import enum
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librocblas_amd.so")


class rocblas_status(enum.IntEnum):
    success = 0
    invalid_handle = 1
    not_implemented = 2
    invalid_pointer = 3
    invalid_size = 4
    memory_error = 5
    internal_error = 6
    perf_degraded = 7
    size_query_mismatch = 8
    size_increased = 9
    size_unchanged = 10
    invalid_value = 11
    continue_ = 12
    check_numerics_fail = 13
    excluded_from_build = 14
    arch_mismatch = 15


class rocblas_operation(enum.IntEnum):
    none = 111
    transpose = 112
    conjugate_transpose = 113


class rocblas_fill(enum.IntEnum):
    upper = 121
    lower = 122
    full = 123


class rocblas_diagonal(enum.IntEnum):
    non_unit = 131
    unit = 132


class rocblas_side(enum.IntEnum):
    left = 141
    right = 142
    both = 143


class rocblas_pointer_mode(enum.IntEnum):
    host = 0
    device = 1


SIDE = {"L": rocblas_side.left, "R": rocblas_side.right}
FILL = {"U": rocblas_fill.upper, "L": rocblas_fill.lower}
OP = {"N": rocblas_operation.none, "T": rocblas_operation.transpose,
      "C": rocblas_operation.conjugate_transpose}
DIAG = {"N": rocblas_diagonal.non_unit, "U": rocblas_diagonal.unit}


class RocblasError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = rocblas_status(status)
        super().__init__(f"{what}: {self.status.name}")


_lib = None
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_i64 = ctypes.c_int64

# (name, restype, argtypes) for every symbol declared in include/rocblas_amd.h
_TRSM_PLAIN = [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _vp, _i32]
_TRSM_BATCHED = [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _vp, _i32, _i32]
_TRTRI_PLAIN = [_vp, _i32, _i32, _i32, _vp, _i32, _vp, _i32]
_TRTRI_BATCHED = [_vp, _i32, _i32, _i32, _vp, _i32, _vp, _i32, _i32]
_TRTRI_STRIDED = [_vp, _i32, _i32, _i32, _vp, _i32, _i64, _vp, _i32, _i64, _i32]
_TRSM_VBATCHED = [_vp, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32]
_TRSM_STRIDED = [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _i64, _vp, _i32, _i64,
                 _i32]
SYMBOLS = [
    ("rocblas_create_handle", _i32, [ctypes.POINTER(_vp)]),
    ("rocblas_destroy_handle", _i32, [_vp]),
    ("rocblas_set_stream", _i32, [_vp, _vp]),
    ("rocblas_get_stream", _i32, [_vp, ctypes.POINTER(_vp)]),
    ("rocblas_set_pointer_mode", _i32, [_vp, _i32]),
    ("rocblas_get_pointer_mode", _i32, [_vp, ctypes.POINTER(_i32)]),
    ("rocblas_status_to_string", ctypes.c_char_p, [_i32]),
    ("rocblas_start_device_memory_size_query", _i32, [_vp]),
    ("rocblas_stop_device_memory_size_query", _i32, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    ("rocblas_is_device_memory_size_query", ctypes.c_bool, [_vp]),
    ("rocblas_get_device_memory_size", _i32, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
    ("rocblas_set_device_memory_size", _i32, [_vp, ctypes.c_size_t]),
    ("rocblas_set_workspace", _i32, [_vp, _vp, ctypes.c_size_t]),
    ("rocblas_is_managing_device_memory", ctypes.c_bool, [_vp]),
    ("rocblas_is_user_managing_device_memory", ctypes.c_bool, [_vp]),
    ("rocblas_strsm", _i32, _TRSM_PLAIN),
    ("rocblas_dtrsm", _i32, _TRSM_PLAIN),
    ("rocblas_strsm_batched", _i32, _TRSM_BATCHED),
    ("rocblas_dtrsm_batched", _i32, _TRSM_BATCHED),
    ("rocblas_strsm_strided_batched", _i32, _TRSM_STRIDED),
    ("rocblas_dtrsm_strided_batched", _i32, _TRSM_STRIDED),
    ("rocblas_strtri", _i32, _TRTRI_PLAIN),
    ("rocblas_dtrtri", _i32, _TRTRI_PLAIN),
    ("rocblas_strtri_batched", _i32, _TRTRI_BATCHED),
    ("rocblas_dtrtri_batched", _i32, _TRTRI_BATCHED),
    ("rocblas_strtri_strided_batched", _i32, _TRTRI_STRIDED),
    ("rocblas_dtrtri_strided_batched", _i32, _TRTRI_STRIDED),
    ("rocblas_strsm_64", _i32, [_vp, _i32, _i32, _i32, _i32, _i64, _i64, _vp, _vp, _i64, _vp, _i64]),
    ("rocblas_strsm_batched_64", _i32, [_vp, _i32, _i32, _i32, _i32, _i64, _i64, _vp, _vp, _i64, _vp,
                                        _i64, _i64]),
    ("rocblas_strsm_strided_batched_64", _i32, [_vp, _i32, _i32, _i32, _i32, _i64, _i64, _vp, _vp,
                                                _i64, _i64, _vp, _i64, _i64, _i64]),
    ("rocblas_dtrsm_64", _i32, [_vp, _i32, _i32, _i32, _i32, _i64, _i64, _vp, _vp, _i64, _vp, _i64]),
    ("rocblas_dtrsm_batched_64", _i32, [_vp, _i32, _i32, _i32, _i32, _i64, _i64, _vp, _vp, _i64, _vp,
                                        _i64, _i64]),
    ("rocblas_dtrsm_strided_batched_64", _i32, [_vp, _i32, _i32, _i32, _i32, _i64, _i64, _vp, _vp,
                                                _i64, _i64, _vp, _i64, _i64, _i64]),
    ("rocblas_trsm_ex", _i32, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _vp, _i32,
                               _vp, _i32, _i32]),
    ("rocblas_trsm_batched_ex", _i32, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32,
                                       _vp, _i32, _i32, _vp, _i32, _i32]),
    ("rocblas_trsm_strided_batched_ex", _i32, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp,
                                               _i32, ctypes.c_int64, _vp, _i32, ctypes.c_int64,
                                               _i32, _vp, _i32, ctypes.c_int64, _i32]),
    # include/rocblas_amd_ext.h
    ("rocblas_amd_strsm_vbatched", _i32, _TRSM_VBATCHED),
    ("rocblas_amd_dtrsm_vbatched", _i32, _TRSM_VBATCHED),
    ("rocblas_amd_dtrsm_vbatched_noalloc", _i32, _TRSM_VBATCHED),
    ("rocblas_amd_vbatched_set_tuning", _i32, [_i32]),
    ("rocblas_amd_vbatched_set_bucket_tuning", _i32, [_i32, _i32]),
    ("rocblas_amd_sgemm", _i32, [_vp, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _vp, _i32, _vp,
                                 _vp, _i32]),
    ("rocblas_amd_dgemm", _i32, [_vp, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _vp, _i32, _vp,
                                 _vp, _i32]),
    ("rocblas_amd_dtrsm_mixed", _i32, [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32,
                                       _vp, _i32, _i32, _i32, _i32, ctypes.c_double, _vp]),
    ("rocblas_amd_server_create", _i32, [_vp, ctypes.c_size_t, _i32, _i32, _i32, _i32]),
    ("rocblas_amd_server_destroy", _i32, [_vp]),
    ("rocblas_amd_server_run", _i32, [_vp, _vp, ctypes.c_int64, _vp, _vp, _vp]),
    ("rocblas_amd_server_result", _i32, [_vp, ctypes.c_int64, _vp, _i32]),
    ("rocblas_amd_dsynth_tri", _i32, [_vp, _i32, _i32, _i32, ctypes.c_uint64, ctypes.c_double,
                                      ctypes.c_double, _vp, _i32]),
    ("rocblas_amd_ssynth_tri", _i32, [_vp, _i32, _i32, _i32, ctypes.c_uint64, ctypes.c_double,
                                      ctypes.c_double, _vp, _i32]),
    ("rocblas_amd_dsynth_rhs", _i32, [_vp, _i32, _i32, ctypes.c_uint64, _i32, _vp, _i32]),
    ("rocblas_amd_ssynth_rhs", _i32, [_vp, _i32, _i32, ctypes.c_uint64, _i32, _vp, _i32]),
    ("rocblas_amd_get_async_status", _i32, [_vp]),
    ("rocblas_amd_set_chain_test_timeout", _i32, [_vp, _i32]),
    ("rocblas_amd_hbm_stream", _i32, [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, _i32,
                                      ctypes.POINTER(ctypes.c_double)]),
    ("rocblas_amd_hbm_stream_step_bytes", _i32, [_vp, ctypes.POINTER(ctypes.c_size_t)]),
]


def lib():
    """Load librocblas_amd.so (torch is imported first so both share one HIP runtime)."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401  -- binds libamdhip64 before our library loads
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m rocblas_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SYMBOLS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_to_string(status):
    return lib().rocblas_status_to_string(int(status)).decode()


def check(status, what="rocblas call"):
    if status != rocblas_status.success:
        raise RocblasError(status, what)
    return status


class Handle:
    """rocblas_handle owner (create on the current device, destroy on close)."""

    def __init__(self):
        h = _vp()
        check(lib().rocblas_create_handle(ctypes.byref(h)), "rocblas_create_handle")
        self.ptr = h.value

    def close(self):
        if self.ptr:
            lib().rocblas_destroy_handle(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        return lib().rocblas_set_stream(self.ptr, stream_ptr)

    def set_pointer_mode(self, mode):
        return lib().rocblas_set_pointer_mode(self.ptr, int(mode))

    def get_async_status(self):
        """rocblas_amd_get_async_status: completes the stream, reports (and clears)
        a wait that a kernel of an earlier call gave up."""
        return rocblas_status(lib().rocblas_amd_get_async_status(self.ptr))

    def set_chain_test_timeout(self, enable):
        return rocblas_status(lib().rocblas_amd_set_chain_test_timeout(self.ptr, int(enable)))


def trsm(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, precision="d",
         alpha_ptr=None):
    """rocblas_{s,d}trsm on raw device pointers (ints).  ``alpha`` is a host
    scalar unless ``alpha_ptr`` (a device pointer) is given.  Returns the status."""
    fn = lib().rocblas_dtrsm if precision == "d" else lib().rocblas_strsm
    a = alpha_ptr
    keep = None
    if a is None and alpha is not None:
        keep = ctypes.c_double(alpha) if precision == "d" else ctypes.c_float(alpha)
        a = ctypes.addressof(keep)
    st = fn(_h(handle), int(SIDE.get(side, side)), int(FILL.get(uplo, uplo)),
            int(OP.get(trans, trans)), int(DIAG.get(diag, diag)), m, n, a, A, lda, B, ldb)
    del keep
    return rocblas_status(st)


def trsm_strided_batched(handle, side, uplo, trans, diag, m, n, alpha, A, lda, stride_a, B, ldb,
                         stride_b, batch, precision="d", alpha_ptr=None):
    fn = (lib().rocblas_dtrsm_strided_batched if precision == "d"
          else lib().rocblas_strsm_strided_batched)
    a = alpha_ptr
    keep = None
    if a is None and alpha is not None:
        keep = ctypes.c_double(alpha) if precision == "d" else ctypes.c_float(alpha)
        a = ctypes.addressof(keep)
    st = fn(_h(handle), int(SIDE.get(side, side)), int(FILL.get(uplo, uplo)),
            int(OP.get(trans, trans)), int(DIAG.get(diag, diag)), m, n, a, A, lda, stride_a, B,
            ldb, stride_b, batch)
    del keep
    return rocblas_status(st)


def trsm_batched(handle, side, uplo, trans, diag, m, n, alpha, A_array, lda, B_array, ldb, batch,
                 precision="d", alpha_ptr=None):
    """``A_array``/``B_array`` are device pointers to device arrays of pointers."""
    fn = lib().rocblas_dtrsm_batched if precision == "d" else lib().rocblas_strsm_batched
    a = alpha_ptr
    keep = None
    if a is None and alpha is not None:
        keep = ctypes.c_double(alpha) if precision == "d" else ctypes.c_float(alpha)
        a = ctypes.addressof(keep)
    st = fn(_h(handle), int(SIDE.get(side, side)), int(FILL.get(uplo, uplo)),
            int(OP.get(trans, trans)), int(DIAG.get(diag, diag)), m, n, a, A_array, lda, B_array,
            ldb, batch)
    del keep
    return rocblas_status(st)


def trtri(handle, uplo, diag, n, A, lda, invA, ldinvA, precision="d"):
    """rocblas_{s,d}trtri on raw device pointers; returns the status."""
    fn = lib().rocblas_dtrtri if precision == "d" else lib().rocblas_strtri
    return rocblas_status(fn(_h(handle), int(FILL.get(uplo, uplo)), int(DIAG.get(diag, diag)), n,
                             A, lda, invA, ldinvA))


def trtri_strided_batched(handle, uplo, diag, n, A, lda, stride_a, invA, ldinvA, stride_invA,
                          batch, precision="d"):
    fn = (lib().rocblas_dtrtri_strided_batched if precision == "d"
          else lib().rocblas_strtri_strided_batched)
    return rocblas_status(fn(_h(handle), int(FILL.get(uplo, uplo)), int(DIAG.get(diag, diag)), n,
                             A, lda, stride_a, invA, ldinvA, stride_invA, batch))


def trtri_batched(handle, uplo, diag, n, A_array, lda, invA_array, ldinvA, batch, precision="d"):
    fn = lib().rocblas_dtrtri_batched if precision == "d" else lib().rocblas_strtri_batched
    return rocblas_status(fn(_h(handle), int(FILL.get(uplo, uplo)), int(DIAG.get(diag, diag)), n,
                             A_array, lda, invA_array, ldinvA, batch))


def trsm_vbatched(handle, side, uplo, trans, diag, m_arr, n_arr, alpha, A_array, lda_arr, B_array,
                  ldb_arr, batch, precision="d", alpha_ptr=None):
    """rocblas_amd_{s,d}trsm_vbatched: every size/pointer argument is a device pointer."""
    fn = lib().rocblas_amd_dtrsm_vbatched if precision == "d" else lib().rocblas_amd_strsm_vbatched
    a = alpha_ptr
    keep = None
    if a is None and alpha is not None:
        keep = ctypes.c_double(alpha) if precision == "d" else ctypes.c_float(alpha)
        a = ctypes.addressof(keep)
    st = fn(_h(handle), int(SIDE.get(side, side)), int(FILL.get(uplo, uplo)),
            int(OP.get(trans, trans)), int(DIAG.get(diag, diag)), m_arr, n_arr, a, A_array,
            lda_arr, B_array, ldb_arr, batch)
    del keep
    return rocblas_status(st)


def trsm_vbatched_noalloc(handle, side, uplo, trans, diag, m_arr, n_arr, alpha, A_array, lda_arr,
                          B_array, ldb_arr, batch, alpha_ptr=None):
    """rocblas_amd_dtrsm_vbatched_noalloc: rocblas_amd_dtrsm_vbatched with its
    planning scratch in a library-owned device array (batch <= 65536)."""
    a = alpha_ptr
    keep = None
    if a is None and alpha is not None:
        keep = ctypes.c_double(alpha)
        a = ctypes.addressof(keep)
    st = lib().rocblas_amd_dtrsm_vbatched_noalloc(
        _h(handle), int(SIDE.get(side, side)), int(FILL.get(uplo, uplo)),
        int(OP.get(trans, trans)), int(DIAG.get(diag, diag)), m_arr, n_arr, a, A_array, lda_arr,
        B_array, ldb_arr, batch)
    del keep
    return rocblas_status(st)


def gemm(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, precision="d"):
    """rocblas_amd_{s,d}gemm with host scalars."""
    ct = ctypes.c_double if precision == "d" else ctypes.c_float
    a, b = ct(alpha), ct(beta)
    fn = lib().rocblas_amd_dgemm if precision == "d" else lib().rocblas_amd_sgemm
    return rocblas_status(fn(_h(handle), int(OP.get(transA, transA)), int(OP.get(transB, transB)),
                             m, n, k, ctypes.addressof(a), A, lda, B, ldb, ctypes.addressof(b), C,
                             ldc))


class RefineInfo(ctypes.Structure):
    """rocblas_amd_refine_info"""
    _fields_ = [("iterations", ctypes.c_int32), ("backward_error", ctypes.c_double),
                ("converged", ctypes.c_int32)]


MIXED_FORMS = {"fp32": 0, "bf16x3": 1}


def trsm_mixed(handle, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, block=0,
               max_refine=10, tol=1e-13, alpha_ptr=None, form="fp32"):
    """rocblas_amd_dtrsm_mixed (FP32 or split-BF16x3 trailing updates + FP64
    refinement).  Returns (status, RefineInfo)."""
    info = RefineInfo()
    keep = None
    a = alpha_ptr
    if a is None and alpha is not None:
        keep = ctypes.c_double(alpha)
        a = ctypes.addressof(keep)
    st = lib().rocblas_amd_dtrsm_mixed(_h(handle), int(SIDE.get(side, side)),
                                       int(FILL.get(uplo, uplo)), int(OP.get(trans, trans)),
                                       int(DIAG.get(diag, diag)), m, n, a, A, lda, B, ldb,
                                       MIXED_FORMS.get(form, form), block, max_refine, tol,
                                       ctypes.byref(info))
    del keep
    return rocblas_status(st), info


def synth_tri(handle, uplo, dist, k, seed, A, lda, scale=1.0, poison=0.0, precision="d"):
    """Device-side synthetic triangle (dist "uniform" or "normal") into A."""
    fn = lib().rocblas_amd_dsynth_tri if precision == "d" else lib().rocblas_amd_ssynth_tri
    return rocblas_status(fn(_h(handle), int(FILL.get(uplo, uplo)),
                             {"uniform": 0, "normal": 1}.get(dist, dist), k, seed, scale, poison,
                             A, lda))


def synth_rhs(handle, m, n, seed, B, ldb, tag=2, precision="d"):
    """Device-side U[-1,1] right-hand side into B."""
    fn = lib().rocblas_amd_dsynth_rhs if precision == "d" else lib().rocblas_amd_ssynth_rhs
    return rocblas_status(fn(_h(handle), m, n, seed, tag, B, ldb))


def hbm_stream_step_bytes(handle):
    """Bytes one access unit of rocblas_amd_hbm_stream moves."""
    b = ctypes.c_size_t()
    check(lib().rocblas_amd_hbm_stream_step_bytes(_h(handle), ctypes.byref(b)),
          "rocblas_amd_hbm_stream_step_bytes")
    return b.value


def hbm_stream(handle, src, read_bytes, dst, write_bytes, iters=10):
    """rocblas_amd_hbm_stream: seconds per pass of the streaming probe that
    reads read_bytes from src and writes write_bytes to dst (device pointers)."""
    sec = ctypes.c_double()
    check(lib().rocblas_amd_hbm_stream(_h(handle), src, read_bytes, dst, write_bytes, iters,
                                       ctypes.byref(sec)), "rocblas_amd_hbm_stream")
    return sec.value


def _h(handle):
    if handle is None:
        return None
    return handle.ptr if isinstance(handle, Handle) else handle
