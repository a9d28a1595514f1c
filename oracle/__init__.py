"""CPU oracle for the triangular-solve path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this package, and only as the checker.  The product package
(``rocblas_amd``) never imports it.

The numerical routines live in ``trsm_oracle.c`` (built to ``liboracle.so`` by
``oracle/Makefile``); this module wraps them for numpy arrays.  Matrices are
passed as numpy arrays of shape ``(ld, cols)`` in Fortran (column-major) order,
which is the rocBLAS storage convention: element (i, j) at ``i + j*ld``.

What each routine restates is listed in the C file's header (reference
file:line citations there).  Pinning of each routine is done in
``tests/test_oracle.py`` against exact-rational golden fixtures
(``tests/golden/``), scipy's LAPACK/BLAS, and closed-form special cases.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SIDE = {"L": 141, "R": 142}
FILL = {"U": 121, "L": 122}
OP = {"N": 111, "T": 112, "C": 113}
DIAG = {"N": 131, "U": 132}

_lib = None


def build():
    """Compile liboracle.so from trsm_oracle.c (gcc, no GPU needed)."""
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        i = ctypes.c_int
        ll = ctypes.c_longlong
        L.oracle_trsm_d.argtypes = [i, i, i, i, i, i, ctypes.c_double, dp, i, dp, i]
        L.oracle_trsm_d.restype = None
        L.oracle_trsm_s.argtypes = [i, i, i, i, i, i, ctypes.c_float, fp, i, fp, i]
        L.oracle_trsm_s.restype = None
        L.oracle_trmm_d.argtypes = [i, i, i, i, i, i, ctypes.c_double, dp, i, dp, i]
        L.oracle_trmm_d.restype = None
        L.oracle_trsm_d_batched_omp.argtypes = [i, i, i, i, i, i, ctypes.c_double, dp, i, ll,
                                                dp, i, ll, i, i]
        L.oracle_trsm_d_batched_omp.restype = None
        L.oracle_trsm_d_omp.argtypes = [i, i, i, i, i, i, ctypes.c_double, dp, i, dp, i, i]
        L.oracle_trsm_d_omp.restype = None
        L.oracle_num_threads.argtypes = []
        L.oracle_num_threads.restype = i
        L.oracle_backward_error_d.argtypes = [i, i, i, i, i, i, ctypes.c_double, dp, i, dp, i,
                                              dp, i]
        L.oracle_backward_error_d.restype = ctypes.c_double
        L.oracle_backward_error_s.argtypes = [i, i, i, i, i, i, ctypes.c_float, fp, i, fp, i,
                                              fp, i]
        L.oracle_backward_error_s.restype = ctypes.c_double
        L.oracle_norm1_rel_diff_d.argtypes = [i, i, dp, i, dp, i]
        L.oracle_norm1_rel_diff_d.restype = ctypes.c_double
        L.oracle_splitmix_key.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_splitmix_key.restype = ctypes.c_uint64
        L.oracle_splitmix_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t,
                                              ctypes.c_double, ctypes.c_double, dp]
        L.oracle_splitmix_uniform.restype = None
        L.oracle_trtri_d.argtypes = [i, i, i, dp, i]
        L.oracle_trtri_d.restype = None
        L.oracle_trtri_s.argtypes = [i, i, i, fp, i]
        L.oracle_trtri_s.restype = None
        L.oracle_cblas_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]
        L.oracle_cblas_load.restype = i
        L.oracle_cblas_dtrsm.argtypes = [i, i, i, i, i, i, ctypes.c_double, dp, i, dp, i, i]
        L.oracle_cblas_dtrsm.restype = i
        L.oracle_cblas_dtrsm_batched_omp.argtypes = [i, i, i, i, i, i, ctypes.c_double, dp, i,
                                                     ll, dp, i, ll, i, i]
        L.oracle_cblas_dtrsm_batched_omp.restype = i
        _lib = L
    return _lib


def _ptr(a):
    if a.dtype == np.float64:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    if a.dtype == np.float32:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    raise TypeError(a.dtype)


def _check_colmajor(a, name):
    if not a.flags.f_contiguous or a.ndim != 2:
        raise ValueError(f"{name} must be a 2-D Fortran-ordered array (ld, cols)")


def _enums(side, uplo, trans, diag):
    return SIDE[side], FILL[uplo], OP[trans], DIAG[diag]


def trsm(side, uplo, trans, diag, m, n, alpha, A, B):
    """In-place B := solution of op(A) X = alpha B (side L) / X op(A) = alpha B (side R).

    Netlib xTRSM loop nest (trsm_oracle.c).  A and B are Fortran-ordered
    ``(ld, cols)`` arrays of the same dtype (float64 or float32)."""
    _check_colmajor(A, "A")
    _check_colmajor(B, "B")
    s, u, t, d = _enums(side, uplo, trans, diag)
    if B.dtype == np.float64:
        lib().oracle_trsm_d(s, u, t, d, m, n, float(alpha), _ptr(A), A.shape[0], _ptr(B),
                            B.shape[0])
    else:
        lib().oracle_trsm_s(s, u, t, d, m, n, float(alpha), _ptr(A), A.shape[0], _ptr(B),
                            B.shape[0])
    return B


def trtri(uplo, diag, n, A):
    """In-place inverse of the ``uplo`` triangle of A (LAPACK xTRTI2 loop nest);
    the other triangle is left untouched, as in LAPACK."""
    _check_colmajor(A, "A")
    fn = lib().oracle_trtri_d if A.dtype == np.float64 else lib().oracle_trtri_s
    fn(FILL[uplo], DIAG[diag], n, _ptr(A), A.shape[0])
    return A


def trmm(side, uplo, trans, diag, m, n, alpha, A, B):
    """In-place B := alpha op(A) B (side L) / alpha B op(A) (side R), long-double accumulate."""
    _check_colmajor(A, "A")
    _check_colmajor(B, "B")
    s, u, t, d = _enums(side, uplo, trans, diag)
    lib().oracle_trmm_d(s, u, t, d, m, n, float(alpha), _ptr(A), A.shape[0], _ptr(B),
                        B.shape[0])
    return B


def trsm_batched_omp(side, uplo, trans, diag, m, n, alpha, A, lda, stride_a, B, ldb, stride_b,
                     batch, threads=0):
    """Strided-batch netlib solve, OpenMP-parallel over instances (flat float64 buffers)."""
    s, u, t, d = _enums(side, uplo, trans, diag)
    lib().oracle_trsm_d_batched_omp(s, u, t, d, m, n, float(alpha), _ptr(A), lda, stride_a,
                                    _ptr(B), ldb, stride_b, batch, threads)
    return B


def trsm_omp(side, uplo, trans, diag, m, n, alpha, A, B, threads=0):
    """Single solve, OpenMP-parallel over independent right-hand sides."""
    _check_colmajor(A, "A")
    _check_colmajor(B, "B")
    s, u, t, d = _enums(side, uplo, trans, diag)
    lib().oracle_trsm_d_omp(s, u, t, d, m, n, float(alpha), _ptr(A), A.shape[0], _ptr(B),
                            B.shape[0], threads)
    return B


def num_threads():
    return lib().oracle_num_threads()


# --------------------------------------------------- host CBLAS (timing) --
_cblas_name = None


def cblas_load():
    """Load a host CBLAS for the second CPU baseline: the OpenBLAS that ships
    with scipy (``scipy_cblas_dtrsm``), else a system ``libcblas``/``libopenblas``.
    Returns a description of what was loaded, or None."""
    global _cblas_name
    if _cblas_name is not None:
        return _cblas_name
    import glob
    cands = []
    try:
        import scipy
        libs = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        for p in sorted(glob.glob(os.path.join(libs, "libscipy_openblas*.so"))):
            cands.append((p, "scipy_cblas_dtrsm", "scipy_openblas_set_num_threads"))
    except ImportError:
        pass
    for p in ("libopenblas.so.0", "libcblas.so.3", "libblas.so.3"):
        cands.append((p, "cblas_dtrsm", "openblas_set_num_threads"))
    for path, sym, thr in cands:
        if lib().oracle_cblas_load(path.encode(), sym.encode(), thr.encode()) == 0:
            _cblas_name = f"{os.path.basename(path)}:{sym}"
            return _cblas_name
    return None


def cblas_trsm(side, uplo, trans, diag, m, n, alpha, A, B, threads=0):
    """In-place solve through the loaded host CBLAS (float64, column-major)."""
    _check_colmajor(A, "A")
    _check_colmajor(B, "B")
    if cblas_load() is None:
        raise RuntimeError("no host CBLAS found")
    s, u, t, d = _enums(side, uplo, trans, diag)
    if lib().oracle_cblas_dtrsm(s, u, t, d, m, n, float(alpha), _ptr(A), A.shape[0], _ptr(B),
                                B.shape[0], threads):
        raise RuntimeError("host CBLAS call failed")
    return B


def cblas_trsm_batched_omp(side, uplo, trans, diag, m, n, alpha, A, lda, stride_a, B, ldb,
                           stride_b, batch, threads=0):
    """Strided batch through the host CBLAS, OpenMP over instances."""
    if cblas_load() is None:
        raise RuntimeError("no host CBLAS found")
    s, u, t, d = _enums(side, uplo, trans, diag)
    if lib().oracle_cblas_dtrsm_batched_omp(s, u, t, d, m, n, float(alpha), _ptr(A), lda,
                                            stride_a, _ptr(B), ldb, stride_b, batch, threads):
        raise RuntimeError("host CBLAS call failed")
    return B


def backward_error(side, uplo, trans, diag, m, n, alpha, A, B0, X):
    """eta = ||alpha B0 - op(A) X||_1 / (||op(A)||_1 ||X||_1 + |alpha| ||B0||_1) (long double)."""
    for a, nm in ((A, "A"), (B0, "B0"), (X, "X")):
        _check_colmajor(a, nm)
    s, u, t, d = _enums(side, uplo, trans, diag)
    if X.dtype == np.float64:
        return lib().oracle_backward_error_d(s, u, t, d, m, n, float(alpha), _ptr(A), A.shape[0],
                                             _ptr(B0), B0.shape[0], _ptr(X), X.shape[0])
    return lib().oracle_backward_error_s(s, u, t, d, m, n, float(alpha), _ptr(A), A.shape[0],
                                         _ptr(B0), B0.shape[0], _ptr(X), X.shape[0])


def norm1_rel_diff(m, n, X, Y):
    """||X - Y||_1 / ||Y||_1 over the leading m x n blocks (float64)."""
    _check_colmajor(X, "X")
    _check_colmajor(Y, "Y")
    return lib().oracle_norm1_rel_diff_d(m, n, _ptr(X), X.shape[0], _ptr(Y), Y.shape[0])


def splitmix_key(seed, tag):
    return int(lib().oracle_splitmix_key(seed, tag))


def splitmix_uniform(key, offset, count, lo=0.0, hi=1.0):
    out = np.empty(count, dtype=np.float64)
    lib().oracle_splitmix_uniform(key, offset, count, lo, hi, _ptr(out))
    return out


def diag_dominant_triangular(k, uplo, rng, lda=None, dtype=np.float64, poison=np.nan):
    """Restatement of the reference test generator
    ``rocblas_client_diagonally_dominant_triangular_matrix``
    (clients/include/rocblas_init.hpp:263-327): random off-diagonal entries in the
    uplo triangle, diagonal = 1.01 * (|row| + |col| off-diagonal sums), or 1 when
    those sums are zero.  The other triangle is filled with ``poison`` so a solver
    that reads it produces NaN.  ``rng`` is a numpy Generator (the reference's
    own values come from std::mt19937 and are not reproduced bit-for-bit)."""
    lda = k if lda is None else lda
    A = np.full((lda, k), poison, dtype=dtype, order="F")
    vals = rng.integers(1, 11, size=(k, k)).astype(dtype) * rng.choice([-1.0, 1.0], size=(k, k))
    # strict uplo triangle (the entries are small integers, so every sum
    # below is exact whatever its order)
    strict = np.triu(np.ones((k, k), dtype=bool), 1) if uplo == "U" else \
        np.tril(np.ones((k, k), dtype=bool), -1)
    A[:k, :k] = np.where(strict, vals, A[:k, :k])
    T = np.where(strict, A[:k, :k], dtype(0))  # in A's precision
    s = np.abs(T).sum(axis=1) + np.abs(T).sum(axis=0)  # row i and column i of the triangle
    A[np.arange(k), np.arange(k)] = np.where(s == 0, dtype(1), (s * 1.01).astype(dtype))
    return A


def make_unit_diagonal(A, uplo, k, poison=np.nan):
    """Restatement of the reference's ``make_unit_diagonal``
    (clients/include/client_utility.hpp:419-444): divide each row (lower) or
    column (upper) of the triangle by its diagonal entry, so the unit-diagonal
    matrix stays diagonally dominant, then overwrite the diagonal with a value
    the solve must never read (here ``poison``)."""
    for i in range(k):
        d = A[i, i]
        if uplo == "L":
            A[i, :i] = A[i, :i] / d
        else:
            A[:i, i] = A[:i, i] / d
        A[i, i] = poison
    return A
