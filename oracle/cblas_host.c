/*
 * cblas_host.c -- a host CBLAS ?trsm, loaded at run time, timed as a second
 * CPU baseline beside the oracle's loop nest.
 *
 * TEST INFRASTRUCTURE ONLY (the CPU-baseline side of the checker; the product
 * never loads it).  The reference times its host comparator cblas_dtrsm
 * (clients/include/blas3/testing_trsm.hpp:571-577, called through
 * clients/common/cblas_interface.cpp:2380-2405); its BLAS (AOCL-BLIS) is not
 * available here, so any CBLAS library found on the host is loaded with
 * dlopen by path and symbol name (e.g. the OpenBLAS that ships with scipy,
 * whose CBLAS symbols carry a "scipy_" prefix).  The CBLAS enum values equal
 * rocBLAS's (order 102 = column major; side 141/142, fill 121/122, op
 * 111/112/113, diag 131/132), so the arguments pass through unchanged.
 */
#include <dlfcn.h>
#include <stddef.h>
#include <stdio.h>

#include <omp.h>

typedef void (*cblas_dtrsm_fn)(int order, int side, int uplo, int trans, int diag, int m, int n,
                               double alpha, const double* A, int lda, double* B, int ldb);
typedef void (*set_threads_fn)(int);

static cblas_dtrsm_fn g_dtrsm   = NULL;
static set_threads_fn g_threads = NULL;

/* Loads `path` and resolves `dtrsm_sym` (a cblas_dtrsm with the LP64 CBLAS
 * signature) and, if given and present, `threads_sym` (an
 * openblas_set_num_threads-style setter).  Returns 0 on success. */
int oracle_cblas_load(const char* path, const char* dtrsm_sym, const char* threads_sym)
{
    void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if(!h)
    {
        fprintf(stderr, "oracle_cblas_load: %s\n", dlerror());
        return 1;
    }
    g_dtrsm = (cblas_dtrsm_fn)dlsym(h, dtrsm_sym);
    if(!g_dtrsm)
        return 2;
    g_threads = threads_sym ? (set_threads_fn)dlsym(h, threads_sym) : NULL;
    return 0;
}

/* One column-major solve through the loaded library with `threads` BLAS
 * threads (0: the library's default). */
int oracle_cblas_dtrsm(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                       const double* A, int lda, double* B, int ldb, int threads)
{
    if(!g_dtrsm)
        return 1;
    if(g_threads && threads > 0)
        g_threads(threads);
    g_dtrsm(102, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb);
    return 0;
}

/* A strided batch, OpenMP over instances, one BLAS thread per instance. */
int oracle_cblas_dtrsm_batched_omp(int side, int uplo, int trans, int diag, int m, int n,
                                   double alpha, const double* A, int lda, long long stride_a,
                                   double* B, int ldb, long long stride_b, int batch, int threads)
{
    int b;
    if(!g_dtrsm)
        return 1;
    if(g_threads)
        g_threads(1);
    if(threads > 0)
        omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
    for(b = 0; b < batch; b++)
        g_dtrsm(102, side, uplo, trans, diag, m, n, alpha, A + (size_t)b * stride_a, lda,
                B + (size_t)b * stride_b, ldb);
    return 0;
}
