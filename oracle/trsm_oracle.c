/*
 * trsm_oracle.c -- CPU oracle for the triangular-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load the library
 * built from it.  The product (rocblas_amd/) never links or calls it.
 *
 * What it restates, and from where:
 *   - oracle_trsm_{d,s}: the textbook reference-BLAS (netlib) xTRSM loop nest,
 *     column-major, all 8 side/uplo/trans cases plus unit/non-unit diagonal.
 *     Semantics follow the rocBLAS contract documented at
 *     library/include/internal/rocblas-functions.h:19302-19386
 *     (op(A)*X = alpha*B or X*op(A) = alpha*B, X overwrites B, only the uplo
 *     triangle of A is referenced, alpha == 0 => B := 0 with A unreferenced),
 *     and the reference's host comparator, which is cblas_?trsm
 *     (clients/common/cblas_interface.cpp:2352-2405; that BLAS -- AOCL-BLIS --
 *     is not in the reference tree and is not installed here).
 *     conjugate-transpose is treated as transpose for real types, as in
 *     library/src/blas3/rocblas_trsm_kernels.hpp:3625-3626.
 *   - oracle_trmm_d: B := alpha*op(A)*B or alpha*B*op(A) (reference BLAS
 *     DTRMM semantics), used by the reference's known-X test procedure
 *     (clients/include/blas3/testing_trsm.hpp:297-315).
 *   - oracle_backward_error_d/s: normwise relative backward error
 *       eta = ||alpha*B0 - op(A)*X||_1 / (||op(A)||_1*||X||_1 + |alpha|*||B0||_1)
 *     (side right: X*op(A)), accumulated in long double.
 *   - oracle_norm1_rel_diff: ||X - Y||_1 / ||Y||_1 (matrix 1-norm), the
 *     measure of clients/include/norm.hpp:449-492 used by the reference's
 *     40*eps*M acceptance test (clients/include/blas3/testing_trsm.hpp:30-31,
 *     433-448).
 *   - oracle_splitmix_uniform: counter-based SplitMix64 stream (Steele, Lea,
 *     Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014),
 *     u_i = mix(key + (i+1)*0x9E3779B97F4A7C15) >> 11 * 2^-53.
 *   - oracle_trsm_*_batched_omp: the same loop nest run over a strided batch,
 *     OpenMP-parallel over instances: the CPU baseline.
 *   - oracle_trtri_{d,s}: in-place triangular inverse, the unblocked LAPACK
 *     xTRTI2 algorithm that the reference's host comparator lapack_xtrtri
 *     (clients/include/lapack_utilities.hpp:625-708) uses on its blocks.
 *
 * Enum values are rocBLAS's (library/include/internal/rocblas-types.h:125-157),
 * which coincide with CBLAS's: side 141/142, fill 121/122, op 111/112/113,
 * diag 131/132.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { OP_N = 111, OP_T = 112, OP_C = 113 };
enum { FILL_U = 121, FILL_L = 122 };
enum { DIAG_N = 131, DIAG_U = 132 };
enum { SIDE_L = 141, SIDE_R = 142 };

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ------------------------------------------------------------------------ */
/* netlib xTRSM loop nest, one template per precision via a macro.           */
/* ------------------------------------------------------------------------ */
#define DEFINE_TRSM(NAME, T)                                                          \
    void NAME(int side, int uplo, int trans, int diag, int m, int n, T alpha,         \
              const T* A, int lda, T* B, int ldb)                                     \
    {                                                                                 \
        int i, j, k;                                                                  \
        const int nounit = (diag == DIAG_N);                                          \
        const int upper  = (uplo == FILL_U);                                          \
        const int notr   = (trans == OP_N);                                           \
        if(m <= 0 || n <= 0)                                                          \
            return;                                                                   \
        if(alpha == (T)0)                                                             \
        {                                                                             \
            for(j = 0; j < n; j++)                                                    \
                for(i = 0; i < m; i++)                                                \
                    B[IDX(i, j, ldb)] = (T)0;                                         \
            return;                                                                   \
        }                                                                             \
        if(side == SIDE_L)                                                            \
        {                                                                             \
            if(notr)                                                                  \
            {                                                                         \
                /* B := alpha*inv(A)*B */                                             \
                if(upper)                                                             \
                {                                                                     \
                    for(j = 0; j < n; j++)                                            \
                    {                                                                 \
                        if(alpha != (T)1)                                             \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, j, ldb)] = alpha * B[IDX(i, j, ldb)];        \
                        for(k = m - 1; k >= 0; k--)                                   \
                        {                                                             \
                            if(B[IDX(k, j, ldb)] != (T)0)                             \
                            {                                                         \
                                if(nounit)                                            \
                                    B[IDX(k, j, ldb)] /= A[IDX(k, k, lda)];           \
                                for(i = 0; i < k; i++)                                \
                                    B[IDX(i, j, ldb)] -= B[IDX(k, j, ldb)]            \
                                                         * A[IDX(i, k, lda)];         \
                            }                                                         \
                        }                                                             \
                    }                                                                 \
                }                                                                     \
                else                                                                  \
                {                                                                     \
                    for(j = 0; j < n; j++)                                            \
                    {                                                                 \
                        if(alpha != (T)1)                                             \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, j, ldb)] = alpha * B[IDX(i, j, ldb)];        \
                        for(k = 0; k < m; k++)                                        \
                        {                                                             \
                            if(B[IDX(k, j, ldb)] != (T)0)                             \
                            {                                                         \
                                if(nounit)                                            \
                                    B[IDX(k, j, ldb)] /= A[IDX(k, k, lda)];           \
                                for(i = k + 1; i < m; i++)                            \
                                    B[IDX(i, j, ldb)] -= B[IDX(k, j, ldb)]            \
                                                         * A[IDX(i, k, lda)];         \
                            }                                                         \
                        }                                                             \
                    }                                                                 \
                }                                                                     \
            }                                                                         \
            else                                                                      \
            {                                                                         \
                /* B := alpha*inv(A**T)*B */                                          \
                if(upper)                                                             \
                {                                                                     \
                    for(j = 0; j < n; j++)                                            \
                        for(i = 0; i < m; i++)                                        \
                        {                                                             \
                            T temp = alpha * B[IDX(i, j, ldb)];                       \
                            for(k = 0; k < i; k++)                                    \
                                temp -= A[IDX(k, i, lda)] * B[IDX(k, j, ldb)];        \
                            if(nounit)                                                \
                                temp /= A[IDX(i, i, lda)];                            \
                            B[IDX(i, j, ldb)] = temp;                                 \
                        }                                                             \
                }                                                                     \
                else                                                                  \
                {                                                                     \
                    for(j = 0; j < n; j++)                                            \
                        for(i = m - 1; i >= 0; i--)                                   \
                        {                                                             \
                            T temp = alpha * B[IDX(i, j, ldb)];                       \
                            for(k = i + 1; k < m; k++)                                \
                                temp -= A[IDX(k, i, lda)] * B[IDX(k, j, ldb)];        \
                            if(nounit)                                                \
                                temp /= A[IDX(i, i, lda)];                            \
                            B[IDX(i, j, ldb)] = temp;                                 \
                        }                                                             \
                }                                                                     \
            }                                                                         \
        }                                                                             \
        else                                                                          \
        {                                                                             \
            if(notr)                                                                  \
            {                                                                         \
                /* B := alpha*B*inv(A) */                                             \
                if(upper)                                                             \
                {                                                                     \
                    for(j = 0; j < n; j++)                                            \
                    {                                                                 \
                        if(alpha != (T)1)                                             \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, j, ldb)] = alpha * B[IDX(i, j, ldb)];        \
                        for(k = 0; k < j; k++)                                        \
                            if(A[IDX(k, j, lda)] != (T)0)                             \
                                for(i = 0; i < m; i++)                                \
                                    B[IDX(i, j, ldb)] -= A[IDX(k, j, lda)]            \
                                                         * B[IDX(i, k, ldb)];         \
                        if(nounit)                                                    \
                        {                                                             \
                            T temp = (T)1 / A[IDX(j, j, lda)];                        \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, j, ldb)] = temp * B[IDX(i, j, ldb)];         \
                        }                                                             \
                    }                                                                 \
                }                                                                     \
                else                                                                  \
                {                                                                     \
                    for(j = n - 1; j >= 0; j--)                                       \
                    {                                                                 \
                        if(alpha != (T)1)                                             \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, j, ldb)] = alpha * B[IDX(i, j, ldb)];        \
                        for(k = j + 1; k < n; k++)                                    \
                            if(A[IDX(k, j, lda)] != (T)0)                             \
                                for(i = 0; i < m; i++)                                \
                                    B[IDX(i, j, ldb)] -= A[IDX(k, j, lda)]            \
                                                         * B[IDX(i, k, ldb)];         \
                        if(nounit)                                                    \
                        {                                                             \
                            T temp = (T)1 / A[IDX(j, j, lda)];                        \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, j, ldb)] = temp * B[IDX(i, j, ldb)];         \
                        }                                                             \
                    }                                                                 \
                }                                                                     \
            }                                                                         \
            else                                                                      \
            {                                                                         \
                /* B := alpha*B*inv(A**T) */                                          \
                if(upper)                                                             \
                {                                                                     \
                    for(k = n - 1; k >= 0; k--)                                       \
                    {                                                                 \
                        if(nounit)                                                    \
                        {                                                             \
                            T temp = (T)1 / A[IDX(k, k, lda)];                        \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, k, ldb)] = temp * B[IDX(i, k, ldb)];         \
                        }                                                             \
                        for(j = 0; j < k; j++)                                        \
                            if(A[IDX(j, k, lda)] != (T)0)                             \
                            {                                                         \
                                T temp = A[IDX(j, k, lda)];                           \
                                for(i = 0; i < m; i++)                                \
                                    B[IDX(i, j, ldb)] -= temp * B[IDX(i, k, ldb)];    \
                            }                                                         \
                        if(alpha != (T)1)                                             \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, k, ldb)] = alpha * B[IDX(i, k, ldb)];        \
                    }                                                                 \
                }                                                                     \
                else                                                                  \
                {                                                                     \
                    for(k = 0; k < n; k++)                                            \
                    {                                                                 \
                        if(nounit)                                                    \
                        {                                                             \
                            T temp = (T)1 / A[IDX(k, k, lda)];                        \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, k, ldb)] = temp * B[IDX(i, k, ldb)];         \
                        }                                                             \
                        for(j = k + 1; j < n; j++)                                    \
                            if(A[IDX(j, k, lda)] != (T)0)                             \
                            {                                                         \
                                T temp = A[IDX(j, k, lda)];                           \
                                for(i = 0; i < m; i++)                                \
                                    B[IDX(i, j, ldb)] -= temp * B[IDX(i, k, ldb)];    \
                            }                                                         \
                        if(alpha != (T)1)                                             \
                            for(i = 0; i < m; i++)                                    \
                                B[IDX(i, k, ldb)] = alpha * B[IDX(i, k, ldb)];        \
                    }                                                                 \
                }                                                                     \
            }                                                                         \
        }                                                                             \
    }

DEFINE_TRSM(oracle_trsm_d, double)
DEFINE_TRSM(oracle_trsm_s, float)

/* Strided batch over the same loop nest; OpenMP-parallel over instances.  This
 * is the CPU baseline: the reference times one host cblas call per case
 * (clients/include/blas3/testing_trsm.hpp:571-577). */
void oracle_trsm_d_batched_omp(int side, int uplo, int trans, int diag, int m, int n,
                               double alpha, const double* A, int lda, long long stride_a,
                               double* B, int ldb, long long stride_b, int batch, int threads)
{
    int b;
#ifdef _OPENMP
    if(threads > 0)
        omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
    for(b = 0; b < batch; b++)
        oracle_trsm_d(side, uplo, trans, diag, m, n, alpha, A + (size_t)b * stride_a, lda,
                      B + (size_t)b * stride_b, ldb);
}

/* Single large solve split over the independent right-hand sides: columns of
 * B for side left, rows for side right (the loop nest is column-oriented, so a
 * row block of B is a sub-matrix with the same ldb). */
void oracle_trsm_d_omp(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                       const double* A, int lda, double* B, int ldb, int threads)
{
    int nt = 1, t;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_max_threads();
#endif
    const int rhs = side == SIDE_L ? n : m;
    if(nt > rhs)
        nt = rhs;
    if(nt < 1)
        nt = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for(t = 0; t < nt; t++)
    {
        int lo = (int)((long long)rhs * t / nt);
        int hi = (int)((long long)rhs * (t + 1) / nt);
        if(hi <= lo)
            continue;
        if(side == SIDE_L)
            oracle_trsm_d(side, uplo, trans, diag, m, hi - lo, alpha, A, lda,
                          B + (size_t)lo * ldb, ldb);
        else
            oracle_trsm_d(side, uplo, trans, diag, hi - lo, n, alpha, A, lda, B + lo, ldb);
    }
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* op(A) element accessor honouring uplo/diag: returns the value op(A)[i][j]  */
/* as the solve sees it (zero outside the triangle, one on a unit diagonal).  */
/* ------------------------------------------------------------------------ */
static long double opA_ld(int uplo, int trans, int diag, const double* A, int lda, int i, int j)
{
    int r = i, c = j;
    if(trans != OP_N)
    {
        r = j;
        c = i;
    }
    if(r == c)
        return diag == DIAG_U ? 1.0L : (long double)A[IDX(r, c, lda)];
    if(uplo == FILL_U ? (r < c) : (r > c))
        return (long double)A[IDX(r, c, lda)];
    return 0.0L;
}

static long double opA_ld_f(int uplo, int trans, int diag, const float* A, int lda, int i, int j)
{
    int r = i, c = j;
    if(trans != OP_N)
    {
        r = j;
        c = i;
    }
    if(r == c)
        return diag == DIAG_U ? 1.0L : (long double)A[IDX(r, c, lda)];
    if(uplo == FILL_U ? (r < c) : (r > c))
        return (long double)A[IDX(r, c, lda)];
    return 0.0L;
}

/* reference BLAS DTRMM semantics: B := alpha*op(A)*B (left) or alpha*B*op(A)
 * (right); computed in long double and rounded once. */
void oracle_trmm_d(int side, int uplo, int trans, int diag, int m, int n, double alpha,
                   const double* A, int lda, double* B, int ldb)
{
    int i, j, k;
    const int kk = side == SIDE_L ? m : n;
    if(m <= 0 || n <= 0)
        return;
    if(side == SIDE_L)
    {
        long double* col = (long double*)malloc(sizeof(long double) * (size_t)kk);
        for(j = 0; j < n; j++)
        {
            for(i = 0; i < m; i++)
            {
                long double s = 0.0L;
                for(k = 0; k < m; k++)
                    s += opA_ld(uplo, trans, diag, A, lda, i, k) * (long double)B[IDX(k, j, ldb)];
                col[i] = s;
            }
            for(i = 0; i < m; i++)
                B[IDX(i, j, ldb)] = (double)((long double)alpha * col[i]);
        }
        free(col);
    }
    else
    {
        long double* row = (long double*)malloc(sizeof(long double) * (size_t)kk);
        for(i = 0; i < m; i++)
        {
            for(j = 0; j < n; j++)
            {
                long double s = 0.0L;
                for(k = 0; k < n; k++)
                    s += (long double)B[IDX(i, k, ldb)] * opA_ld(uplo, trans, diag, A, lda, k, j);
                row[j] = s;
            }
            for(j = 0; j < n; j++)
                B[IDX(i, j, ldb)] = (double)((long double)alpha * row[j]);
        }
        free(row);
    }
}

/* matrix 1-norm (max column absolute sum) of op(A) as the solve sees it */
static long double norm1_opA(int uplo, int trans, int diag, const double* A, int lda, int k)
{
    long double best = 0.0L;
    for(int j = 0; j < k; j++)
    {
        long double s = 0.0L;
        for(int i = 0; i < k; i++)
            s += fabsl(opA_ld(uplo, trans, diag, A, lda, i, j));
        if(s > best)
            best = s;
    }
    return best;
}

static long double norm1_opA_f(int uplo, int trans, int diag, const float* A, int lda, int k)
{
    long double best = 0.0L;
    for(int j = 0; j < k; j++)
    {
        long double s = 0.0L;
        for(int i = 0; i < k; i++)
            s += fabsl(opA_ld_f(uplo, trans, diag, A, lda, i, j));
        if(s > best)
            best = s;
    }
    return best;
}

static long double norm1_mat(const double* X, int ldx, int m, int n)
{
    long double best = 0.0L;
    for(int j = 0; j < n; j++)
    {
        long double s = 0.0L;
        for(int i = 0; i < m; i++)
            s += fabsl((long double)X[IDX(i, j, ldx)]);
        if(s > best)
            best = s;
    }
    return best;
}

static long double norm1_mat_f(const float* X, int ldx, int m, int n)
{
    long double best = 0.0L;
    for(int j = 0; j < n; j++)
    {
        long double s = 0.0L;
        for(int i = 0; i < m; i++)
            s += fabsl((long double)X[IDX(i, j, ldx)]);
        if(s > best)
            best = s;
    }
    return best;
}

/* eta = ||alpha*B0 - op(A)*X||_1 / (||op(A)||_1 ||X||_1 + |alpha| ||B0||_1),
 * side right uses X*op(A).  Returns 0 for an empty problem. */
double oracle_backward_error_d(int side, int uplo, int trans, int diag, int m, int n,
                               double alpha, const double* A, int lda, const double* B0, int ldb0,
                               const double* X, int ldx)
{
    const int k = side == SIDE_L ? m : n;
    if(m <= 0 || n <= 0)
        return 0.0;
    long double rnorm = 0.0L;
    for(int j = 0; j < n; j++)
    {
        long double colsum = 0.0L;
        for(int i = 0; i < m; i++)
        {
            long double s = (long double)alpha * (long double)B0[IDX(i, j, ldb0)];
            if(side == SIDE_L)
                for(int p = 0; p < m; p++)
                    s -= opA_ld(uplo, trans, diag, A, lda, i, p) * (long double)X[IDX(p, j, ldx)];
            else
                for(int p = 0; p < n; p++)
                    s -= (long double)X[IDX(i, p, ldx)] * opA_ld(uplo, trans, diag, A, lda, p, j);
            colsum += fabsl(s);
        }
        if(colsum > rnorm)
            rnorm = colsum;
    }
    long double denom = norm1_opA(uplo, trans, diag, A, lda, k) * norm1_mat(X, ldx, m, n)
                        + fabsl((long double)alpha) * norm1_mat(B0, ldb0, m, n);
    if(denom == 0.0L)
        return rnorm == 0.0L ? 0.0 : INFINITY;
    return (double)(rnorm / denom);
}

double oracle_backward_error_s(int side, int uplo, int trans, int diag, int m, int n, float alpha,
                               const float* A, int lda, const float* B0, int ldb0, const float* X,
                               int ldx)
{
    const int k = side == SIDE_L ? m : n;
    if(m <= 0 || n <= 0)
        return 0.0;
    long double rnorm = 0.0L;
    for(int j = 0; j < n; j++)
    {
        long double colsum = 0.0L;
        for(int i = 0; i < m; i++)
        {
            long double s = (long double)alpha * (long double)B0[IDX(i, j, ldb0)];
            if(side == SIDE_L)
                for(int p = 0; p < m; p++)
                    s -= opA_ld_f(uplo, trans, diag, A, lda, i, p) * (long double)X[IDX(p, j, ldx)];
            else
                for(int p = 0; p < n; p++)
                    s -= (long double)X[IDX(i, p, ldx)] * opA_ld_f(uplo, trans, diag, A, lda, p, j);
            colsum += fabsl(s);
        }
        if(colsum > rnorm)
            rnorm = colsum;
    }
    long double denom = norm1_opA_f(uplo, trans, diag, A, lda, k) * norm1_mat_f(X, ldx, m, n)
                        + fabsl((long double)alpha) * norm1_mat_f(B0, ldb0, m, n);
    if(denom == 0.0L)
        return rnorm == 0.0L ? 0.0 : INFINITY;
    return (double)(rnorm / denom);
}

/* ||X - Y||_1 / ||Y||_1 with the matrix 1-norm (clients/include/norm.hpp:449-492) */
double oracle_norm1_rel_diff_d(int m, int n, const double* X, int ldx, const double* Y, int ldy)
{
    long double num = 0.0L, den = 0.0L;
    for(int j = 0; j < n; j++)
    {
        long double a = 0.0L, b = 0.0L;
        for(int i = 0; i < m; i++)
        {
            a += fabsl((long double)X[IDX(i, j, ldx)] - (long double)Y[IDX(i, j, ldy)]);
            b += fabsl((long double)Y[IDX(i, j, ldy)]);
        }
        if(a > num)
            num = a;
        if(b > den)
            den = b;
    }
    if(den == 0.0L)
        return num == 0.0L ? 0.0 : INFINITY;
    return (double)(num / den);
}

/* ------------------------------------------------------------------------ */
/* SplitMix64, counter form                                                   */
/* ------------------------------------------------------------------------ */
static inline uint64_t sm64_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t oracle_splitmix_key(uint64_t seed, uint64_t tag)
{
    return sm64_mix(seed + tag * 0xD1B54A32D192ED03ULL);
}

/* out[i] = lo + (hi - lo) * u_{offset+i}, u in [0,1) with 53 random bits */
void oracle_splitmix_uniform(uint64_t key, uint64_t offset, size_t count, double lo, double hi,
                             double* out)
{
    const double scale = 1.0 / 9007199254740992.0; /* 2^-53 */
    for(size_t i = 0; i < count; i++)
    {
        uint64_t z = sm64_mix(key + (offset + i + 1) * 0x9E3779B97F4A7C15ULL);
        double   u = (double)(z >> 11) * scale;
        out[i]     = lo + (hi - lo) * u;
    }
}

/* ------------------------------------------------------------------------ */
/* Triangular inverse, in place: the unblocked LAPACK xTRTI2 algorithm,     */
/* which the reference's host comparator lapack_xtrtri                      */
/* (clients/include/lapack_utilities.hpp:625-708) applies per 64-wide       */
/* block.  Column j of the inverse is the already-inverted triangle times  */
/* column j (an xTRMV), scaled by -1/A(j,j).                                */
/* ------------------------------------------------------------------------ */
#define DEFINE_TRTRI(NAME, T)                                                        \
    void NAME(int uplo, int diag, int n, T* A, int lda)                              \
    {                                                                                \
        const int nounit = (diag == DIAG_N);                                         \
        int       i, j, k;                                                           \
        if(uplo == FILL_U)                                                           \
        {                                                                            \
            for(j = 0; j < n; j++)                                                   \
            {                                                                        \
                T ajj;                                                               \
                if(nounit)                                                           \
                {                                                                    \
                    A[IDX(j, j, lda)] = (T)1 / A[IDX(j, j, lda)];                    \
                    ajj               = -A[IDX(j, j, lda)];                          \
                }                                                                    \
                else                                                                 \
                    ajj = (T)-1;                                                     \
                /* x := U * x, U = inverted leading j x j block, x = A(0:j, j) */   \
                for(k = 0; k < j; k++)                                               \
                {                                                                    \
                    T temp = A[IDX(k, j, lda)];                                      \
                    if(temp != (T)0)                                                 \
                    {                                                                \
                        for(i = 0; i < k; i++)                                       \
                            A[IDX(i, j, lda)] += temp * A[IDX(i, k, lda)];           \
                        if(nounit)                                                   \
                            A[IDX(k, j, lda)] = temp * A[IDX(k, k, lda)];            \
                    }                                                                \
                }                                                                    \
                for(i = 0; i < j; i++)                                               \
                    A[IDX(i, j, lda)] *= ajj;                                        \
            }                                                                        \
        }                                                                            \
        else                                                                         \
        {                                                                            \
            for(j = n - 1; j >= 0; j--)                                              \
            {                                                                        \
                T ajj;                                                               \
                if(nounit)                                                           \
                {                                                                    \
                    A[IDX(j, j, lda)] = (T)1 / A[IDX(j, j, lda)];                    \
                    ajj               = -A[IDX(j, j, lda)];                          \
                }                                                                    \
                else                                                                 \
                    ajj = (T)-1;                                                     \
                if(j < n - 1)                                                        \
                {                                                                    \
                    /* x := L * x, L = inverted trailing block, x = A(j+1:n, j) */  \
                    for(k = n - 1; k > j; k--)                                       \
                    {                                                                \
                        T temp = A[IDX(k, j, lda)];                                  \
                        if(temp != (T)0)                                             \
                        {                                                            \
                            for(i = n - 1; i > k; i--)                               \
                                A[IDX(i, j, lda)] += temp * A[IDX(i, k, lda)];       \
                            if(nounit)                                               \
                                A[IDX(k, j, lda)] = temp * A[IDX(k, k, lda)];        \
                        }                                                            \
                    }                                                                \
                    for(i = j + 1; i < n; i++)                                       \
                        A[IDX(i, j, lda)] *= ajj;                                    \
                }                                                                    \
            }                                                                        \
        }                                                                            \
    }

DEFINE_TRTRI(oracle_trtri_d, double)
DEFINE_TRTRI(oracle_trtri_s, float)
