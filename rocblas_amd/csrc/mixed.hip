// mixed.hip -- FP64 triangular solve with low-precision trailing updates and
// FP64 iterative refinement:
//
//   rocblas_amd_dtrsm_mixed(handle, side, uplo, transA, diag, m, n, alpha,
//                           A, lda, B, ldb, form, block, max_refine, tol, info)
//
// Same problem and argument rules as rocblas_dtrsm; B is overwritten with X.
//
//   1. FP64 inverses of the NB x NB diagonal blocks of A (identity + the FP64
//      blocked solve, i.e. trtri), and an FP32 copy of A's triangle.
//   2. Low-precision solve X0 of op(A) X = alpha B: recursive over NB blocks;
//      a leaf is X_I = op(invA_II) * Y_I (FP64 MFMA GEMM), a trailing update
//      Y_2 -= op(A)_21 * X_1 runs on FP32 MFMA (the FP64 X_1 is rounded on its
//      way into LDS, accumulation in FP32, the update is added to FP64 Y).
//   3. Refinement: R = alpha B0 - op(A) X in FP64 (one triangular-masked FP64
//      MFMA GEMM), eta = |R|_1 / (|op(A)|_1 |X|_1 + |alpha| |B0|_1); stop when
//      eta <= tol, else solve op(A) D = R as in 2 and X += D.
//
// The FP64 work is the diagonal-block inverses, the leaf products and the
// residual; FP32 carries the trailing updates.  Every returned X has its
// backward error measured by the residual that accepted it.
#include "device_common.hpp"
#include "gemm_bf16x3.hpp"
#include "mixed_tuning.hpp"

#include "../../include/rocblas_amd_ext.h"

#include <cmath>
#include <new>
#include <vector>

namespace rbamd
{
namespace
{
// ---------------------------------------------------------------------------
// elementwise kernels (2-D grids: x over rows, y over columns)
// ---------------------------------------------------------------------------
inline dim3 grid2d(int m, int n)
{
    return dim3(unsigned((m + 255) / 256), unsigned(std::min(n, 65535)));
}

// D := alpha * S  (alpha == 0 gives exact zeros)
__global__ void __launch_bounds__(256) scale_copy_kernel(int m, int n, const double* S, int lds,
                                                         double* D, int ldd, double alpha)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if(i >= m)
        return;
    for(int j = blockIdx.y; j < n; j += gridDim.y)
    {
        const double v          = S[i + int64_t(j) * lds];
        D[i + int64_t(j) * ldd] = alpha == 0.0 ? 0.0 : alpha * v;
    }
}

// X += D
__global__ void __launch_bounds__(256) add_kernel(int m, int n, const double* D, int ldd,
                                                  double* X, int ldx)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if(i >= m)
        return;
    for(int j = blockIdx.y; j < n; j += gridDim.y)
        X[i + int64_t(j) * ldx] += D[i + int64_t(j) * ldd];
}

// L := float(triangle of A), zero elsewhere (k x k, leading dimension k); the
// diagonal is never used by the trailing updates and is stored as zero
__global__ void __launch_bounds__(256) lowcopy_kernel(int k, const double* A, int lda, float* L,
                                                      bool upper)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if(i >= k)
        return;
    for(int j = blockIdx.y; j < k; j += gridDim.y)
    {
        const bool in = upper ? i < j : i > j;
        L[i + int64_t(j) * k] = in ? float(A[i + int64_t(j) * lda]) : 0.0f;
    }
}

// sums[j] = sum_i |op(M)(i, j)| for an m x n block: column sums when !rows,
// row sums (sums[i] += over columns) when rows.  Optional triangle mask of a
// square M (tri: 0 none, 1 lower, 2 upper; unit: diagonal counts as 1).
// Column sums: one wavefront per column; row sums: one thread per row and a
// 64-column chunk per grid column, accumulated with atomics.
__global__ void __launch_bounds__(64) colsum_kernel(int m, int n, const double* M, int ldm,
                                                   int tri, bool unit, double* sums)
{
    for(int j = blockIdx.x; j < n; j += gridDim.x)
    {
        double s = 0.0;
        for(int i = threadIdx.x; i < m; i += 64)
        {
            if((tri == 1 && i < j) || (tri == 2 && i > j))
                continue;
            s += (unit && i == j) ? 1.0 : fabs(M[i + int64_t(j) * ldm]);
        }
        for(int o = 32; o > 0; o >>= 1)
            s += __shfl_xor(s, o);
        if(threadIdx.x == 0)
            sums[j] = s;
    }
}

__global__ void __launch_bounds__(256) rowsum_kernel(int m, int n, const double* M, int ldm,
                                                     int tri, bool unit, double* sums)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if(i >= m)
        return;
    const int j0 = blockIdx.y * 64;
    const int j1 = min(n, j0 + 64);
    double    s  = 0.0;
    for(int j = j0; j < j1; ++j)
    {
        if((tri == 1 && i < j) || (tri == 2 && i > j))
            continue;
        s += (unit && i == j) ? 1.0 : fabs(M[i + int64_t(j) * ldm]);
    }
    if(j1 > j0)
        atomicAdd(&sums[i], s);
}

__global__ void __launch_bounds__(256) fill_kernel(int n, double* p, double v)
{
    for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        p[i] = v;
}

// out[slot] = max(v[0..n))
__global__ void __launch_bounds__(256) max_kernel(int n, const double* v, double* out)
{
    __shared__ double red[256];
    double            s = 0.0;
    for(int i = threadIdx.x; i < n; i += 256)
        s = fmax(s, v[i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for(int o = 128; o > 0; o >>= 1)
    {
        if(threadIdx.x < o)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if(threadIdx.x == 0)
        *out = red[0];
}

// ---------------------------------------------------------------------------
// the solver
// ---------------------------------------------------------------------------
struct Mixed
{
    rocblas_handle    h;
    rocblas_side      side;
    rocblas_fill      uplo;
    rocblas_operation trans;
    rocblas_diagonal  diag;
    int               m, n, K, NB;
    const double*     A;
    int               lda;
    bool              sideL, tA, rev, upperA;
    bool              bf16x3; // updates as split-BF16x3 products (else FP32 MFMA)
    // workspace
    float*  Alp;   // K x K, ld K
    double* Ainv;  // nslots x NB x NB
    double* W;     // leaf product (NB x n or m x NB)
    double* norms; // max(m, n, K) partial sums + 4 scalars
    int     nfull, part;

    int real_lo(int a, int b) const
    {
        return rev ? K - b : a;
    }
    int slot_of(int r0, int nb) const
    {
        if(nb != NB)
            return nfull; // the partial block
        return rev && part ? (r0 - part) / NB : r0 / NB;
    }

    rocblas_status check_launch() const
    {
        return launch_status();
    }

    // FP64 inverses of the diagonal blocks (trtri = identity + FP64 solve)
    rocblas_status invert_diagonal()
    {
        rocblas_status st;
        const int64_t  nb2 = int64_t(NB) * NB;
        if(nfull)
        {
            const int off0 = rev ? part : 0;
            BatchPtr<double> X{nullptr, Ainv, nb2, 0};
            if((st = identity_launch<double>(h, NB, X, NB, nfull)) != rocblas_status_success)
                return st;
            TrsmProblem<double> p;
            p.side       = rocblas_side_left;
            p.uplo       = uplo;
            p.trans      = rocblas_operation_none;
            p.diag       = diag;
            p.m = p.n    = NB;
            p.lda        = lda;
            p.ldb        = NB;
            p.A          = BatchPtr<const double>{nullptr, A, int64_t(NB) * (lda + 1),
                                         int64_t(off0) * (lda + 1)};
            p.B          = X;
            p.batch      = nfull;
            p.alpha_host = 1.0;
            p.alpha_dev  = nullptr;
            if((st = trsm_launch<double>(h, p)) != rocblas_status_success)
                return st;
        }
        if(part)
        {
            const int        r0 = rev ? 0 : nfull * NB;
            BatchPtr<double> X{nullptr, Ainv + int64_t(nfull) * nb2, 0, 0};
            if((st = identity_launch<double>(h, part, X, NB, 1)) != rocblas_status_success)
                return st;
            TrsmProblem<double> p;
            p.side       = rocblas_side_left;
            p.uplo       = uplo;
            p.trans      = rocblas_operation_none;
            p.diag       = diag;
            p.m = p.n    = part;
            p.lda        = lda;
            p.ldb        = NB;
            p.A          = BatchPtr<const double>{nullptr, A, 0, int64_t(r0) * (lda + 1)};
            p.B          = X;
            p.batch      = 1;
            p.alpha_host = 1.0;
            p.alpha_dev  = nullptr;
            if((st = trsm_launch<double>(h, p)) != rocblas_status_success)
                return st;
        }
        return rocblas_status_success;
    }

    // Y[block a..b) := op(invA_block) Y[block]  (FP64)
    rocblas_status leaf(double* Y, int ldy, int a, int b)
    {
        const int     nb   = b - a;
        const int     r0   = real_lo(a, b);
        const double* inv  = Ainv + int64_t(slot_of(r0, nb)) * NB * NB;
        const bool    opT  = trans != rocblas_operation_none;
        GemmArgs<double> g{};
        g.alpha = 1.0;
        g.beta  = 0.0;
        g.batch = 1;
        if(sideL)
        {
            g.M   = nb;
            g.N   = n;
            g.K   = nb;
            g.opA = opT;
            g.A   = BatchPtr<const double>{nullptr, inv, 0, 0};
            g.lda = NB;
            g.opB = 0;
            g.B   = BatchPtr<const double>{nullptr, Y, 0, int64_t(r0)};
            g.ldb = ldy;
            g.C   = BatchPtr<double>{nullptr, W, 0, 0};
            g.ldc = NB;
        }
        else
        {
            g.M   = m;
            g.N   = nb;
            g.K   = nb;
            g.opA = 0;
            g.A   = BatchPtr<const double>{nullptr, Y, 0, int64_t(r0) * ldy};
            g.lda = ldy;
            g.opB = opT;
            g.B   = BatchPtr<const double>{nullptr, inv, 0, 0};
            g.ldb = NB;
            g.C   = BatchPtr<double>{nullptr, W, 0, 0};
            g.ldc = m;
        }
        rocblas_status st = launch_gemm(h, g);
        if(st != rocblas_status_success)
            return st;
        if(sideL)
            hipLaunchKernelGGL(scale_copy_kernel, grid2d(nb, n), dim3(256), 0, h->stream, nb, n,
                               W, NB, Y + r0, ldy, 1.0);
        else
            hipLaunchKernelGGL(scale_copy_kernel, grid2d(m, nb), dim3(256), 0, h->stream, m, nb,
                               W, m, Y + int64_t(r0) * ldy, ldy, 1.0);
        return check_launch();
    }

    // Y[I2] -= M[I2, I1] Y[I1] with the FP32 copy of A, on FP32 MFMA or as
    // split-BF16x3 products (gemm_bf16x3.hip)
    rocblas_status update(double* Y, int ldy, int a1, int b1, int a2, int b2)
    {
        const int     R1 = real_lo(a1, b1), n1 = b1 - a1;
        const int     R2 = real_lo(a2, b2), n2 = b2 - a2;
        const int64_t a_off = tA ? int64_t(R1) + int64_t(R2) * K : int64_t(R2) + int64_t(R1) * K;
        if(sideL)
        {
            GemmArgsT<float, double, double> g{};
            g.alpha = -1.0;
            g.beta  = 1.0;
            g.batch = 1;
            g.M     = n2;
            g.N     = n;
            g.K     = n1;
            g.opA   = tA ? 1 : 0;
            g.A     = BatchPtr<const float>{nullptr, Alp, 0, a_off};
            g.lda   = K;
            g.opB   = 0;
            g.B     = BatchPtr<const double>{nullptr, Y, 0, int64_t(R1)};
            g.ldb   = ldy;
            g.C     = BatchPtr<double>{nullptr, Y, 0, int64_t(R2)};
            g.ldc   = ldy;
            return bf16x3 ? launch_gemm_bf16x3(h, g) : launch_gemm(h, g);
        }
        GemmArgsT<double, float, double> g{};
        g.alpha = -1.0;
        g.beta  = 1.0;
        g.batch = 1;
        g.M     = m;
        g.N     = n2;
        g.K     = n1;
        g.opA   = 0;
        g.A     = BatchPtr<const double>{nullptr, Y, 0, int64_t(R1) * ldy};
        g.lda   = ldy;
        g.opB   = tA ? 0 : 1;
        g.B     = BatchPtr<const float>{nullptr, Alp, 0, a_off};
        g.ldb   = K;
        g.C     = BatchPtr<double>{nullptr, Y, 0, int64_t(R2) * ldy};
        g.ldc   = ldy;
        return bf16x3 ? launch_gemm_bf16x3(h, g) : launch_gemm(h, g);
    }

    rocblas_status solve(double* Y, int ldy, int a, int b)
    {
        const int kk = b - a;
        if(kk <= NB)
            return leaf(Y, ldy, a, b);
        const int nb = (kk + NB - 1) / NB;
        const int K1 = ((nb + 1) / 2) * NB;
        rocblas_status st = solve(Y, ldy, a, a + K1);
        if(st != rocblas_status_success)
            return st;
        if((st = update(Y, ldy, a, a + K1, a + K1, b)) != rocblas_status_success)
            return st;
        return solve(Y, ldy, a + K1, b);
    }

    // R := alpha B0 - op(A) X   (FP64, triangular-masked GEMM)
    rocblas_status residual(const double* B0, const double* X, int ldx, double alpha, double* R)
    {
        hipLaunchKernelGGL(scale_copy_kernel, grid2d(m, n), dim3(256), 0, h->stream, m, n, B0, m,
                           R, m, alpha);
        rocblas_status st = check_launch();
        if(st != rocblas_status_success)
            return st;
        const bool opT     = trans != rocblas_operation_none;
        const bool opLower = (uplo == rocblas_fill_lower) != opT;
        const int  unit    = diag == rocblas_diagonal_unit ? GT_UNIT : 0;
        GemmArgs<double> g{};
        g.alpha = -1.0;
        g.beta  = 1.0;
        g.batch = 1;
        g.M     = m;
        g.N     = n;
        g.C     = BatchPtr<double>{nullptr, R, 0, 0};
        g.ldc   = m;
        if(sideL)
        {
            g.K   = m;
            g.opA = opT;
            g.A   = BatchPtr<const double>{nullptr, A, 0, 0};
            g.lda = lda;
            g.opB = 0;
            g.B   = BatchPtr<const double>{nullptr, X, 0, 0};
            g.ldb = ldx;
            g.tri = (opLower ? GT_A_LOWER : GT_A_UPPER) | unit;
        }
        else
        {
            g.K   = n;
            g.opA = 0;
            g.A   = BatchPtr<const double>{nullptr, X, 0, 0};
            g.lda = ldx;
            g.opB = opT;
            g.B   = BatchPtr<const double>{nullptr, A, 0, 0};
            g.ldb = lda;
            g.tri = (opLower ? GT_B_LOWER : GT_B_UPPER) | unit;
        }
        return launch_gemm(h, g);
    }

    // norms[slot] = |M|_1 (max column sum) of an m x n matrix
    rocblas_status norm1(const double* M, int rows, int cols, int ld, int slot)
    {
        double* sums = norms + 8;
        hipLaunchKernelGGL(colsum_kernel, dim3(unsigned(std::min(cols, 65535))), dim3(64), 0,
                           h->stream, rows, cols, M, ld, 0, false, sums);
        hipLaunchKernelGGL(max_kernel, dim3(1), dim3(256), 0, h->stream, cols, sums,
                           norms + slot);
        return check_launch();
    }

    // norms[slot] = |op(A)|_1 over the referenced triangle
    rocblas_status norm1_opA(int slot)
    {
        double*    sums = norms + 8;
        const int  tri  = uplo == rocblas_fill_lower ? 1 : 2;
        const bool unit = diag == rocblas_diagonal_unit;
        if(trans == rocblas_operation_none)
            hipLaunchKernelGGL(colsum_kernel, dim3(unsigned(std::min(K, 65535))), dim3(64), 0,
                               h->stream, K, K, A, lda, tri, unit, sums);
        else
        {
            hipLaunchKernelGGL(fill_kernel, dim3(unsigned(std::min((K + 255) / 256, 1024))),
                               dim3(256), 0, h->stream, K, sums, 0.0);
            hipLaunchKernelGGL(rowsum_kernel, dim3(unsigned((K + 255) / 256),
                                                   unsigned((K + 63) / 64)),
                               dim3(256), 0, h->stream, K, K, A, lda, tri, unit, sums);
        }
        hipLaunchKernelGGL(max_kernel, dim3(1), dim3(256), 0, h->stream, K, sums, norms + slot);
        return check_launch();
    }
};

template <size_t N>
int table_block(const MixedBlockEntry (&t)[N], int K)
{
    for(const auto& e : t)
        if(K <= e.max_k)
            return e.nb;
    return t[N - 1].nb;
}

int tuned_block(int K, bool bf16x3)
{
    return bf16x3 ? table_block(kMixedBlockTableBf16x3, K) : table_block(kMixedBlockTableFp32, K);
}

rocblas_status mixed_entry(rocblas_handle             handle,
                           rocblas_side               side,
                           rocblas_fill               uplo,
                           rocblas_operation          transA,
                           rocblas_diagonal           diag,
                           rocblas_int                m,
                           rocblas_int                n,
                           const double*              alpha,
                           const double*              A,
                           rocblas_int                lda,
                           double*                    B,
                           rocblas_int                ldb,
                           rocblas_amd_precision_form form,
                           rocblas_int                block,
                           rocblas_int                max_refine,
                           double                     tol,
                           rocblas_amd_refine_info*   info)
{
    // argument checks in the order of rocblas_dtrsm
    if(!handle)
        return rocblas_status_invalid_handle;
    if(side != rocblas_side_left && side != rocblas_side_right)
        return rocblas_status_invalid_value;
    if(uplo != rocblas_fill_lower && uplo != rocblas_fill_upper)
        return rocblas_status_invalid_value;
    if(transA != rocblas_operation_none && transA != rocblas_operation_transpose
       && transA != rocblas_operation_conjugate_transpose)
        return rocblas_status_invalid_value;
    if(diag != rocblas_diagonal_non_unit && diag != rocblas_diagonal_unit)
        return rocblas_status_invalid_value;
    if(form != rocblas_amd_form_fp32 && form != rocblas_amd_form_bf16x3)
        return rocblas_status_invalid_value;
    const int K = side == rocblas_side_left ? m : n;
    if(m < 0 || n < 0 || lda < K || lda < 1 || ldb < m || ldb < 1)
        return rocblas_status_invalid_size;
    if(block < 0 || max_refine < 0)
        return rocblas_status_invalid_size;
    if(info)
    {
        info->iterations     = 0;
        info->backward_error = 0.0;
        info->converged      = 1;
    }
    if(!m || !n)
        return handle->size_query ? rocblas_status_size_unchanged : rocblas_status_success;
    if(handle->size_query)
        return rocblas_status_size_unchanged; // sized on demand (library-managed workspace)
    if(!B || !alpha)
        return rocblas_status_invalid_pointer;
    if(transA == rocblas_operation_conjugate_transpose)
        transA = rocblas_operation_transpose;

    DeviceGuard dg(handle->device);
    double      a_host = 0.0;
    if(handle->pointer_mode == rocblas_pointer_mode_host)
        a_host = *alpha;
    else
    {
        hipError_t e = hipMemcpyAsync(&a_host, alpha, sizeof(double), hipMemcpyDeviceToHost,
                                      handle->stream);
        if(e == hipSuccess)
            e = hipStreamSynchronize(handle->stream);
        if(e != hipSuccess)
            return hip_to_status(e);
    }
    if(a_host == 0.0)
        return zero_matrix_launch<double>(handle, m, n, BatchPtr<double>{nullptr, B, 0, 0}, ldb,
                                          1);
    if(!A)
        return rocblas_status_invalid_pointer;

    Mixed s{};
    s.h      = handle;
    s.side   = side;
    s.uplo   = uplo;
    s.trans  = transA;
    s.diag   = diag;
    s.m      = m;
    s.n      = n;
    s.K      = K;
    s.bf16x3 = form == rocblas_amd_form_bf16x3;
    s.NB     = std::min(block ? block : tuned_block(K, s.bf16x3), K);
    s.A      = A;
    s.lda    = lda;
    s.sideL  = side == rocblas_side_left;
    s.tA     = s.sideL ? (transA != rocblas_operation_none) : (transA == rocblas_operation_none);
    s.upperA = uplo == rocblas_fill_upper;
    s.rev    = !((!s.upperA) != s.tA);
    s.nfull  = K / s.NB;
    s.part   = K % s.NB;

    // workspace: B0, R (m x n), W, Ainv, norms, Alp
    const int64_t mn     = int64_t(m) * n;
    const int64_t wsz    = s.sideL ? int64_t(s.NB) * n : int64_t(m) * s.NB;
    const int64_t nslots = s.nfull + (s.part ? 1 : 0);
    const int64_t nrm    = 8 + std::max<int64_t>(std::max(m, n), K);
    size_t        bytes  = 0;
    auto          carve  = [&](size_t b) {
        const size_t off = bytes;
        bytes += roundup_ws(b);
        return off;
    };
    const size_t o_B0   = carve(sizeof(double) * mn);
    const size_t o_R    = carve(sizeof(double) * mn);
    const size_t o_W    = carve(sizeof(double) * wsz);
    const size_t o_inv  = carve(sizeof(double) * nslots * s.NB * s.NB);
    const size_t o_norm = carve(sizeof(double) * nrm);
    const size_t o_Alp  = carve(sizeof(float) * int64_t(K) * K);
    WorkspaceLease lease(handle, bytes);
    if(!lease.ptr)
        return rocblas_status_memory_error;
    char*   base = static_cast<char*>(lease.ptr);
    double* B0   = reinterpret_cast<double*>(base + o_B0);
    double* R    = reinterpret_cast<double*>(base + o_R);
    s.W          = reinterpret_cast<double*>(base + o_W);
    s.Ainv       = reinterpret_cast<double*>(base + o_inv);
    s.norms      = reinterpret_cast<double*>(base + o_norm);
    s.Alp        = reinterpret_cast<float*>(base + o_Alp);
    hipStream_t st_ = handle->stream;

    rocblas_status st;
    // B0 := B (ld m); B := alpha B; FP32 triangle; FP64 diagonal-block inverses
    hipLaunchKernelGGL(scale_copy_kernel, grid2d(m, n), dim3(256), 0, st_, m, n, B, ldb, B0, m,
                       1.0);
    hipLaunchKernelGGL(scale_copy_kernel, grid2d(m, n), dim3(256), 0, st_, m, n, B, ldb, B, ldb,
                       a_host);
    hipLaunchKernelGGL(lowcopy_kernel, grid2d(K, K), dim3(256), 0, st_, K, A, lda, s.Alp,
                       s.upperA);
    if((st = launch_status()) != rocblas_status_success)
        return st;
    if((st = s.invert_diagonal()) != rocblas_status_success)
        return st;
    // X0
    if((st = s.solve(B, ldb, 0, K)) != rocblas_status_success)
        return st;

    const bool check = tol > 0.0;
    if(check)
    {
        if((st = s.norm1_opA(0)) != rocblas_status_success)
            return st;
        if((st = s.norm1(B0, m, n, m, 1)) != rocblas_status_success)
            return st;
    }
    int    it  = 0;
    double eta = 0.0;
    bool   ok  = !check;
    for(;; ++it)
    {
        if(check || it < max_refine)
        {
            if((st = s.residual(B0, B, ldb, a_host, R)) != rocblas_status_success)
                return st;
        }
        if(check)
        {
            if((st = s.norm1(R, m, n, m, 2)) != rocblas_status_success)
                return st;
            if((st = s.norm1(B, m, n, ldb, 3)) != rocblas_status_success)
                return st;
            double     nv[4];
            hipError_t e = hipMemcpyAsync(nv, s.norms, sizeof(nv), hipMemcpyDeviceToHost, st_);
            if(e == hipSuccess)
                e = hipStreamSynchronize(st_);
            if(e != hipSuccess)
                return hip_to_status(e);
            const double den = nv[0] * nv[3] + std::fabs(a_host) * nv[1];
            eta              = den > 0.0 ? nv[2] / den : nv[2];
            if(!(eta > tol)) // also stops on NaN-free zero residual
            {
                ok = std::isfinite(eta);
                break;
            }
        }
        if(it >= max_refine)
            break;
        // D: op(A) D = R in low precision, X += D
        if((st = s.solve(R, m, 0, K)) != rocblas_status_success)
            return st;
        hipLaunchKernelGGL(add_kernel, grid2d(m, n), dim3(256), 0, st_, m, n, R, m, B, ldb);
        if((st = launch_status()) != rocblas_status_success)
            return st;
    }
    if(info)
    {
        info->iterations     = it;
        info->backward_error = check ? eta : -1.0;
        info->converged      = ok ? 1 : 0;
    }
    return rocblas_status_success;
}
} // namespace
} // namespace rbamd

using namespace rbamd;

extern "C" rocblas_status rocblas_amd_dtrsm_mixed(rocblas_handle             handle,
                                                  rocblas_side               side,
                                                  rocblas_fill               uplo,
                                                  rocblas_operation          transA,
                                                  rocblas_diagonal           diag,
                                                  rocblas_int                m,
                                                  rocblas_int                n,
                                                  const double*              alpha,
                                                  const double*              A,
                                                  rocblas_int                lda,
                                                  double*                    B,
                                                  rocblas_int                ldb,
                                                  rocblas_amd_precision_form form,
                                                  rocblas_int                block,
                                                  rocblas_int                max_refine,
                                                  double                     tol,
                                                  rocblas_amd_refine_info*   info)
try
{
    return mixed_entry(handle, side, uplo, transA, diag, m, n, alpha, A, lda, B, ldb, form, block,
                       max_refine, tol, info);
}
catch(const std::bad_alloc&)
{
    return rocblas_status_memory_error;
}
catch(...)
{
    return rocblas_status_internal_error;
}
