// trtri_api.cpp -- extern "C" rocblas_{s,d}trtri{,_batched,_strided_batched}.
//
// Checks and quick returns follow library/src/blas3/rocblas_trtri.hpp:31-62
// (the size query is answered first, as in library/src/blas3/rocblas_trtri.cpp:55-59).
// The inverse is built in k^3/3 flops (trtri_core below): the diagonal
// blocks by the inversion kernel of trtri_kernels.hip, the rest by doubling
// with triangular-masked MFMA GEMMs.  The triangle opposite `uplo` of invA
// comes out exactly zero ("may be cleared", rocblas-functions.h:19064-19071).
// When invA is A itself (in-place call), A is first copied to the handle
// workspace.
#include "internal.hpp"
#include "trtri_kernels.hpp"
#include "device_common.hpp"

#include <new>
#include <optional>
#include <type_traits>

namespace rbamd
{
// Workspace the balanced split products of an order-n inversion use: the
// largest need of any FP64 doubling product (trtri_core's levels; equal
// groups as the batch when `groups_batched`), 0 for FP32.
template <typename T>
size_t trtri_split_bytes(rocblas_handle h, int n, int batch, bool groups_batched)
{
    if constexpr(!std::is_same_v<T, double>)
        return 0;
    else
    {
        size_t need = 0;
        for(int64_t s = (n <= 16 ? 16 : kTrtriNB); s < n; s *= 2)
        {
            const int64_t full = n / (2 * s), n2last = n - full * 2 * s - s;
            if(full > 0)
            {
                const int gb = batch == 1 && full > 1 && groups_batched ? int(full) : batch;
                for(int tri : {GT_A_LOWER, GT_B_LOWER})
                    need = std::max(need, gemm_tri_split_bytes(h, int(s), int(s), int(s), tri, gb));
            }
            if(n2last > 0)
            {
                need = std::max(need, gemm_tri_split_bytes(h, int(n2last), int(s), int(n2last),
                                                           GT_A_LOWER, batch));
                need = std::max(need, gemm_tri_split_bytes(h, int(s), int(n2last), int(s),
                                                           GT_A_UPPER, batch));
            }
        }
        return need;
    }
}

namespace
{

template <typename T>
rocblas_status trtri_entry(rocblas_handle    handle,
                           rocblas_fill      uplo,
                           rocblas_diagonal  diag,
                           rocblas_int       n,
                           BatchPtr<const T> A,
                           bool              A_null,
                           rocblas_int       lda,
                           BatchPtr<T>       invA,
                           bool              invA_null,
                           rocblas_int       ldinvA,
                           rocblas_int       batch_count,
                           bool              in_place)
{
    if(!handle)
        return rocblas_status_invalid_handle;

    const size_t copy_bytes
        = in_place && n > 0 && batch_count > 0 ? size_t(batch_count) * n * n * sizeof(T) : 0;
    const size_t split_bytes = n > 0 && batch_count > 0 ? trtri_split_bytes<T>(
                                   handle, n, batch_count, !A.arr && !invA.arr)
                                                        : 0;
    const size_t ws_bytes = roundup_ws(copy_bytes) + split_bytes;
    if(handle->size_query)
    {
        const size_t need = roundup_ws(ws_bytes);
        if(!need || need <= handle->query_size)
            return rocblas_status_size_unchanged;
        handle->query_size = need;
        return rocblas_status_size_increased;
    }

    if(uplo != rocblas_fill_lower && uplo != rocblas_fill_upper)
        return rocblas_status_invalid_value;
    if(diag != rocblas_diagonal_non_unit && diag != rocblas_diagonal_unit)
        return rocblas_status_invalid_value;
    if(batch_count < 0 || n < 0 || lda < n || ldinvA < n)
        return rocblas_status_invalid_size;
    if(!n || !batch_count)
        return rocblas_status_success;
    if(A_null || invA_null)
        return rocblas_status_invalid_pointer;

    // the split products' partial tiles after the in-place copy; without
    // room for them (a user-managed workspace too small, a captured stream)
    // the products run without (other last bits)
    std::optional<WorkspaceLease> lease;
    lease.emplace(handle, ws_bytes);
    bool has_split = split_bytes && lease->ptr;
    if(!lease->ptr && split_bytes)
    {
        lease.reset();
        lease.emplace(handle, copy_bytes);
    }
    if(in_place && !lease->ptr)
        return rocblas_status_memory_error;
    char* const ws = static_cast<char*>(lease->ptr);
    return trtri_core<T>(handle, uplo, diag, n, A, lda, invA, ldinvA, batch_count,
                         in_place ? reinterpret_cast<T*>(ws) : nullptr,
                         has_split ? ws + roundup_ws(copy_bytes) : nullptr,
                         has_split ? split_bytes : 0);
}

} // namespace

namespace
{
// One doubling step over `groups` equal groups [r0, r0 + n1 + n2) spaced
// `gstep` rows apart (batch instances each, or, for one instance, the groups
// as the batch): the inverses X11 (n1) and X22 (n2) of the group's diagonal
// halves are done, and the off-diagonal block is formed as
//   lower:  X21 = -X22 (A21 X11)      upper:  X12 = -X11 (A12 X22)
// with the intermediate product held in the group's opposite block of X
// (cleared afterwards): stored as is when that block is square (n1 == n2),
// transposed otherwise.  Triangular masks keep both products to the
// triangles' flops (2 (n1^2 n2 + n1 n2^2) / 2).
template <typename T>
rocblas_status trtri_double_step(rocblas_handle    h,
                                 bool              upper,
                                 int               r0,
                                 int               n1,
                                 int               n2,
                                 int               groups,
                                 int64_t           gstep,
                                 BatchPtr<const T> A,
                                 int               lda,
                                 BatchPtr<T>       X,
                                 int               ldx,
                                 int               batch,
                                 void*             split_ws,
                                 size_t            split_bytes)
{
    // operand views at element offset `off` of each instance (or of each
    // group, when the groups are the batch)
    const bool multi  = groups > 1;
    auto       cview  = [&](const BatchPtr<const T>& P, int64_t off, int64_t ld) {
        return multi ? BatchPtr<const T>{nullptr, P.base, gstep * (ld + 1), P.offset + off}
                     : P.shifted(off);
    };
    auto xview = [&](int64_t off) {
        return multi ? BatchPtr<T>{nullptr, X.base, gstep * (int64_t(ldx) + 1), X.offset + off}
                     : X.shifted(off);
    };
    auto xcview = [&](int64_t off) {
        const BatchPtr<const T> Xc{(const T* const*)X.arr, X.base, X.stride, X.offset};
        return cview(Xc, off, ldx);
    };
    const int64_t ldx64 = ldx, lda64 = lda;
    const int64_t d11   = int64_t(r0) * (ldx64 + 1);        // X11 in X
    const int64_t d22   = int64_t(r0 + n1) * (ldx64 + 1);   // X22 in X
    const int64_t x21   = int64_t(r0 + n1) + r0 * ldx64;    // lower-left block of X
    const int64_t x12   = int64_t(r0) + int64_t(r0 + n1) * ldx64; // upper-right block of X
    const bool    sq    = n1 == n2;

    GemmArgs<T> g{};
    g.beta         = T(0);
    g.skip_if_zero = nullptr;
    g.split_ok     = 1; // products with few tiles may split their K ranges
    g.split_ws       = static_cast<double*>(split_ws);
    g.split_ws_tiles = int64_t(split_bytes / (128 * 128 * sizeof(double)));
    g.batch        = multi ? groups : batch;
    g.lda          = lda;
    g.ldb          = ldx;
    g.ldc          = ldx;
    // product 1: W = A21 X11 (n2 x n1) or W = A12 X22 (n1 x n2), into the
    // opposite block; when that block has the other shape, W^T is stored
    g.alpha = T(1);
    if(!upper)
    {
        const int64_t a21 = int64_t(r0 + n1) + r0 * lda64;
        if(sq)
        { // W (n2 x n1) = A21 (n2 x n1) * X11 (n1 x n1, lower)
            g.M = n2, g.N = n1, g.K = n1, g.opA = 0, g.opB = 0;
            g.A = cview(A, a21, lda), g.lda = lda;
            g.B = xcview(d11), g.ldb = ldx;
            g.tri = GT_B_LOWER;
        }
        else
        { // W^T (n1 x n2) = X11^T (upper) * A21^T
            g.M = n1, g.N = n2, g.K = n1, g.opA = 1, g.opB = 1;
            g.A = xcview(d11), g.lda = ldx;
            g.B = cview(A, a21, lda), g.ldb = lda;
            g.tri = GT_A_UPPER;
        }
        g.C = xview(x12);
    }
    else
    {
        const int64_t a12 = int64_t(r0) + int64_t(r0 + n1) * lda64;
        if(sq)
        { // W (n1 x n2) = A12 (n1 x n2) * X22 (n2 x n2, upper)
            g.M = n1, g.N = n2, g.K = n2, g.opA = 0, g.opB = 0;
            g.A = cview(A, a12, lda), g.lda = lda;
            g.B = xcview(d22), g.ldb = ldx;
            g.tri = GT_B_UPPER;
        }
        else
        { // W^T (n2 x n1) = X22^T (lower) * A12^T
            g.M = n2, g.N = n1, g.K = n2, g.opA = 1, g.opB = 1;
            g.A = xcview(d22), g.lda = ldx;
            g.B = cview(A, a12, lda), g.ldb = lda;
            g.tri = GT_A_LOWER;
        }
        g.C = xview(x21);
    }
    rocblas_status st = launch_gemm<T, T, T>(h, g);
    if(st != rocblas_status_success)
        return st;

    // product 2: X21 = -X22 W (n2 x n1) or X12 = -X11 W (n1 x n2)
    g.alpha = T(-1);
    g.lda   = ldx;
    g.ldb   = ldx;
    g.opA   = 0;
    g.opB   = sq ? 0 : 1;
    if(!upper)
    {
        g.M = n2, g.N = n1, g.K = n2;
        g.A   = xcview(d22);
        g.B   = xcview(x12);
        g.tri = GT_A_LOWER;
        g.C   = xview(x21);
    }
    else
    {
        g.M = n1, g.N = n2, g.K = n1;
        g.A   = xcview(d11);
        g.B   = xcview(x21);
        g.tri = GT_A_UPPER;
        g.C   = xview(x12);
    }
    return launch_gemm<T, T, T>(h, g);
}
} // namespace

// Blocked inversion in k^3/3 flops: the diagonal blocks of order kTrtriNB
// are inverted by trtri_diag_kernel in one launch, then the inverse is built
// bottom-up by doubling: at level s the groups of 2s rows (the last may be
// shorter) get their off-diagonal block from two triangular-masked GEMMs
// (trtri_double_step), all equal groups of a level in one batched launch
// when there is one instance.  Finally the opposite triangle, used as
// scratch by the doubling, is cleared.
template <typename T>
rocblas_status trtri_core(rocblas_handle    handle,
                          rocblas_fill      uplo,
                          rocblas_diagonal  diag,
                          int               n,
                          BatchPtr<const T> A,
                          int               lda,
                          BatchPtr<T>       invA,
                          int               ldinvA,
                          int               batch_count,
                          T*                scratch,
                          void*             split_ws,
                          size_t            split_bytes)
{
    // a split product of an earlier call gave up a wait: report it here
    if(rocblas_status st = take_async_error(handle); st != rocblas_status_success)
        return st;
    BatchPtr<const T> Asrc = A;
    int               ldas = lda;
    if(scratch)
    {
        BatchPtr<T> W{nullptr, scratch, int64_t(n) * n, 0};
        rocblas_status st = copy_launch<T>(handle, n, n, A, lda, W, n, batch_count);
        if(st != rocblas_status_success)
            return st;
        Asrc = BatchPtr<const T>{nullptr, W.base, W.stride, 0};
        ldas = n;
    }

    const bool      upper = uplo == rocblas_fill_upper;
    TrtriDiagArgs<T> d{};
    d.A     = Asrc;
    d.lda   = ldas;
    d.X     = invA;
    d.ldx   = ldinvA;
    d.n     = n;
    d.nb    = n <= 16 ? 16 : kTrtriNB;
    d.upper = upper;
    d.unit  = diag == rocblas_diagonal_unit;
    d.batch = batch_count;
    rocblas_status st = trtri_diag_launch<T>(handle, d);
    if(st != rocblas_status_success || n <= d.nb)
        return st;

    for(int64_t s = d.nb; s < n; s *= 2)
    {
        const int64_t span   = 2 * s;
        const int64_t full   = n / span;              // groups with n2 == s
        const int64_t r_last = full * span;           // start of a shorter group
        const int64_t n2last = n - r_last - s;        // its n2 (> 0 if it exists)
        if(full > 0)
        {
            if(batch_count == 1 && full > 1 && !Asrc.arr && !invA.arr)
                st = trtri_double_step<T>(handle, upper, 0, int(s), int(s), int(full), span, Asrc,
                                          ldas, invA, ldinvA, 1, split_ws, split_bytes);
            else
                for(int64_t gi = 0; gi < full && st == rocblas_status_success; ++gi)
                    st = trtri_double_step<T>(handle, upper, int(gi * span), int(s), int(s), 1,
                                              span, Asrc, ldas, invA, ldinvA, batch_count,
                                              split_ws, split_bytes);
            if(st != rocblas_status_success)
                return st;
        }
        if(n2last > 0)
        {
            st = trtri_double_step<T>(handle, upper, int(r_last), int(s), int(n2last), 1, span,
                                      Asrc, ldas, invA, ldinvA, batch_count, split_ws,
                                      split_bytes);
            if(st != rocblas_status_success)
                return st;
        }
    }
    return trtri_zero_opposite_launch<T>(handle, d);
}

template rocblas_status trtri_core<float>(rocblas_handle, rocblas_fill, rocblas_diagonal, int,
                                          BatchPtr<const float>, int, BatchPtr<float>, int, int,
                                          float*, void*, size_t);
template rocblas_status trtri_core<double>(rocblas_handle, rocblas_fill, rocblas_diagonal, int,
                                           BatchPtr<const double>, int, BatchPtr<double>, int,
                                           int, double*, void*, size_t);
template size_t trtri_split_bytes<float>(rocblas_handle, int, int, bool);
template size_t trtri_split_bytes<double>(rocblas_handle, int, int, bool);

} // namespace rbamd

using namespace rbamd;

#define RBAMD_TRY_BEGIN try {
#define RBAMD_TRY_END                         \
    }                                         \
    catch(const std::bad_alloc&)              \
    {                                         \
        return rocblas_status_memory_error;   \
    }                                         \
    catch(...)                                \
    {                                         \
        return rocblas_status_internal_error; \
    }

// rocblas-functions.h:19081-19097; library/src/blas3/rocblas_trtri.cpp:42-175
#define RBAMD_TRTRI_PLAIN(NAME, T)                                                            \
    rocblas_status NAME(rocblas_handle   handle,                                              \
                        rocblas_fill     uplo,                                                \
                        rocblas_diagonal diag,                                                \
                        rocblas_int      n,                                                   \
                        const T*         A,                                                   \
                        rocblas_int      lda,                                                 \
                        T*               invA,                                                \
                        rocblas_int      ldinvA)                                              \
    {                                                                                         \
        RBAMD_TRY_BEGIN                                                                       \
        return trtri_entry<T>(handle, uplo, diag, n, BatchPtr<const T>{nullptr, A, 0, 0},     \
                              A == nullptr, lda, BatchPtr<T>{nullptr, invA, 0, 0},            \
                              invA == nullptr, ldinvA, 1,                                     \
                              A != nullptr && (const void*)A == (const void*)invA);           \
        RBAMD_TRY_END                                                                         \
    }

// rocblas-functions.h:19159-19180; library/src/blas3/rocblas_trtri_batched.cpp
#define RBAMD_TRTRI_BATCHED(NAME, T)                                                          \
    rocblas_status NAME(rocblas_handle   handle,                                              \
                        rocblas_fill     uplo,                                                \
                        rocblas_diagonal diag,                                                \
                        rocblas_int      n,                                                   \
                        const T* const   A[],                                                 \
                        rocblas_int      lda,                                                 \
                        T* const         invA[],                                              \
                        rocblas_int      ldinvA,                                              \
                        rocblas_int      batch_count)                                         \
    {                                                                                         \
        RBAMD_TRY_BEGIN                                                                       \
        return trtri_entry<T>(handle, uplo, diag, n, BatchPtr<const T>{A, nullptr, 0, 0},     \
                              A == nullptr, lda, BatchPtr<T>{invA, nullptr, 0, 0},            \
                              invA == nullptr, ldinvA, batch_count,                           \
                              A != nullptr && (const void*)A == (const void*)invA);           \
        RBAMD_TRY_END                                                                         \
    }

// rocblas-functions.h:19249-19272; library/src/blas3/rocblas_trtri_strided_batched.cpp
#define RBAMD_TRTRI_STRIDED(NAME, T)                                                          \
    rocblas_status NAME(rocblas_handle   handle,                                              \
                        rocblas_fill     uplo,                                                \
                        rocblas_diagonal diag,                                                \
                        rocblas_int      n,                                                   \
                        const T*         A,                                                   \
                        rocblas_int      lda,                                                 \
                        rocblas_stride   stride_a,                                            \
                        T*               invA,                                                \
                        rocblas_int      ldinvA,                                              \
                        rocblas_stride   stride_invA,                                         \
                        rocblas_int      batch_count)                                         \
    {                                                                                         \
        RBAMD_TRY_BEGIN                                                                       \
        return trtri_entry<T>(handle, uplo, diag, n,                                          \
                              BatchPtr<const T>{nullptr, A, stride_a, 0}, A == nullptr, lda,  \
                              BatchPtr<T>{nullptr, invA, stride_invA, 0}, invA == nullptr,    \
                              ldinvA, batch_count,                                            \
                              A != nullptr && (const void*)A == (const void*)invA);           \
        RBAMD_TRY_END                                                                         \
    }

extern "C" {
RBAMD_TRTRI_PLAIN(rocblas_strtri, float)
RBAMD_TRTRI_PLAIN(rocblas_dtrtri, double)
RBAMD_TRTRI_BATCHED(rocblas_strtri_batched, float)
RBAMD_TRTRI_BATCHED(rocblas_dtrtri_batched, double)
RBAMD_TRTRI_STRIDED(rocblas_strtri_strided_batched, float)
RBAMD_TRTRI_STRIDED(rocblas_dtrtri_strided_batched, double)
}
