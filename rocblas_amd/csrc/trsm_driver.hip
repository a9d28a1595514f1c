// trsm_driver.hip -- dispatch of rocblas_?trsm onto the gfx950 kernels.
//
//   k <= 64: one launch of the small substitution kernel (small_trsm.hpp),
//            all instances and right-hand sides of the call in one grid.
//   k >  64: recursive blocked solve on the canonical index range (the
//            reduction of every side/uplo/trans case to forward substitution
//            on L = P M P is described in small_trsm.hpp):
//              y1 = solve(L11, c1);  c2 -= L21 y1 (MFMA GEMM, gemm.hip);
//              y2 = solve(L22, c2)
//            with leaves of at most 64 solved in place by the small kernel.
//            The GEMM updates at the top of the recursion are large and
//            square-ish, which is where almost all of the m*n*k flops are.
//
// alpha: with a device pointer every kernel reads alpha itself (no host
// synchronisation, graph-capturable); alpha == 0 writes B := 0 and leaves A
// unread (library/include/internal/rocblas-functions.h:19357-19361).
#include "device_common.hpp"

namespace rbamd
{
// ---------------------------------------------------------------------------
// B := 0 over a batch (alpha == 0 in host pointer mode)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) zero_kernel(int m, int n, BatchPtr<T> B, int ldb,
                                                    int batch)
{
    const int64_t total = int64_t(m) * n;
    for(int64_t b = blockIdx.y; b < batch; b += gridDim.y)
    {
        T* p = B.get(b);
        for(int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total;
            e += int64_t(gridDim.x) * 256)
        {
            int64_t j = e / m, i = e - j * m;
            p[i + j * ldb] = T(0);
        }
    }
}

template <typename T>
rocblas_status zero_matrix_launch(rocblas_handle h, int m, int n, BatchPtr<T> B, int ldb,
                                  int batch)
{
    const int64_t total  = int64_t(m) * n;
    const unsigned bx    = unsigned(std::min<int64_t>((total + 255) / 256, 4096));
    DeviceGuard    g(h->device);
    hipLaunchKernelGGL(zero_kernel<T>, dim3(bx, grid_batch(batch)), dim3(256), 0, h->stream, m,
                       n, B, ldb, batch);
    return launch_status();
}

// ---------------------------------------------------------------------------
// B := alpha * B, alpha from the host or the device; alpha == 0 gives exact 0.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) scale_kernel(int m, int n, BatchPtr<T> B, int ldb,
                                                     int batch, T alpha_host, const T* alpha_dev)
{
    const T       alpha = alpha_dev ? *alpha_dev : alpha_host;
    const int64_t total = int64_t(m) * n;
    for(int64_t b = blockIdx.y; b < batch; b += gridDim.y)
    {
        T* p = B.get(b);
        for(int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total;
            e += int64_t(gridDim.x) * 256)
        {
            int64_t j = e / m, i = e - j * m;
            T&      v = p[i + j * ldb];
            v         = alpha == T(0) ? T(0) : alpha * v;
        }
    }
}

// ---------------------------------------------------------------------------
// X := I (n x n, leading dimension ldx) over a batch, and a plain matrix copy;
// used by trtri, which solves op(A) X = I in place on invA.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) identity_kernel(int n, BatchPtr<T> X, int ldx, int batch)
{
    const int64_t total = int64_t(n) * n;
    for(int64_t b = blockIdx.y; b < batch; b += gridDim.y)
    {
        T* p = X.get(b);
        for(int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total;
            e += int64_t(gridDim.x) * 256)
        {
            int64_t j = e / n, i = e - j * n;
            p[i + j * ldx] = i == j ? T(1) : T(0);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) copy_kernel(int m, int n, BatchPtr<const T> S,
                                                    int64_t lds, BatchPtr<T> D, int64_t ldd,
                                                    int batch)
{
    const int64_t total = int64_t(m) * n;
    for(int64_t b = blockIdx.y; b < batch; b += gridDim.y)
    {
        const T* s = S.get(b);
        T*       d = D.get(b);
        for(int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < total;
            e += int64_t(gridDim.x) * 256)
        {
            int64_t j = e / m, i = e - j * m;
            d[i + j * ldd] = s[i + j * lds];
        }
    }
}

template <typename T>
rocblas_status identity_launch(rocblas_handle h, int n, BatchPtr<T> X, int ldx, int batch)
{
    const int64_t  total = int64_t(n) * n;
    const unsigned bx    = unsigned(std::min<int64_t>((total + 255) / 256, 4096));
    DeviceGuard    g(h->device);
    hipLaunchKernelGGL(identity_kernel<T>, dim3(bx, grid_batch(batch)), dim3(256), 0, h->stream,
                       n, X, ldx, batch);
    return launch_status();
}

template <typename T>
rocblas_status copy_launch(rocblas_handle h, int m, int n, BatchPtr<const T> S, int64_t lds,
                           BatchPtr<T> D, int64_t ldd, int batch)
{
    const int64_t  total = int64_t(m) * n;
    const unsigned bx    = unsigned(std::min<int64_t>((total + 255) / 256, 4096));
    DeviceGuard    g(h->device);
    hipLaunchKernelGGL(copy_kernel<T>, dim3(bx, grid_batch(batch)), dim3(256), 0, h->stream, m,
                       n, S, lds, D, ldd, batch);
    return launch_status();
}

template rocblas_status identity_launch<float>(rocblas_handle, int, BatchPtr<float>, int, int);
template rocblas_status identity_launch<double>(rocblas_handle, int, BatchPtr<double>, int, int);
template rocblas_status copy_launch<float>(rocblas_handle, int, int, BatchPtr<const float>,
                                           int64_t, BatchPtr<float>, int64_t, int);
template rocblas_status copy_launch<double>(rocblas_handle, int, int, BatchPtr<const double>,
                                            int64_t, BatchPtr<double>, int64_t, int);

// ---------------------------------------------------------------------------
// Small-solve launch plan
// ---------------------------------------------------------------------------
struct SmallPlan
{
    int    kmax;
    int    vp_log2;
    int    G;
    int    slot_stride;
    int    bs_stride;
    size_t lds_elems;
    size_t lds_bytes;
};

// Two kernel variants (small_trsm.hpp):
//   DPP  (KMAX <= 32 and more than 8 vectors per instance): four slots of
//        16 vectors per wavefront, triangles in registers; LDS only stages
//        side-L right-hand sides.
//   LDS  (otherwise): VP = next power of two >= nvec, G = 64 / VP slots per
//        wavefront capped by small_gmax(KMAX) (by widening the slots);
//        triangles in LDS.
template <typename T>
static SmallPlan plan_small(int K, int nvec, bool cols, bool& dpp)
{
    SmallPlan s;
    s.kmax        = K <= 8 ? 8 : K <= 16 ? 16 : K <= 32 ? 32 : K <= 48 ? 48 : 64;
    s.slot_stride = slot_elems(s.kmax);
    s.bs_stride   = s.kmax + 1;
    dpp           = s.kmax <= 32 && nvec > 8;
    int vp        = 1;
    while(vp < std::min(nvec, WAVE))
        vp <<= 1;
    if(dpp)
        vp = 16;
    int G = WAVE / vp;
    while(!dpp && G > small_gmax(s.kmax))
    {
        G >>= 1;
        vp <<= 1;
    }
    s.G       = G;
    s.vp_log2 = 0;
    while((1 << s.vp_log2) < vp)
        ++s.vp_log2;
    // staging: side L for the DPP variant, both sides for the LDS variant
    const size_t staging = (cols || !dpp) ? size_t(WAVE) * s.bs_stride : 0;
    s.lds_elems          = (dpp ? 0 : size_t(G) * s.slot_stride) + staging;
    s.lds_bytes          = s.lds_elems * sizeof(T);
    return s;
}

template <typename T>
static rocblas_status launch_small(rocblas_handle h, const SmallArgs<T>& args_in, int batch)
{
    SmallArgs<T>    a    = args_in;
    const bool      cols = a.flags & SF_COLS;
    bool            dpp  = false;
    const SmallPlan s    = plan_small<T>(a.K, a.nvec, cols, dpp);
    if(dpp)
        a.flags |= SF_DPP;
    const int       VP   = 1 << s.vp_log2;
    a.vp_log2            = s.vp_log2;
    a.chunks             = (a.nvec + VP - 1) / VP;
    a.units              = int64_t(batch) * a.chunks;
    a.slot_stride        = s.slot_stride;
    a.bs_stride          = s.bs_stride;
    if(dpp && a.K == s.kmax && a.nvec % VP == 0 && a.units % s.G == 0)
        a.flags |= SF_FULL;
    const int64_t blocks = (a.units + s.G - 1) / s.G;
    if(blocks > int64_t(0x7fffffff))
        return rocblas_status_invalid_size;
    const dim3 grid{unsigned(blocks)};
    switch(s.kmax)
    {
    case 8:
        return launch_small_kmax<T, 8>(h, a, grid, s.lds_bytes);
    case 16:
        return launch_small_kmax<T, 16>(h, a, grid, s.lds_bytes);
    case 32:
        return launch_small_kmax<T, 32>(h, a, grid, s.lds_bytes);
    case 48:
        return launch_small_kmax<T, 48>(h, a, grid, s.lds_bytes);
    default:
        return launch_small_kmax<T, 64>(h, a, grid, s.lds_bytes);
    }
}

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------
constexpr int SMALL_K = 64; // largest k solved by the small kernel / leaf size

// A solve with k <= SMALL_K takes the small kernels (triangle in registers or
// in one wavefront's LDS) except where the panel solve is as fast or faster:
// 48 < k, or 32 < k with at most 16 vectors.  Measured on MI355X, side L,
// strided batch of 20000 (small kernels / panel kernel behind the variable-
// size batch's planning launch):
//   k = 64, n = 16: 1000 / 416 us;  k = 48, n = 16: 380 / 377 us;
//   k = 64, n = 64: 1076 / 1097 us; k = 48, n = 64: 548 / 1027 us;
//   one instance, k = 64, n = 16: 38 / 17 us; k = 64, n = 64: 39 / 31 us.
// The choice depends on k and the vector count only, so a batched solve and
// the single solves of its instances take the same kernels (and bits).
inline bool small_kernel_preferred(int K, int nvec)
{
    return K <= 32 || (K <= 48 && nvec > 16);
}
// leaves of the recursion for right-hand sides of at least kWidePanels panels
constexpr int kWideLeafK  = 256;
constexpr int kWidePanels = 64;

template <typename T>
struct Solver
{
    rocblas_handle        h;
    const TrsmProblem<T>& p;
    int                   Kt; // total k
    bool                  sideL, tA, rev, upperA;

    // canonical range [a, b) -> first real index
    int real_lo(int a, int b) const
    {
        return rev ? Kt - b : a;
    }

    rocblas_status leaf(int a, int b)
    {
        const int     K  = b - a;
        const int     r0 = real_lo(a, b);
        SmallArgs<T>  s{};
        s.A          = p.A.shifted(int64_t(r0) + int64_t(r0) * p.lda);
        s.B          = p.B.shifted(sideL ? int64_t(r0) : int64_t(r0) * p.ldb);
        s.K          = K;
        s.nvec       = sideL ? p.n : p.m;
        s.lda        = p.lda;
        s.ldb        = p.ldb;
        s.flags      = (tA ? SF_TA : 0) | (upperA ? SF_UPPER : 0) | (rev ? SF_REV : 0)
                  | (p.diag == rocblas_diagonal_unit ? SF_UNIT : 0) | (sideL ? SF_COLS : 0);
        s.alpha_host = T(1);
        s.alpha_dev  = p.alpha_dev; // only the alpha == 0 guard
        return launch_small<T>(h, s, p.batch);
    }

    // c[I2] -= L[I2, I1] y[I1]  with canonical I2 = [a2, b2), I1 = [a1, b1)
    rocblas_status update(int a1, int b1, int a2, int b2)
    {
        const int R1 = real_lo(a1, b1), n1 = b1 - a1;
        const int R2 = real_lo(a2, b2), n2 = b2 - a2;
        GemmArgs<T> g{};
        g.alpha        = T(-1);
        g.beta         = T(1);
        g.skip_if_zero = p.alpha_dev;
        g.batch        = p.batch;
        // block of A holding M[R2, R1] (tA: A[R1, R2]^T)
        const int64_t a_off = tA ? int64_t(R1) + int64_t(R2) * p.lda
                                 : int64_t(R2) + int64_t(R1) * p.lda;
        if(sideL)
        {
            // B[R2, :] -= M[R2, R1] * B[R1, :]
            g.M   = n2;
            g.N   = p.n;
            g.K   = n1;
            g.opA = tA ? 1 : 0;
            g.A   = p.A.shifted(a_off);
            g.lda = p.lda;
            g.opB = 0;
            g.B   = BatchPtr<const T>{(const T* const*)p.B.arr, p.B.base, p.B.stride,
                                    p.B.offset + R1};
            g.ldb = p.ldb;
            g.C   = p.B.shifted(R2);
            g.ldc = p.ldb;
        }
        else
        {
            // B[:, R2] -= B[:, R1] * M[R2, R1]^T
            g.M   = p.m;
            g.N   = n2;
            g.K   = n1;
            g.opA = 0;
            g.A   = BatchPtr<const T>{(const T* const*)p.B.arr, p.B.base, p.B.stride,
                                    p.B.offset + int64_t(R1) * p.ldb};
            g.lda = p.ldb;
            g.opB = tA ? 0 : 1;
            g.B   = p.A.shifted(a_off);
            g.ldb = p.lda;
            g.C   = p.B.shifted(int64_t(R2) * p.ldb);
            g.ldc = p.ldb;
        }
        return launch_gemm(h, g);
    }

    // SMALL_K < k <= kPanelMaxK: one launch solves the whole range -- the
    // chained kernel for one instance whose leaf has at most two vector
    // panels per 64-row block (a 512-row leaf of 256 vectors: 8 x 16
    // workgroups instead of 16; with more panels the chain's waits cost more
    // than the extra workgroups gain: 2048^2 769 vs 499 us), else one
    // workgroup per panel; both give the same bits
    rocblas_status panel_leaf(int a, int b)
    {
        const int    K  = b - a;
        const int    r0 = real_lo(a, b);
        SmallArgs<T> s{};
        s.A          = p.A.shifted(int64_t(r0) + int64_t(r0) * p.lda);
        s.B          = p.B.shifted(sideL ? int64_t(r0) : int64_t(r0) * p.ldb);
        s.K          = K;
        s.nvec       = sideL ? p.n : p.m;
        s.lda        = p.lda;
        s.ldb        = p.ldb;
        s.flags      = (tA ? SF_TA : 0) | (upperA ? SF_UPPER : 0) | (rev ? SF_REV : 0)
                  | (p.diag == rocblas_diagonal_unit ? SF_UNIT : 0) | (sideL ? SF_COLS : 0);
        s.alpha_host = T(1);
        s.alpha_dev  = p.alpha_dev; // only the alpha == 0 guard
        const int blocks = (K + 63) / 64, panels = (s.nvec + 15) / 16;
        if(panels <= 2 * blocks && chain_fits<T>(h, K, s.nvec, p.batch))
            return launch_chain<T>(h, s);
        return launch_panel<T>(h, s, p.batch);
    }

    rocblas_status solve(int a, int b)
    {
        const int K = b - a;
        if(K <= SMALL_K)
            return leaf(a, b);
        // A panel leaf streams its whole triangle once per 16-vector panel;
        // with many panels the work is better spent in GEMM updates, so wide
        // right-hand sides stop at leaves of kWideLeafK rows (measured at
        // m = n = 1024..16384: 186 vs 202 us at 1024^2, 506 vs 544 us at 2048^2).
        const int64_t panels = int64_t((sideL ? p.n : p.m) + 15) / 16 * p.batch;
        if(panel_fits<T>(h, K) && (K <= kWideLeafK || panels < kWidePanels))
            return panel_leaf(a, b);
        const int nb = (K + SMALL_K - 1) / SMALL_K;
        const int K1 = ((nb + 1) / 2) * SMALL_K;
        rocblas_status st = solve(a, a + K1);
        if(st != rocblas_status_success)
            return st;
        st = update(a, a + K1, a + K1, b);
        if(st != rocblas_status_success)
            return st;
        return solve(a + K1, b);
    }
};

template <typename T>
size_t trsm_workspace_bytes(const TrsmProblem<T>&)
{
    return 0; // every path works in place on B
}

template <typename T>
rocblas_status trsm_launch(rocblas_handle h, const TrsmProblem<T>& p)
{
    // a chained launch of an earlier call gave up a wait: report it here
    if(rocblas_status st = take_async_error(h); st != rocblas_status_success)
        return st;
    DeviceGuard dg(h->device);
    const bool  sideL  = p.side == rocblas_side_left;
    const bool  tA     = sideL ? (p.trans != rocblas_operation_none)
                               : (p.trans == rocblas_operation_none);
    const bool  upperA = p.uplo == rocblas_fill_upper;
    const bool  lowerM = (!upperA) != tA;
    const int   K      = sideL ? p.m : p.n;

    if(K <= SMALL_K && small_kernel_preferred(K, sideL ? p.n : p.m))
    {
        SmallArgs<T> s{};
        s.A          = p.A;
        s.B          = p.B;
        s.K          = K;
        s.nvec       = sideL ? p.n : p.m;
        s.lda        = p.lda;
        s.ldb        = p.ldb;
        s.flags      = (tA ? SF_TA : 0) | (upperA ? SF_UPPER : 0) | (lowerM ? 0 : SF_REV)
                  | (p.diag == rocblas_diagonal_unit ? SF_UNIT : 0) | (sideL ? SF_COLS : 0)
                  | SF_SCALE;
        s.alpha_host = p.alpha_host;
        s.alpha_dev  = p.alpha_dev;
        return launch_small<T>(h, s, p.batch);
    }

    if(chain_fits<T>(h, K, sideL ? p.n : p.m, p.batch) || panel_fits<T>(h, K))
    {
        // one launch that scales by alpha itself: chained blocks for one
        // instance with few vector panels, else one workgroup per panel
        SmallArgs<T> s{};
        s.A          = p.A;
        s.B          = p.B;
        s.K          = K;
        s.nvec       = sideL ? p.n : p.m;
        s.lda        = p.lda;
        s.ldb        = p.ldb;
        s.flags      = (tA ? SF_TA : 0) | (upperA ? SF_UPPER : 0) | (lowerM ? 0 : SF_REV)
                  | (p.diag == rocblas_diagonal_unit ? SF_UNIT : 0) | (sideL ? SF_COLS : 0)
                  | SF_SCALE;
        s.alpha_host = p.alpha_host;
        s.alpha_dev  = p.alpha_dev;
        if(chain_fits<T>(h, K, s.nvec, p.batch))
            return launch_chain<T>(h, s);
        return launch_panel<T>(h, s, p.batch);
    }

    if(p.alpha_dev || p.alpha_host != T(1))
    {
        const int64_t total = int64_t(p.m) * p.n;
        const unsigned bx   = unsigned(std::min<int64_t>((total + 255) / 256, 4096));
        hipLaunchKernelGGL(scale_kernel<T>, dim3(bx, grid_batch(p.batch)), dim3(256), 0, h->stream,
                           p.m, p.n, p.B, p.ldb, p.batch, p.alpha_host, p.alpha_dev);
        rocblas_status st = launch_status();
        if(st != rocblas_status_success)
            return st;
    }
    Solver<T> s{h, p, K, sideL, tA, !lowerM, upperA};
    return s.solve(0, K);
}

template rocblas_status trsm_launch<float>(rocblas_handle, const TrsmProblem<float>&);
template rocblas_status trsm_launch<double>(rocblas_handle, const TrsmProblem<double>&);
template size_t         trsm_workspace_bytes<float>(const TrsmProblem<float>&);
template size_t         trsm_workspace_bytes<double>(const TrsmProblem<double>&);
template rocblas_status zero_matrix_launch<float>(rocblas_handle, int, int, BatchPtr<float>, int,
                                                  int);
template rocblas_status zero_matrix_launch<double>(rocblas_handle, int, int, BatchPtr<double>, int,
                                                   int);

} // namespace rbamd
