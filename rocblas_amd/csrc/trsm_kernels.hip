// trsm_kernels.hip -- gfx950 kernels and the launcher for the triangular solve.
//
// Design (this library's own; the reference's kernels are not ported):
//
//   Every one of the 8 side/uplo/trans cases is reduced to one canonical
//   problem: forward substitution  L y = c  with L lower triangular, solved
//   independently for many vectors c.
//     side L:  op(A) X = alpha B   -> the vectors are the columns of B, M = op(A)
//     side R:  X op(A) = alpha B   <=> op(A)^T X^T = alpha B^T
//                                   -> the vectors are the rows of B, M = op(A)^T
//   M[i][j] = tA ? A[j][i] : A[i][j].  If M is upper triangular, indices are
//   reversed (i -> K-1-i), which turns backward substitution into forward
//   substitution on L = P M P.
//
//   k <= 64 ("small"): one 64-lane wavefront per workgroup solves the vectors
//     of one or more instances.  The referenced triangle is read once from HBM
//     (coalesced, column by column) and scattered into LDS in the canonical
//     packed-lower layout with reciprocals on the diagonal; each lane owns one
//     vector in registers and runs a fully unrolled substitution whose L reads
//     are LDS broadcasts.  Side-L vectors are columns of B, so they are staged
//     through LDS to keep HBM accesses coalesced; side-R vectors (rows of B) are
//     read directly, already coalesced across lanes.
//
//   k > 64: recursive blocked solve on the canonical index range,
//     y1 = solve(L11, c1);  c2 -= L21 y1 (MFMA GEMM);  y2 = solve(L22, c2),
//     with leaves of at most 64 solved by the small kernel in place.  The
//     GEMM updates are large and square-ish at the top of the recursion, which
//     is where almost all of the m*n*k flops are.
//
//   alpha: with a device pointer every kernel reads alpha itself (no host
//   synchronisation, graph-capturable); alpha == 0 writes B := 0 and leaves A
//   unread (rocblas-functions.h:19357-19361).
#include "internal.hpp"

#include <algorithm>

namespace rbamd
{
constexpr int WAVE = 64;

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float  f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double>
{
    using acc_t = d4_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<float>
{
    using acc_t = f4_t;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

static inline rocblas_status launch_status()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? rocblas_status_success : hip_to_status(e);
}

static inline unsigned grid_batch(int batch)
{
    return unsigned(std::min(batch, 65535));
}

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
// Small solve: k <= KMAX <= 64, one wavefront per workgroup.
// ---------------------------------------------------------------------------
enum SmallFlags : int
{
    SF_TA     = 1,  // M[i][j] = A[j][i]
    SF_UPPER  = 2,  // stored triangle of A is the upper one
    SF_REV    = 4,  // canonical index i is real index K-1-i
    SF_UNIT   = 8,  // unit diagonal: the diagonal of A is never read
    SF_COLS   = 16, // the vectors are columns of B (side left)
    SF_SCALE  = 32  // multiply the right-hand side by alpha
};

template <typename T>
struct SmallArgs
{
    BatchPtr<const T> A;
    BatchPtr<T>       B;
    int               K;
    int               nvec;
    int               lda;
    int               ldb;
    int               vp_log2;     // vectors per slot = 1 << vp_log2
    int               chunks;      // ceil(nvec / vectors per slot)
    int64_t           units;       // batch * chunks
    int               flags;
    int               slot_stride; // elements per slot of the packed triangle in LDS
    int               bs_stride;   // elements per staged vector (side L)
    T                 alpha_host;
    const T*          alpha_dev;
};

// column q of an upper-stored triangle holds rows 0..q at offset q(q+1)/2
__device__ __forceinline__ void tri_decode_upper(int t, int& p, int& q)
{
    int qq = int((sqrtf(8.0f * float(t) + 1.0f) - 1.0f) * 0.5f);
    while(qq * (qq + 1) / 2 > t)
        --qq;
    while((qq + 1) * (qq + 2) / 2 <= t)
        ++qq;
    q = qq;
    p = t - qq * (qq + 1) / 2;
}

// packed canonical lower layout: element (i, j), i >= j, at colstart(j) + i - j
template <int KMAX>
__host__ __device__ constexpr int colstart(int j)
{
    return j * KMAX - j * (j - 1) / 2;
}

template <int KMAX>
constexpr int slot_elems()
{
    // odd, so consecutive slots start in different LDS banks
    return (KMAX * (KMAX + 1) / 2) | 1;
}

template <typename T, int KMAX>
__global__ void __launch_bounds__(WAVE) small_trsm_kernel(const SmallArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* const sm = reinterpret_cast<T*>(smem_raw);

    constexpr int BATCHLD = 16; // independent global loads in flight per lane

    const int  lane    = threadIdx.x;
    const int  VP      = 1 << a.vp_log2;
    const int  G       = WAVE >> a.vp_log2;
    const int  g       = lane >> a.vp_log2;
    const int  v       = lane & (VP - 1);
    const int  K       = a.K;
    const bool tA      = a.flags & SF_TA;
    const bool upperA  = a.flags & SF_UPPER;
    const bool rev     = a.flags & SF_REV;
    const bool unit    = a.flags & SF_UNIT;
    const bool cols    = a.flags & SF_COLS;
    const bool scale   = a.flags & SF_SCALE;
    const int  lda     = a.lda;
    const int  ldb     = a.ldb;

    const T alpha = a.alpha_dev ? *a.alpha_dev : a.alpha_host;

    const int64_t unit0   = int64_t(blockIdx.x) * G;
    const int64_t my_unit = unit0 + g;
    const bool    slot_ok = my_unit < a.units;
    const int64_t inst    = slot_ok ? my_unit / a.chunks : 0;
    const int     chunk   = slot_ok ? int(my_unit - inst * a.chunks) : 0;
    const int     vec     = chunk * VP + v;
    const bool    valid   = slot_ok && vec < a.nvec;

    if(alpha == T(0))
    {
        // B := 0 and A is not referenced (uniform branch: alpha is global)
        if(valid)
        {
            T* Bp = a.B.get(inst);
            for(int i = 0; i < K; ++i)
                Bp[cols ? (i + int64_t(vec) * ldb) : (vec + int64_t(i) * ldb)] = T(0);
        }
        return;
    }

    T* const Ls = sm;
    T* const Bs = sm + G * a.slot_stride;

    // ---- 1. referenced triangle of A -> LDS, canonical packed lower layout ----
    const int Tn    = K * (K + 1) / 2;
    const int totalA = G * Tn;
    for(int t0 = 0; t0 < totalA; t0 += WAVE * BATCHLD)
    {
        T   val[BATCHLD];
        int dst[BATCHLD];
#pragma unroll
        for(int s = 0; s < BATCHLD; ++s)
        {
            const int t = t0 + lane + WAVE * s;
            dst[s]      = -1;
            val[s]      = T(0);
            if(t < totalA)
            {
                const int     g2 = t / Tn;
                const int     tt = t - g2 * Tn;
                const int64_t u2 = unit0 + g2;
                if(u2 < a.units)
                {
                    int p, q;
                    if(upperA)
                        tri_decode_upper(tt, p, q);
                    else
                    {
                        int pp, qq;
                        tri_decode_upper(Tn - 1 - tt, pp, qq);
                        p = K - 1 - pp;
                        q = K - 1 - qq;
                    }
                    const int mi = tA ? q : p;
                    const int mj = tA ? p : q;
                    const int i  = rev ? K - 1 - mi : mi;
                    const int j  = rev ? K - 1 - mj : mj;
                    dst[s] = g2 * a.slot_stride + colstart<KMAX>(j) + (i - j);
                    if(!(unit && p == q))
                        val[s] = a.A.get(u2 / a.chunks)[p + int64_t(q) * lda];
                    if(p == q)
                        val[s] = unit ? T(1) : T(1) / val[s];
                }
            }
        }
#pragma unroll
        for(int s = 0; s < BATCHLD; ++s)
            if(dst[s] >= 0)
                Ls[dst[s]] = val[s];
    }

    // ---- 2. right-hand side vector -> registers (canonical order) ----
    T c[KMAX];
    if(cols)
    {
        const int per_slot = VP * K;
        const int totalB   = G * per_slot;
        for(int t0 = 0; t0 < totalB; t0 += WAVE * BATCHLD)
        {
            T   val[BATCHLD];
            int dst[BATCHLD];
#pragma unroll
            for(int s = 0; s < BATCHLD; ++s)
            {
                const int t = t0 + lane + WAVE * s;
                dst[s]      = -1;
                val[s]      = T(0);
                if(t < totalB)
                {
                    const int     g2  = t / per_slot;
                    const int     rem = t - g2 * per_slot;
                    const int     vv  = rem / K;
                    const int     r   = rem - vv * K;
                    const int64_t u2  = unit0 + g2;
                    if(u2 < a.units)
                    {
                        const int64_t in2  = u2 / a.chunks;
                        const int     vec2 = int(u2 - in2 * a.chunks) * VP + vv;
                        if(vec2 < a.nvec)
                        {
                            val[s] = a.B.get(in2)[r + int64_t(vec2) * ldb];
                            dst[s] = (g2 * VP + vv) * a.bs_stride + (rev ? K - 1 - r : r);
                        }
                    }
                }
            }
#pragma unroll
            for(int s = 0; s < BATCHLD; ++s)
                if(dst[s] >= 0)
                    Bs[dst[s]] = val[s];
        }
        __syncthreads();
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            c[i] = Bs[lane * a.bs_stride + i];
    }
    else
    {
        const T* Bp = valid ? a.B.get(inst) : nullptr;
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
        {
            c[i] = T(0);
            if(valid && i < K)
                c[i] = Bp[vec + int64_t(rev ? K - 1 - i : i) * ldb];
        }
        __syncthreads();
    }
    if(scale)
    {
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            c[i] *= alpha;
    }

    // ---- 3. forward substitution on the canonical lower triangle ----
    // Rows >= K (padding when K < KMAX) only ever feed rows >= K, so their
    // (uninitialised) LDS entries never reach a stored value.
    const T* L = Ls + g * a.slot_stride;
#pragma unroll
    for(int j = 0; j < KMAX; ++j)
    {
        const int cs = colstart<KMAX>(j);
        const T   x  = c[j] * L[cs];
        c[j]         = x;
#pragma unroll
        for(int i = j + 1; i < KMAX; ++i)
            c[i] = fma(-L[cs + i - j], x, c[i]);
    }

    // ---- 4. store ----
    if(cols)
    {
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            Bs[lane * a.bs_stride + i] = c[i];
        __syncthreads();
        const int per_slot = VP * K;
        const int totalB   = G * per_slot;
        for(int t = lane; t < totalB; t += WAVE)
        {
            const int     g2  = t / per_slot;
            const int     rem = t - g2 * per_slot;
            const int     vv  = rem / K;
            const int     r   = rem - vv * K;
            const int64_t u2  = unit0 + g2;
            if(u2 < a.units)
            {
                const int64_t in2  = u2 / a.chunks;
                const int     vec2 = int(u2 - in2 * a.chunks) * VP + vv;
                if(vec2 < a.nvec)
                    a.B.get(in2)[r + int64_t(vec2) * ldb]
                        = Bs[(g2 * VP + vv) * a.bs_stride + (rev ? K - 1 - r : r)];
            }
        }
    }
    else if(valid)
    {
        T* Bp = a.B.get(inst);
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            if(i < K)
                Bp[vec + int64_t(rev ? K - 1 - i : i) * ldb] = c[i];
    }
}

struct SmallPlan
{
    int    kmax;
    int    vp_log2;
    int    G;
    int    slot_stride;
    int    bs_stride;
    size_t lds_bytes;
};

template <typename T>
static int small_slot_stride(int kmax)
{
    switch(kmax)
    {
    case 8:
        return slot_elems<8>();
    case 16:
        return slot_elems<16>();
    case 32:
        return slot_elems<32>();
    case 48:
        return slot_elems<48>();
    default:
        return slot_elems<64>();
    }
}

template <typename T>
static SmallPlan plan_small(int K, int nvec, bool cols)
{
    SmallPlan s;
    s.kmax        = K <= 8 ? 8 : K <= 16 ? 16 : K <= 32 ? 32 : K <= 48 ? 48 : 64;
    s.slot_stride = small_slot_stride<T>(s.kmax);
    s.bs_stride   = s.kmax + 1;
    // LDS budget for the packed triangles of one workgroup: at most 36 KiB, and
    // at most 64 KiB including the side-left staging area
    const size_t staging = cols ? size_t(WAVE) * s.bs_stride * sizeof(T) : 0;
    const size_t budget  = std::min<size_t>(36 * 1024, 64 * 1024 - staging);
    int          gmax    = int(budget / (size_t(s.slot_stride) * sizeof(T)));
    gmax                = std::max(1, std::min(gmax, WAVE));
    int vp              = 1;
    while(vp < std::min(nvec, WAVE))
        vp <<= 1;
    int G = WAVE / vp;
    while(G > gmax)
    {
        G >>= 1;
        vp <<= 1;
    }
    s.G       = G;
    s.vp_log2 = 0;
    while((1 << s.vp_log2) < vp)
        ++s.vp_log2;
    s.lds_bytes = (size_t(G) * s.slot_stride + (cols ? size_t(WAVE) * s.bs_stride : 0)) * sizeof(T);
    return s;
}

template <typename T>
static rocblas_status launch_small(rocblas_handle h, const SmallArgs<T>& args_in, int batch)
{
    SmallArgs<T>    a    = args_in;
    const bool      cols = a.flags & SF_COLS;
    const SmallPlan s    = plan_small<T>(a.K, a.nvec, cols);
    const int       VP   = 1 << s.vp_log2;
    a.vp_log2            = s.vp_log2;
    a.chunks             = (a.nvec + VP - 1) / VP;
    a.units              = int64_t(batch) * a.chunks;
    a.slot_stride        = s.slot_stride;
    a.bs_stride          = s.bs_stride;
    const int64_t blocks = (a.units + s.G - 1) / s.G;
    if(blocks > int64_t(0x7fffffff))
        return rocblas_status_invalid_size;
    const dim3 grid{unsigned(blocks)};
    const dim3 block{unsigned(WAVE)};
    switch(s.kmax)
    {
    case 8:
        hipLaunchKernelGGL((small_trsm_kernel<T, 8>), grid, block, s.lds_bytes, h->stream, a);
        break;
    case 16:
        hipLaunchKernelGGL((small_trsm_kernel<T, 16>), grid, block, s.lds_bytes, h->stream, a);
        break;
    case 32:
        hipLaunchKernelGGL((small_trsm_kernel<T, 32>), grid, block, s.lds_bytes, h->stream, a);
        break;
    case 48:
        hipLaunchKernelGGL((small_trsm_kernel<T, 48>), grid, block, s.lds_bytes, h->stream, a);
        break;
    default:
        hipLaunchKernelGGL((small_trsm_kernel<T, 64>), grid, block, s.lds_bytes, h->stream, a);
        break;
    }
    return launch_status();
}

// ---------------------------------------------------------------------------
// MFMA GEMM:  C := alpha * op(A) * op(B) + beta * C   (column-major, batched)
//   op(A) is M x K, op(B) is K x N.  Workgroup tile BM x BN, K step BK,
//   WGM x WGN wavefronts each owning a (BM/WGM) x (BN/WGN) block of 16x16
//   MFMA tiles.  Global -> register prefetch of the next K step overlaps the
//   MFMAs of the current one; LDS is double buffered (one barrier per step).
// ---------------------------------------------------------------------------
template <typename T>
struct GemmArgs
{
    int               M, N, K;
    int               opA, opB; // 0 = none, 1 = transpose
    BatchPtr<const T> A;
    int               lda;
    BatchPtr<const T> B;
    int               ldb;
    BatchPtr<T>       C;
    int               ldc;
    T                 alpha, beta;
    const T*          skip_if_zero; // device scalar; == 0 -> no work (alpha == 0 in trsm)
    int               tiles_m, tiles_n;
    int               batch;
};

template <typename T, int BM, int BN, int BK, int WGM, int WGN>
__global__ void __launch_bounds__(WAVE * WGM * WGN) gemm_kernel(const GemmArgs<T> a)
{
    using acc_t              = typename Mfma<T>::acc_t;
    constexpr int NT         = WAVE * WGM * WGN;
    constexpr int WTM        = BM / WGM;
    constexpr int WTN        = BN / WGN;
    constexpr int TM         = WTM / 16;
    constexpr int TN         = WTN / 16;
    constexpr int LDA_S      = BM + 2;
    constexpr int LDB_S      = BN + 2;
    constexpr int EA         = BM * BK / NT;
    constexpr int EB         = BK * BN / NT;
    static_assert(EA * NT == BM * BK && EB * NT == BK * BN, "tile/threads mismatch");
    static_assert(TM >= 1 && TN >= 1 && BK % 4 == 0, "bad tile");

    __shared__ T As[2][BK * LDA_S];
    __shared__ T Bs[2][BK * LDB_S];

    if(a.skip_if_zero && *a.skip_if_zero == T(0))
        return;

    const int tid  = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w    = tid / WAVE;
    const int wm   = w % WGM;
    const int wn   = w / WGM;

    // tile index: spread consecutive tiles over the 8 XCDs' L2s in contiguous
    // runs, then group row-panels so neighbouring workgroups share A panels.
    const int total = a.tiles_m * a.tiles_n;
    int       bid   = blockIdx.x;
    {
        const int q = total / 8, r = total % 8;
        const int x = bid % 8, l = bid / 8;
        bid         = x * q + min(x, r) + l;
    }
    constexpr int GROUP = 8;
    const int     grp   = bid / (GROUP * a.tiles_n);
    const int     fm    = grp * GROUP;
    const int     gsz   = min(a.tiles_m - fm, GROUP);
    const int     tm    = fm + (bid % gsz);
    const int     tn    = (bid % (GROUP * a.tiles_n)) / gsz;
    const int     m0    = tm * BM;
    const int     n0    = tn * BN;

    for(int b = blockIdx.y; b < a.batch; b += gridDim.y)
    {
        const T* Ap = a.A.get(b);
        const T* Bp = a.B.get(b);
        T*       Cp = a.C.get(b);

        acc_t acc[TM][TN];
#pragma unroll
        for(int i = 0; i < TM; ++i)
#pragma unroll
            for(int j = 0; j < TN; ++j)
                acc[i][j] = acc_t{0, 0, 0, 0};

        T ra[EA], rb[EB];

        auto load_tiles = [&](int k0) {
#pragma unroll
            for(int s = 0; s < EA; ++s)
            {
                const int e = tid + NT * s;
                int       i, kk;
                if(a.opA == 0)
                {
                    i  = e % BM;
                    kk = e / BM;
                }
                else
                {
                    kk = e % BK;
                    i  = e / BK;
                }
                const int gi = m0 + i, gk = k0 + kk;
                T         v  = T(0);
                if(gi < a.M && gk < a.K)
                    v = a.opA == 0 ? Ap[gi + int64_t(gk) * a.lda] : Ap[gk + int64_t(gi) * a.lda];
                ra[s] = v;
            }
#pragma unroll
            for(int s = 0; s < EB; ++s)
            {
                const int e = tid + NT * s;
                int       j, kk;
                if(a.opB == 0)
                {
                    kk = e % BK;
                    j  = e / BK;
                }
                else
                {
                    j  = e % BN;
                    kk = e / BN;
                }
                const int gj = n0 + j, gk = k0 + kk;
                T         v  = T(0);
                if(gj < a.N && gk < a.K)
                    v = a.opB == 0 ? Bp[gk + int64_t(gj) * a.ldb] : Bp[gj + int64_t(gk) * a.ldb];
                rb[s] = v;
            }
        };
        auto store_tiles = [&](int buf) {
#pragma unroll
            for(int s = 0; s < EA; ++s)
            {
                const int e = tid + NT * s;
                int       i, kk;
                if(a.opA == 0)
                {
                    i  = e % BM;
                    kk = e / BM;
                }
                else
                {
                    kk = e % BK;
                    i  = e / BK;
                }
                As[buf][kk * LDA_S + i] = ra[s];
            }
#pragma unroll
            for(int s = 0; s < EB; ++s)
            {
                const int e = tid + NT * s;
                int       j, kk;
                if(a.opB == 0)
                {
                    kk = e % BK;
                    j  = e / BK;
                }
                else
                {
                    j  = e % BN;
                    kk = e / BN;
                }
                Bs[buf][kk * LDB_S + j] = rb[s];
            }
        };

        const int ktiles = (a.K + BK - 1) / BK;
        load_tiles(0);
        for(int kt = 0; kt < ktiles; ++kt)
        {
            const int buf = kt & 1;
            store_tiles(buf);
            __syncthreads();
            if(kt + 1 < ktiles)
                load_tiles((kt + 1) * BK);
            const T* as = As[buf];
            const T* bs = Bs[buf];
#pragma unroll
            for(int k4 = 0; k4 < BK / 4; ++k4)
            {
                const int kr = k4 * 4 + lane / 16;
                T         af[TM], bf[TN];
#pragma unroll
                for(int i = 0; i < TM; ++i)
                    af[i] = as[kr * LDA_S + wm * WTM + i * 16 + (lane & 15)];
#pragma unroll
                for(int j = 0; j < TN; ++j)
                    bf[j] = bs[kr * LDB_S + wn * WTN + j * 16 + (lane & 15)];
#pragma unroll
                for(int i = 0; i < TM; ++i)
#pragma unroll
                    for(int j = 0; j < TN; ++j)
                        acc[i][j] = Mfma<T>::run(af[i], bf[j], acc[i][j]);
            }
        }

        // epilogue: lane holds rows 4*(lane/16)+r, column lane%16 of each tile
#pragma unroll
        for(int i = 0; i < TM; ++i)
#pragma unroll
            for(int j = 0; j < TN; ++j)
            {
                const int col = n0 + wn * WTN + j * 16 + (lane & 15);
                if(col >= a.N)
                    continue;
#pragma unroll
                for(int r = 0; r < 4; ++r)
                {
                    const int row = m0 + wm * WTM + i * 16 + 4 * (lane / 16) + r;
                    if(row < a.M)
                    {
                        T& cref = Cp[row + int64_t(col) * a.ldc];
                        T  v    = a.alpha * acc[i][j][r];
                        if(a.beta != T(0))
                            v = fma(a.beta, cref, v);
                        cref = v;
                    }
                }
            }
        __syncthreads(); // LDS reuse across batch iterations
    }
}

template <typename T, int BM, int BN, int BK, int WGM, int WGN>
static rocblas_status launch_gemm_cfg(rocblas_handle h, GemmArgs<T> a)
{
    a.tiles_m         = (a.M + BM - 1) / BM;
    a.tiles_n         = (a.N + BN - 1) / BN;
    const int64_t nt  = int64_t(a.tiles_m) * a.tiles_n;
    if(nt > int64_t(0x7fffffff))
        return rocblas_status_invalid_size;
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, BK, WGM, WGN>), dim3(unsigned(nt), grid_batch(a.batch)),
                       dim3(WAVE * WGM * WGN), 0, h->stream, a);
    return launch_status();
}

// Pick the largest tile that still gives every CU work.
template <typename T>
static rocblas_status launch_gemm(rocblas_handle h, const GemmArgs<T>& a)
{
    if(a.M <= 0 || a.N <= 0)
        return rocblas_status_success;
    const int64_t cus   = h->num_cus;
    auto          tiles = [&](int bm, int bn) {
        return int64_t((a.M + bm - 1) / bm) * ((a.N + bn - 1) / bn) * a.batch;
    };
    if(a.N <= 32)
        return tiles(128, 32) >= cus ? launch_gemm_cfg<T, 128, 32, 16, 4, 1>(h, a)
                                     : launch_gemm_cfg<T, 32, 32, 16, 2, 2>(h, a);
    if(a.M <= 32)
        return tiles(32, 128) >= cus ? launch_gemm_cfg<T, 32, 128, 16, 1, 4>(h, a)
                                     : launch_gemm_cfg<T, 32, 32, 16, 2, 2>(h, a);
    if(tiles(128, 128) >= cus)
        return launch_gemm_cfg<T, 128, 128, 16, 2, 2>(h, a);
    if(tiles(64, 64) >= cus)
        return launch_gemm_cfg<T, 64, 64, 16, 2, 2>(h, a);
    return launch_gemm_cfg<T, 32, 32, 16, 2, 2>(h, a);
}

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------
constexpr int SMALL_K = 64; // largest k solved by the small kernel / leaf size

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
        return launch_gemm<T>(h, g);
    }

    rocblas_status solve(int a, int b)
    {
        const int K = b - a;
        if(K <= SMALL_K)
            return leaf(a, b);
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
    DeviceGuard dg(h->device);
    const bool  sideL  = p.side == rocblas_side_left;
    const bool  tA     = sideL ? (p.trans != rocblas_operation_none)
                               : (p.trans == rocblas_operation_none);
    const bool  upperA = p.uplo == rocblas_fill_upper;
    const bool  lowerM = (!upperA) != tA;
    const int   K      = sideL ? p.m : p.n;

    if(K <= SMALL_K)
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
