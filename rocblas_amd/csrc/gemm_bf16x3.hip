// gemm_bf16x3.hip -- split-BF16x3 form of the trailing update of the
// mixed-precision solve:
//
//   C (FP64) += alpha op(A) op(B),  A and B FP32 or FP64
//
// Every operand value x is split into two BF16 numbers, hi = bf16(x) and
// lo = bf16(x - hi) (both rounded to nearest even, x taken in FP32), and the
// product is hi_a hi_b + hi_a lo_b + lo_a hi_b, three v_mfma_f32_16x16x32_bf16
// accumulated in FP32 (the lo_a lo_b term, about 2^-16 of the product, is
// dropped): about 16 significant bits per product at the BF16 matrix rate
// (2.5 PFLOP/s dense on MI355X), against 24 bits at the FP32 rate (157
// TFLOP/s) for the FP32 form.  No reference counterpart: rocBLAS's trsm
// updates are FP64 GEMMs (library/src/blas3/rocblas_trsm_kernels.hpp:1203-1325).
//
// Tiles of 128 x 128 x 32, 256 threads (2 x 2 wavefronts of 64 x 64, 4 x 4
// MFMA tiles each).  The next tile's operands are loaded into registers
// while the current one is multiplied; they are split and stored to LDS as
// four BF16 planes (A hi/lo, B hi/lo), each row of a plane holding its 32 k
// values contiguously, so one 16-byte LDS read gives a lane its eight k
// values of an MFMA operand.  MFMA operand layout of 16x16x32 (gfx950):
// lane l holds row (A) or column (B) l % 16, k = 8 (l / 16) .. 8 (l / 16) + 7;
// result element r of lane l is D[4 (l / 16) + r][l % 16].  The operands are
// passed swapped (B fragment first), so the MFMA computes the transposed tile
// and a lane's results run down a column of C (coalesced FP64 stores).
#include "device_common.hpp"
#include "gemm_bf16x3.hpp"

namespace rbamd
{
namespace
{
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float  f4v_t __attribute__((ext_vector_type(4)));

constexpr int XB_M  = 128;
constexpr int XB_N  = 128;
constexpr int XB_K  = 32;
constexpr int XB_NT = 256;
constexpr int XB_LD = XB_K + 8; // bf16 per LDS row: 80 bytes, 16-byte aligned rows
constexpr int XB_EA = XB_M * XB_K / XB_NT; // operand elements per thread per tile
constexpr int XB_EB = XB_N * XB_K / XB_NT;

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo)
{
    hi = __bf16(x);
    lo = __bf16(x - float(hi));
}

template <typename TA, typename TB, int OPA, int OPB>
__global__ void __launch_bounds__(XB_NT) __attribute__((amdgpu_waves_per_eu(2)))
    gemm_bf16x3_kernel(const GemmArgsT<TA, TB, double> a)
{
    __shared__ __align__(16) __bf16 Ahi[XB_M * XB_LD];
    __shared__ __align__(16) __bf16 Alo[XB_M * XB_LD];
    __shared__ __align__(16) __bf16 Bhi[XB_N * XB_LD];
    __shared__ __align__(16) __bf16 Blo[XB_N * XB_LD];

    if(a.skip_if_zero && *a.skip_if_zero == 0.0)
        return;

    const int tid  = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w    = tid / WAVE;
    const int wm   = w % 2;
    const int wn   = w / 2;

    // XCD-contiguous runs of 8-row-panel groups (as the FP64 kernels)
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
    const int     m0    = (fm + (bid % gsz)) * XB_M;
    const int     n0    = ((bid % (GROUP * a.tiles_n)) / gsz) * XB_N;
    const int     ktiles = (a.K + XB_K - 1) / XB_K;

    // load geometry: consecutive threads read consecutive addresses
    auto ia = [&](int s) { return OPA == 0 ? tid % XB_M : tid / XB_K + s * (XB_NT / XB_K); };
    auto ka = [&](int s) { return OPA == 0 ? tid / XB_M + s * (XB_NT / XB_M) : tid % XB_K; };
    auto kb = [&](int s) { return OPB == 0 ? tid % XB_K : tid / XB_N + s * (XB_NT / XB_N); };
    auto jb = [&](int s) { return OPB == 0 ? tid / XB_K + s * (XB_NT / XB_K) : tid % XB_N; };

    for(int b = blockIdx.y; b < a.batch; b += gridDim.y)
    {
        const TA* Ap = a.A.get(b);
        const TB* Bp = a.B.get(b);
        double*   Cp = a.C.get(b);

        f4v_t acc[4][4];
#pragma unroll
        for(int i = 0; i < 4; ++i)
#pragma unroll
            for(int j = 0; j < 4; ++j)
                acc[i][j] = f4v_t{0.f, 0.f, 0.f, 0.f};

        TA  ra[XB_EA];
        TB  rb[XB_EB];
        int k_ld = 0;
        // branch-free loads of clamped addresses; bounds applied at the split
        auto load_tiles = [&](int k0) {
            k_ld = k0;
#pragma unroll
            for(int s = 0; s < XB_EA; ++s)
            {
                const int gi = min(m0 + ia(s), a.M - 1);
                const int gk = min(k0 + ka(s), a.K - 1);
                ra[s] = gload(Ap + (OPA == 0 ? gi + int64_t(gk) * a.lda : gk + int64_t(gi) * a.lda));
            }
#pragma unroll
            for(int s = 0; s < XB_EB; ++s)
            {
                const int gj = min(n0 + jb(s), a.N - 1);
                const int gk = min(k0 + kb(s), a.K - 1);
                rb[s] = gload(Bp + (OPB == 0 ? gk + int64_t(gj) * a.ldb : gj + int64_t(gk) * a.ldb));
            }
        };
        auto store_tiles = [&]() {
#pragma unroll
            for(int s = 0; s < XB_EA; ++s)
            {
                const bool in = m0 + ia(s) < a.M && k_ld + ka(s) < a.K;
                __bf16     hi, lo;
                split_bf16(in ? float(ra[s]) : 0.f, hi, lo);
                Ahi[ia(s) * XB_LD + ka(s)] = hi;
                Alo[ia(s) * XB_LD + ka(s)] = lo;
            }
#pragma unroll
            for(int s = 0; s < XB_EB; ++s)
            {
                const bool in = n0 + jb(s) < a.N && k_ld + kb(s) < a.K;
                __bf16     hi, lo;
                split_bf16(in ? float(rb[s]) : 0.f, hi, lo);
                Bhi[jb(s) * XB_LD + kb(s)] = hi;
                Blo[jb(s) * XB_LD + kb(s)] = lo;
            }
        };

        if(ktiles > 0)
            load_tiles(0);
        for(int kt = 0; kt < ktiles; ++kt)
        {
            __syncthreads(); // the previous tile's readers are done
            store_tiles();
            __syncthreads();
            // the load past the last tile reads clamped addresses and is discarded
            load_tiles((kt + 1) * XB_K);
            const int koff = 8 * (lane / 16);
            bf16x8_t  bh[4], bl[4];
#pragma unroll
            for(int j = 0; j < 4; ++j)
            {
                const int col = wn * 64 + j * 16 + (lane & 15);
                bh[j] = *reinterpret_cast<const bf16x8_t*>(&Bhi[col * XB_LD + koff]);
                bl[j] = *reinterpret_cast<const bf16x8_t*>(&Blo[col * XB_LD + koff]);
            }
#pragma unroll
            for(int i = 0; i < 4; ++i)
            {
                const int      row = wm * 64 + i * 16 + (lane & 15);
                const bf16x8_t ah  = *reinterpret_cast<const bf16x8_t*>(&Ahi[row * XB_LD + koff]);
                const bf16x8_t al  = *reinterpret_cast<const bf16x8_t*>(&Alo[row * XB_LD + koff]);
#pragma unroll
                for(int j = 0; j < 4; ++j)
                {
                    // transposed tile: first operand B (rows = columns of C)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah, acc[i][j], 0, 0, 0);
                }
            }
        }

        // epilogue: lane l of tile (i, j) holds C[row][col + 4 (l / 16) + r]
#pragma unroll
        for(int i = 0; i < 4; ++i)
        {
            const int row = m0 + wm * 64 + i * 16 + (lane & 15);
#pragma unroll
            for(int j = 0; j < 4; ++j)
#pragma unroll
                for(int r = 0; r < 4; ++r)
                {
                    const int col = n0 + wn * 64 + j * 16 + 4 * (lane / 16) + r;
                    if(row < a.M && col < a.N)
                    {
                        double& c = Cp[row + int64_t(col) * a.ldc];
                        double  v = a.alpha * double(acc[i][j][r]);
                        if(a.beta != 0.0)
                            v = fma(a.beta, c, v);
                        c = v;
                    }
                }
        }
        __syncthreads(); // LDS reuse across batch iterations
    }
}

template <typename TA, typename TB, int OPA, int OPB>
void launch_ops(rocblas_handle h, const GemmArgsT<TA, TB, double>& a, dim3 grid)
{
    hipLaunchKernelGGL((gemm_bf16x3_kernel<TA, TB, OPA, OPB>), grid, dim3(XB_NT), 0, h->stream,
                       a);
}

} // namespace

template <typename TA, typename TB>
rocblas_status launch_gemm_bf16x3(rocblas_handle h, GemmArgsT<TA, TB, double> a)
{
    if(a.M <= 0 || a.N <= 0)
        return rocblas_status_success;
    if(a.tri)
        return rocblas_status_not_implemented;
    a.tiles_m        = (a.M + XB_M - 1) / XB_M;
    a.tiles_n        = (a.N + XB_N - 1) / XB_N;
    const int64_t nt = int64_t(a.tiles_m) * a.tiles_n;
    if(nt > int64_t(0x7fffffff))
        return rocblas_status_invalid_size;
    const dim3  grid(unsigned(nt), grid_batch(a.batch));
    DeviceGuard g(h->device);
    if(a.opA == 0 && a.opB == 0)
        launch_ops<TA, TB, 0, 0>(h, a, grid);
    else if(a.opA == 1 && a.opB == 0)
        launch_ops<TA, TB, 1, 0>(h, a, grid);
    else if(a.opA == 0 && a.opB == 1)
        launch_ops<TA, TB, 0, 1>(h, a, grid);
    else
        launch_ops<TA, TB, 1, 1>(h, a, grid);
    return launch_status();
}

template rocblas_status launch_gemm_bf16x3<float, double>(rocblas_handle,
                                                          GemmArgsT<float, double, double>);
template rocblas_status launch_gemm_bf16x3<double, float>(rocblas_handle,
                                                          GemmArgsT<double, float, double>);
template rocblas_status launch_gemm_bf16x3<double, double>(rocblas_handle,
                                                           GemmArgsT<double, double, double>);

} // namespace rbamd
