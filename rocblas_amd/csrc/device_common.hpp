// device_common.hpp -- shared device-side definitions: MFMA wrappers, kernel
// argument blocks, and the launch interfaces between the driver and the
// kernel translation units.
#pragma once

#include "internal.hpp"

#include <algorithm>

namespace rbamd
{
constexpr int WAVE = 64;

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float  f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

// Operand layout of both 16x16x4 forms: lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  Result layout differs (measured on gfx950): element r of the
// f64 result in lane l is D[l/16 + 4r][l%16]; of the f32 result, D[4(l/16) + r][l%16].
template <>
struct Mfma<double>
{
    using acc_t = d4_t;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int r)
    {
        return (lane >> 4) + 4 * r;
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
    static __device__ __forceinline__ int row(int lane, int r)
    {
        return 4 * (lane >> 4) + r;
    }
};

// Wave-uniform copies of a value every lane holds (the compiler then keeps
// it, and what is derived from it, in scalar registers)
template <typename P>
__device__ __forceinline__ P* uniform_ptr(P* p)
{
    const uint64_t v  = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return reinterpret_cast<P*>((uint64_t(hi) << 32) | lo);
}

__device__ __forceinline__ int uniform_int(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ double uniform_val(double x)
{
    const uint64_t v  = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
    return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

__device__ __forceinline__ float uniform_val(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, x)));
}

// Explicit global-address-space access: pointers read from LDS tables are
// generic, and generic (flat) memory instructions also count against the LDS
// wait counter, which would serialise LDS waits behind HBM loads.
template <typename T>
__device__ __forceinline__ T gload(const T* p)
{
    return *(const __attribute__((address_space(1))) T*)(p);
}

template <typename T>
__device__ __forceinline__ void gstore(T* p, T v)
{
    *(__attribute__((address_space(1))) T*)(p) = v;
}

// ---------------------------------------------------------------------------
// Row broadcasts (row_newbcast: lane N of each 16-lane row to the whole row)
// as inline assembly, shared by the small solve and the panel substitutions.
// ---------------------------------------------------------------------------
// Broadcast of lane N of the row, as ordered inline assembly: the compiler
// can neither hoist it (the diagonal broadcasts would all be computed up
// front and held in registers) nor merge it.
template <int N>
__device__ __forceinline__ double row_bcast_ordered(double x)
{
    double r;
    asm volatile("v_mov_b64_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "=v"(r)
                 : "v"(x), "n"(N));
    return r;
}

template <int N>
__device__ __forceinline__ float row_bcast_ordered(float x)
{
    float r;
    asm volatile("v_mov_b32_dpp %0, %1 row_newbcast:%2 row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "=v"(r)
                 : "v"(x), "n"(N));
    return r;
}

// c += (lane N of the row's l) * x in one instruction.  W > 0 pads W wait
// states first: a DPP instruction reading a VGPR that one of the two
// preceding VALU instructions wrote needs two (five after an EXEC write), and
// the compiler does not insert them in front of inline assembly.
template <int N, int W>
__device__ __forceinline__ void fmac_bcast(double& c, double l, double x)
{
    if constexpr(W > 0)
        asm volatile("s_nop %4\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf "
                     "bank_mask:0xf bound_ctrl:1"
                     : "+v"(c)
                     : "v"(l), "v"(x), "n"(N), "n"(W - 1));
    else
        asm volatile(
            "v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
            : "+v"(c)
            : "v"(l), "v"(x), "n"(N));
}

template <int N, int W>
__device__ __forceinline__ void fmac_bcast(float& c, float l, float x)
{
    if constexpr(W > 0)
        asm volatile("s_nop %4\n\tv_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf "
                     "bank_mask:0xf bound_ctrl:1"
                     : "+v"(c)
                     : "v"(l), "v"(x), "n"(N), "n"(W - 1));
    else
        asm volatile(
            "v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf bound_ctrl:1"
            : "+v"(c)
            : "v"(l), "v"(x), "n"(N));
}

inline rocblas_status launch_status()
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? rocblas_status_success : hip_to_status(e);
}

inline unsigned grid_batch(int batch)
{
    return unsigned(std::min(batch, 65535));
}

// ---------------------------------------------------------------------------
// Small solve (k <= 64)
// ---------------------------------------------------------------------------
enum SmallFlags : int
{
    SF_TA    = 1,  // M[i][j] = A[j][i]
    SF_UPPER = 2,  // stored triangle of A is the upper one
    SF_REV   = 4,  // canonical index i is real index K-1-i
    SF_UNIT  = 8,  // unit diagonal: the diagonal of A is never read
    SF_COLS  = 16, // the vectors are columns of B (side left)
    SF_SCALE = 32, // multiply the right-hand side by alpha
    SF_DPP   = 64, // register-resident triangles with DPP row broadcasts (KMAX <= 32)
    SF_FULL  = 128 // DPP launch with K == KMAX and every slot holding 16 vectors
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

// Largest number of instance slots per wavefront of the LDS variant for a
// given KMAX: keeps the triangles' LDS region near 17-19 KiB.
constexpr int small_gmax(int kmax)
{
    return kmax <= 8 ? 64 : kmax <= 16 ? 16 : kmax <= 32 ? 4 : kmax <= 48 ? 2 : 1;
}

// packed canonical lower layout: element (i, j), i >= j, at colstart(j) + i - j
template <int KMAX>
__host__ __device__ constexpr int colstart(int j)
{
    return j * KMAX - j * (j - 1) / 2;
}

constexpr int slot_elems(int kmax)
{
    // odd, so consecutive slots start in different LDS banks
    return (kmax * (kmax + 1) / 2) | 1;
}

// Single-launch panel solve for k <= kPanelMaxK (panel_trsm.hip; below 64 only
// where the driver prefers it to the small kernels):
// one workgroup per 16 vectors, same SmallArgs (K, nvec, lda, ldb, flags,
// alpha); `units` is ignored (the batch count is passed separately).
constexpr int kPanelMaxK = 512;

template <typename T>
bool panel_fits(rocblas_handle h, int K);

template <typename T>
rocblas_status launch_panel(rocblas_handle h, const SmallArgs<T>& a, int batch);

// Chained single-instance solve (chain_trsm.hip): one workgroup per (64-row
// block, 16-vector panel), blocks chained through per-panel completion
// counters in the handle's chain_flags buffer (kChainFlagsBytes: one 64-bit
// counter per panel, then an error word).
constexpr int kChainMaxPanels = int(kChainFlagsBytes / 8) - 1;

template <typename T>
bool chain_fits(rocblas_handle h, int K, int nvec, int batch);

template <typename T>
rocblas_status launch_chain(rocblas_handle h, const SmallArgs<T>& a);

// Launch one small-solve grid; defined per KMAX in small_k*.hip.
template <typename T, int KMAX>
rocblas_status launch_small_kmax(rocblas_handle h, const SmallArgs<T>& a, dim3 grid, size_t lds);

// ---------------------------------------------------------------------------
// GEMM (trailing updates)
// ---------------------------------------------------------------------------
// Triangular operand masks of the GEMM (residual B - op(A) X): op(A) / op(B)
// restricted to its lower or upper triangle in global coordinates, with an
// implicit unit diagonal when GT_UNIT is set.  Tiles wholly outside the
// triangle are skipped.
enum GemmTri : int
{
    GT_A_LOWER = 1,
    GT_A_UPPER = 2,
    GT_B_LOWER = 4,
    GT_B_UPPER = 8,
    GT_UNIT    = 16
};

// Operand storage types may differ from the accumulator's: the mixed-precision
// solve multiplies an FP32 copy of A with FP64 right-hand sides on FP32 MFMA
// and accumulates into FP64 C.
template <typename TA, typename TB = TA, typename TC = TA>
struct GemmArgsT
{
    int                M, N, K;
    int                opA, opB; // 0 = none, 1 = transpose
    BatchPtr<const TA> A;
    int                lda;
    BatchPtr<const TB> B;
    int                ldb;
    BatchPtr<TC>       C;
    int                ldc;
    TC                 alpha, beta;
    const TC*          skip_if_zero; // device scalar; == 0 -> no work (alpha == 0 in trsm)
    int                tri;          // GemmTri bits, 0 = dense operands
    int                tiles_m, tiles_n;
    int                batch;
    // Split of the K range over several workgroups per tile (FP64 DMA kernel,
    // triangular operands).  split_ok: the caller allows it (the handle's
    // counters and error word are then used).  The launcher sets the rest:
    // kc = the longest K piece (0 = no split), the counters (one per tile and
    // instance, zero between launches), the error word a given-up wait
    // raises, and the wait controls.
    int                split_ok;
    int                kc;
    // balanced triangular products (gemm_dma_balanced_kernel): the workspace
    // of 2 * split_w partial tiles of 128 x 128 (nullptr: not available), its
    // capacity in tiles, and the launch's workgroup count
    double*            split_ws;
    int64_t            split_ws_tiles;
    int                split_w;
    unsigned*          split_cnt;
    unsigned*          split_err;
    int                spin_limit;
    unsigned           need_bias;
};

template <typename T>
using GemmArgs = GemmArgsT<T, T, T>;

template <typename TA, typename TB, typename TC>
rocblas_status launch_gemm(rocblas_handle h, const GemmArgsT<TA, TB, TC>& a);

// Workspace bytes an FP64 product of this shape with one triangular operand
// (tri: one GemmTri mask bit) uses when its K steps are balanced over the
// workgroups (split_ok and split_ws set); 0 if it would not be.
size_t gemm_tri_split_bytes(rocblas_handle h, int M, int N, int K, int tri, int batch);

} // namespace rbamd
