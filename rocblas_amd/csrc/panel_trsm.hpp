// panel_trsm.hpp -- the panel solve shared by the single-launch k <= 512
// path (panel_trsm.hip) and the persistent variable-size batch (vbatched.hip).
//
// panel_solve() solves 16 right-hand-side vectors [v0, v0 + 16) of one
// instance with all 256 threads of the workgroup.  Canonical form as in
// small_trsm.hpp: forward substitution L y = c on the canonical lower triangle
// (side L: vectors are columns of B, M = op(A); side R: vectors are rows of B,
// M = op(A)^T; SF_REV reverses the index).  The triangle is walked in 64-row
// blocks I, left-looking:
//
//   c_I = scale * b_I - sum_{J<I} L_IJ y_J    (FP64/FP32 MFMA, y_J from LDS)
//   y_I = L_II^{-1} c_I                       (substitution, DPP multiply-adds)
//
// Four wavefronts: in the update each owns 16 rows of the block (one 16x16
// MFMA tile over the panel's 16 vectors); in the substitution each owns 4
// vectors, a 16-lane row per vector, lane g holding rows g, g+16, g+32, g+48,
// and every update of row i by the solved value of row j is one
// v_fmac_dpp that broadcasts it from lane j%16 (row_newbcast).  Tiles of L
// are staged through LDS with coalesced
// loads (either storage order of A), the next tile's loads in flight during
// the current tile's MFMAs.
#pragma once

#include "device_common.hpp"

namespace rbamd
{
namespace panel
{
constexpr int PNB   = 64;       // rows per block
constexpr int PNC   = 16;       // vectors per panel
constexpr int PLDA  = PNB + 1;  // LDS row stride of a staged tile
constexpr int PLDY  = PNC + 1;  // LDS row stride of the solved panel
constexpr int PNT   = 256;      // threads
constexpr int PEPT  = PNB * PNB / PNT; // tile elements per thread

// Fused substitution: Ln holds -L of the diagonal block with
// its diagonal and upper part zero, rd[r] = 1 / L_ii of this lane's rows
// i = g + 16 r.  Step J: the owner lane (g == J % 16) holds
// x_J = c_J * rd_J in register J / 16, and each update is one
// v_fmac_{f64,f32}_dpp broadcasting x_J from that lane:
// c[r] += x_J * (-L[g + 16 r][J]).  Rows <= J are unchanged (their entries of
// Ln are zero), so after the last step row i still holds c_i and
// y_i = c_i * rd_i, which is x_i exactly as the step that owned row i computed it.
//
// The inline-assembly updates are scheduling barriers, so the LDS reads of
// column J are issued by hand FUSED_AHEAD steps early, into a ring of
// FUSED_AHEAD + 1 register quadruples.
constexpr int FUSED_AHEAD = 4;
constexpr int FUSED_RING  = FUSED_AHEAD + 1;

template <typename T, int J>
__device__ __forceinline__ void fused_load(T (&l)[4], const T* Ln, int g)
{
#pragma unroll
    for(int r = 0; r < 4; ++r)
        l[r] = Ln[(g + 16 * r) * PLDA + J];
}

template <typename T, int J>
struct FusedSolve
{
    static __device__ __forceinline__ void run(T (&c)[4], const T (&rd)[4], const T* Ln, int g,
                                               int nrows, T (&lb)[FUSED_RING][4])
    {
        if constexpr(J < PNB)
        {
            // rows past the block's last real row hold c = 0: their steps
            // would add exact zeros (uniform branch)
            if(J >= nrows)
                return;
            if constexpr(J + FUSED_AHEAD < PNB)
                fused_load<T, J + FUSED_AHEAD>(lb[(J + FUSED_AHEAD) % FUSED_RING], Ln, g);
            constexpr int q  = J / 16;
            constexpr int jl = J % 16;
            const T       x  = c[q] * rd[q];
            T(&l)[4]         = lb[J % FUSED_RING];
            // x was written just before the step's first update; the solve's
            // very first update also covers a possible EXEC write
            fmac_bcast<jl, J == 0 ? 5 : 2>(c[0], x, l[0]);
            fmac_bcast<jl, 0>(c[1], x, l[1]);
            fmac_bcast<jl, 0>(c[2], x, l[2]);
            fmac_bcast<jl, 0>(c[3], x, l[3]);
            FusedSolve<T, J + 1>::run(c, rd, Ln, g, nrows, lb);
        }
    }
};

// nrows: real rows of the block (steps nrows..63 are skipped)
template <typename T>
__device__ __forceinline__ void fused_solve(T (&c)[4], const T (&rd)[4], const T* Ln, int g,
                                            int nrows)
{
    T lb[FUSED_RING][4];
    fused_load<T, 0>(lb[0], Ln, g);
    fused_load<T, 1>(lb[1], Ln, g);
    fused_load<T, 2>(lb[2], Ln, g);
    fused_load<T, 3>(lb[3], Ln, g);
    static_assert(FUSED_AHEAD == 4, "prologue loads FUSED_AHEAD columns");
    FusedSolve<T, 0>::run(c, rd, Ln, g, nrows, lb);
}

// LDS bytes of one workgroup solving triangles of size up to K
template <typename T>
__host__ __device__ constexpr size_t lds_bytes(int K)
{
    return sizeof(T) * (size_t(PNB) * PLDA + size_t(PNB) * PNC + PNB + size_t(K) * PLDY);
}

// B := 0 for the panel's vectors (alpha == 0; A is not referenced)
template <typename T>
__device__ void panel_zero(T* Bp, int K, int nvec, int ldb, int flags, int v0)
{
    const bool rev  = flags & SF_REV;
    const bool cols = flags & SF_COLS;
    for(int e = threadIdx.x; e < K * PNC; e += PNT)
    {
        const int i = e % K, v = v0 + e / K;
        const int r = rev ? K - 1 - i : i;
        if(v < nvec)
            gstore(Bp + (cols ? int64_t(r) + int64_t(v) * ldb : int64_t(v) + int64_t(r) * ldb),
                   T(0));
    }
}

template <typename T>
__device__ void panel_solve(const T* __restrict__ Ap, T* Bp, int K, int nvec, int lda, int ldb,
                            int flags, T scale, int v0, unsigned char* smem)
{
    T* Ls    = reinterpret_cast<T*>(smem); // PNB x PLDA staged tile
    T* Cs    = Ls + PNB * PLDA;            // PNB x PNC accumulated update
    T* rdiag = Cs + PNB * PNC;             // PNB reciprocal diagonal
    T* Ys    = rdiag + PNB;                // K x PLDY solved panel

    const int  tid  = threadIdx.x;
    const int  lane = tid & (WAVE - 1);
    const int  w    = tid / WAVE;
    const bool tA   = flags & SF_TA;
    const bool rev  = flags & SF_REV;
    const bool unit = flags & SF_UNIT;
    const bool cols = flags & SF_COLS;
    const int  nblk = (K + PNB - 1) / PNB;

    // substitution layout: vector vl = 4w + lane/16, rows g + 16r of a block
    const int  g   = lane & 15;
    const int  vl  = 4 * w + (lane >> 4);
    const int  vec = v0 + vl;
    const bool vok = vec < nvec;

    auto boff = [&](int i, int v) -> int64_t {
        const int r = rev ? K - 1 - i : i;
        return cols ? int64_t(r) + int64_t(v) * ldb : int64_t(v) + int64_t(r) * ldb;
    };

    // Tile (I, J) of canonical L, element s of this thread: canonical
    // (i, j) chosen so that consecutive threads read consecutive addresses.
    auto tile_ij = [&](int s, int& i, int& j) {
        const int e = tid + PNT * s;
        if(!tA)
        {
            i = e % PNB; // contiguous real rows (reversed when rev)
            j = e / PNB;
        }
        else
        {
            j = e % PNB; // A^T: contiguous along canonical columns
            i = e / PNB;
        }
    };
    // raw values of the tile in flight; the triangle mask is applied when the
    // tile is written to LDS, so no load is predicated on it
    T    reg[PEPT];
    int  tI = 0, tJ = 0;
    auto load_tile = [&](int I, int J) {
        tI = I;
        tJ = J;
#pragma unroll
        for(int s = 0; s < PEPT; ++s)
        {
            int i, j;
            tile_ij(s, i, j);
            const int ci = min(I * PNB + i, K - 1), cj = min(J * PNB + j, K - 1);
            const int mi = rev ? K - 1 - ci : ci, mj = rev ? K - 1 - cj : cj;
            reg[s] = gload(Ap + (tA ? int64_t(mj) + int64_t(mi) * lda
                                    : int64_t(mi) + int64_t(mj) * lda));
        }
    };
    auto store_tile = [&]() {
#pragma unroll
        for(int s = 0; s < PEPT; ++s)
        {
            int i, j;
            tile_ij(s, i, j);
            const int  gi = tI * PNB + i, gj = tJ * PNB + j;
            const bool in = gi < K && gj <= gi && !(unit && gi == gj);
            Ls[i * PLDA + j] = in ? reg[s] : T(0);
        }
    };
    // the diagonal block: -L strictly below the diagonal (the fused
    // substitution's operand), the diagonal into rdiag
    auto store_diag = [&]() {
#pragma unroll
        for(int s = 0; s < PEPT; ++s)
        {
            int i, j;
            tile_ij(s, i, j);
            const int gi = tI * PNB + i;
            Ls[i * PLDA + j] = (gi < K && j < i) ? -reg[s] : T(0);
            if(i == j)
                rdiag[i] = (gi < K && !unit) ? reg[s] : T(1);
        }
    };

    // tile sequence: (0,0), (1,0), (1,1), (2,0), ... ; (I, I) is the diagonal block
    load_tile(0, 0);
    for(int I = 0; I < nblk; ++I)
    {
        // right-hand sides of this block, consumed after the accumulation
        T bv[4];
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const int gi = I * PNB + g + 16 * r;
            bv[r]        = gload(Bp + ((vok && gi < K) ? boff(gi, vec) : 0));
        }
        // ---- c_I accumulation: sum_{J<I} L_IJ y_J on MFMA (two chains) ----
        typename Mfma<T>::acc_t acc0{0, 0, 0, 0}, acc1{0, 0, 0, 0};
        for(int J = 0; J < I; ++J)
        {
            __syncthreads(); // previous readers of Ls are done
            store_tile();
            __syncthreads();
            load_tile(I, J + 1); // next: (I, J+1), the diagonal block when J+1 == I
#pragma unroll
            for(int k4 = 0; k4 < PNB / 4; k4 += 2)
            {
                const int kk0 = 4 * k4 + (lane >> 4);
                const int kk1 = kk0 + 4;
                const T   a0  = Ls[(16 * w + (lane & 15)) * PLDA + kk0];
                const T   y0  = Ys[(J * PNB + kk0) * PLDY + (lane & 15)];
                const T   a1  = Ls[(16 * w + (lane & 15)) * PLDA + kk1];
                const T   y1  = Ys[(J * PNB + kk1) * PLDY + (lane & 15)];
                acc0          = Mfma<T>::run(a0, y0, acc0);
                acc1          = Mfma<T>::run(a1, y1, acc1);
            }
        }
        __syncthreads();
        // diagonal block (I, I): -L below the diagonal, reciprocal diagonal
        store_diag();
#pragma unroll
        for(int r = 0; r < 4; ++r)
            Cs[(16 * w + Mfma<T>::row(lane, r)) * PNC + (lane & 15)] = acc0[r] + acc1[r];
        __syncthreads();
        if(tid < PNB)
            rdiag[tid] = T(1) / rdiag[tid];
        __syncthreads();
        // prefetch the first tile of the next block during the substitution
        if(I + 1 < nblk)
            load_tile(I + 1, 0);

        // ---- y_I = L_II^{-1} c_I ----
        T c[4], rd[4];
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const int  gi = I * PNB + g + 16 * r;
            const bool ok = vok && gi < K;
            c[r]          = ok ? scale * bv[r] - Cs[(g + 16 * r) * PNC + vl] : T(0);
            rd[r]         = rdiag[g + 16 * r];
        }
        // every operand of the DPP updates is final here (an empty asm that
        // redefines it), so no VALU instruction producing one of them is
        // sunk in front of its first DPP reader
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            asm volatile("" : "+v"(c[r]));
            asm volatile("" : "+v"(rd[r]));
        }
        fused_solve<T>(c, rd, Ls, g, min(PNB, K - I * PNB));
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const int gi = I * PNB + g + 16 * r;
            const T   y  = c[r] * rd[r];
            if(gi < K)
            {
                Ys[gi * PLDY + vl] = y;
                if(vok)
                    gstore(Bp + boff(gi, vec), y);
            }
        }
    }
    __syncthreads(); // LDS reuse by the caller's next panel
}
} // namespace panel
} // namespace rbamd
