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
//   c_I = scale * b_I - sum_{J<I} L_IJ y_J    (FP64/FP32 MFMA)
//   y_I = L_II^{-1} c_I                       (substitution, DPP multiply-adds)
//
// Four wavefronts: in the update each owns 16 rows of the block (one 16x16
// MFMA tile over the panel's 16 vectors); in the substitution each owns 4
// vectors, a 16-lane row per vector, lane g holding rows g, g+16, g+32, g+48,
// and every update of row i by the solved value of row j is one
// v_fmac_dpp that broadcasts it from lane j%16 (row_newbcast).  Tiles of L
// are staged through LDS with coalesced loads (either storage order of A),
// the next tile's loads in flight during the current tile's MFMAs or
// substitution.  The solved rows y_J are not kept: each block row reads them
// back from B (where the substitution stored them), one tile ahead, so the
// workgroup's LDS does not grow with K and two workgroups share a CU (which
// hides each one's waits better than a deeper prefetch: the registers of a
// second tile in flight would not fit two waves per SIMD).
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
constexpr int MCH   = 2;        // MFMA accumulator chains of a tile update

// Fused substitution: Ln holds -L of the diagonal block with
// its diagonal and upper part zero, rd[r] = 1 / L_ii of this lane's rows
// i = g + 16 r.  Step J: the owner lane (g == J % 16) holds
// x_J = c_J * rd_J in register q = J / 16, and each update is one
// v_fmac_{f64,f32}_dpp broadcasting x_J from that lane:
// c[r] += x_J * (-L[g + 16 r][J]).  Rows <= J are unchanged (their entries of
// Ln are zero), so after the last step row i still holds c_i and
// y_i = c_i * rd_i, which is x_i exactly as the step that owned row i computed it.
// Registers r < q hold rows above J only, which are final: step J updates
// (and loads the column entries of) registers q..3 alone.
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
    for(int r = J / 16; r < 4; ++r)
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
            fmac_bcast<jl, J == 0 ? 5 : 2>(c[q], x, l[q]);
            if constexpr(q < 1)
                fmac_bcast<jl, 0>(c[1], x, l[1]);
            if constexpr(q < 2)
                fmac_bcast<jl, 0>(c[2], x, l[2]);
            if constexpr(q < 3)
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

// One 64 x 64 tile's update of a block's accumulators, sum_k L[row][k] y[k][v]
// on MFMA: lrow = the tile in LDS at (16 w + lane % 16, lane / 16), ycol = the
// solved block in LDS at (lane / 16, lane % 16) with row stride PLDY.  Each
// half of the fragments is read before its first product; the products
// alternate over MCH accumulator chains (fold_acc sums them).  The chained and panel solves both
// accumulate through this function, so they round identically.
template <typename T>
__device__ __forceinline__ void mfma_tile(typename Mfma<T>::acc_t (&acc)[MCH], const T* lrow,
                                          const T* ycol)
{
    constexpr int HALF = PNB / 8;
#pragma unroll
    for(int h = 0; h < 2; ++h)
    {
        T af[HALF], yf[HALF];
#pragma unroll
        for(int k = 0; k < HALF; ++k)
        {
            af[k] = lrow[4 * (h * HALF + k)];
            yf[k] = ycol[4 * (h * HALF + k) * PLDY];
        }
#pragma unroll
        for(int k = 0; k < HALF; ++k)
            acc[k % MCH] = Mfma<T>::run(af[k], yf[k], acc[k % MCH]);
    }
}

template <typename T>
__device__ __forceinline__ T fold_acc(const typename Mfma<T>::acc_t (&acc)[MCH], int r)
{
    static_assert(MCH == 2 || MCH == 4, "two or four chains");
    if constexpr(MCH == 2)
        return acc[0][r] + acc[1][r];
    else
        return (acc[0][r] + acc[1][r]) + (acc[2][r] + acc[3][r]);
}

// LDS bytes of one workgroup (the same for every triangle size)
template <typename T>
__host__ __device__ constexpr size_t lds_bytes(int /*K*/)
{
    return sizeof(T) * (size_t(PNB) * PLDA + size_t(PNB) * PNC + PNB + size_t(PNB) * PLDY);
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
    T* Yt    = rdiag + PNB;                // PNB x PLDY solved block y_J of the tile

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
    // (i, j) = (lane, wu + 4s) (A read as is) or (wu + 4s, lane) (A^T), so
    // that consecutive lanes read consecutive addresses.  The lane's index
    // fixes the address within a column of A's storage and the wavefront's
    // (wu + 4s) picks the column, which is wave-uniform: each element is one
    // load at a scalar column address plus the lane's fixed offset.
    const int wu      = __builtin_amdgcn_readfirstlane(w);
    auto      tile_ij = [&](int s, int& i, int& j) {
        i = tA ? wu + 4 * s : lane;
        j = tA ? lane : wu + 4 * s;
    };
    auto fold = [&](int x) { return rev ? K - 1 - x : x; };
    // Raw values of a tile in flight.  Indices past K - 1 read row (column)
    // K - 1 instead, so every load is in bounds and unpredicated; only the
    // last block row has such rows, and their results are never used.
    auto load_tile = [&](T(&reg)[PEPT], int I, int J) {
        const int p  = fold(min((tA ? J : I) * PNB + lane, K - 1)); // lane's row in A
        const int qb = (tA ? I : J) * PNB + wu;
#pragma unroll
        for(int s = 0; s < PEPT; ++s)
        {
            const int q = fold(min(qb + 4 * s, K - 1)); // the column, wave-uniform
            reg[s]      = gload(Ap + int64_t(q) * lda + p);
        }
    };
    // An off-diagonal tile (J < I) is dense and its columns are all below K;
    // rows past K (last block row) carry row K - 1's values into MFMA result
    // rows that are discarded, so it is stored without a mask.
    // (two loops, so each element's LDS address is an immediate offset from
    // one base register)
    T* const lds0       = Ls + (tA ? wu * PLDA + lane : lane * PLDA + wu); // element 0
    auto     store_tile = [&](const T(&reg)[PEPT]) {
        if(tA)
        {
#pragma unroll
            for(int s = 0; s < PEPT; ++s)
                lds0[s * 4 * PLDA] = reg[s];
        }
        else
        {
#pragma unroll
            for(int s = 0; s < PEPT; ++s)
                lds0[s * 4] = reg[s];
        }
    };
    // the diagonal block: -L strictly below the diagonal (the fused
    // substitution's operand), the diagonal into rdiag
    auto store_diag = [&](const T(&reg)[PEPT], int I) {
#pragma unroll
        for(int s = 0; s < PEPT; ++s)
        {
            int i, j;
            tile_ij(s, i, j);
            const int gi = I * PNB + i;
            Ls[i * PLDA + j] = (gi < K && j < i) ? -reg[s] : T(0);
            if(i == j)
                rdiag[i] = (gi < K && !unit) ? reg[s] : T(1);
        }
    };
    // right-hand sides of block row I; rows past K and vectors past nvec
    // read an element in bounds (row K - 1, vector nvec - 1), discarded
    // later, so the loads need no branch
    const int vclamp = min(vec, nvec - 1);
    auto      load_rhs = [&](T(&bv)[4], int I) {
#pragma unroll
        for(int r = 0; r < 4; ++r)
            bv[r] = gload(Bp + boff(min(I * PNB + g + 16 * r, K - 1), vclamp));
    };

    // y_J of block row J for the panel's vectors, read back from B: 64 rows x
    // 16 vectors, four per thread, consecutive lanes along B's contiguous
    // dimension.  Rows past K and vectors past nvec read row K - 1 / vector
    // nvec - 1 (in bounds; such values never reach a stored result).
    auto y_rv = [&](int s, int& r, int& v) {
        const int e = tid + PNT * s;
        r           = cols ? e % PNB : e / PNC;
        v           = cols ? e / PNB : e % PNC;
    };
    auto load_y = [&](T(&yv)[4], int J) {
#pragma unroll
        for(int s = 0; s < 4; ++s)
        {
            int r, v;
            y_rv(s, r, v);
            yv[s] = gload(Bp + boff(min(J * PNB + r, K - 1), min(v0 + v, nvec - 1)));
        }
    };
    auto store_y = [&](const T(&yv)[4]) {
#pragma unroll
        for(int s = 0; s < 4; ++s)
        {
            int r, v;
            y_rv(s, r, v);
            Yt[r * PLDY + v] = yv[s];
        }
    };
    static_assert(PNB * PNC == 4 * PNT, "four y values per thread");

    // The tiles are consumed in the order (0,0), (1,0), (1,1), (2,0), ...,
    // (I, I) being the diagonal block that ends block row I.  The register
    // buffer ra is refilled with the next tile as soon as its tile is in
    // LDS, so each tile's load latency overlaps the step before it is needed.
    // Every step issues exactly one tile's loads (past the last tile, the
    // last tile again, never consumed), so each path through a step leaves
    // the same number of loads in flight and the compiler's wait counts
    // retire only the tile being consumed.
    int  iI = 0, iJ = 0; // next tile to load
    auto issue = [&](T(&reg)[PEPT]) {
        const bool past = iI >= nblk;
        load_tile(reg, past ? nblk - 1 : iI, past ? nblk - 1 : iJ);
        if(!past && ++iJ > iI)
        {
            ++iI;
            iJ = 0;
        }
    };
    T ra[PEPT];
    issue(ra);
    T yv[4]; // y of the next off-diagonal tile (loaded one step ahead)

    int cI = 0, cJ = 0; // tile being consumed
    typename Mfma<T>::acc_t acc[MCH];
#pragma unroll
    for(int a = 0; a < MCH; ++a)
        acc[a] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    const T* lrow = Ls + (16 * w + (lane & 15)) * PLDA + (lane >> 4);
    const T* ycol = Yt + (lane >> 4) * PLDY + (lane & 15);

    // consume the tile held in `reg`; true once the last block row is solved
    auto step = [&](T(&reg)[PEPT]) -> bool {
        if(cJ < cI)
        {
            // ---- c_I accumulation: L_IJ y_J on MFMA ----
            __syncthreads(); // previous readers of Ls and Yt are done
            store_tile(reg);
            store_y(yv);
            __syncthreads();
            issue(reg);
            // y of the next tile: y_{J+1}, or (the diagonal tile is next) a
            // load nobody reads, so every step issues the same loads
            load_y(yv, cJ + 1 < cI ? cJ + 1 : 0);
            mfma_tile<T>(acc, lrow, ycol);
            ++cJ;
            return false;
        }
        // diagonal block (I, I): -L below the diagonal, reciprocal diagonal.
        // The block row's right-hand sides are loaded first: only the next
        // tile's loads follow them, so waiting for them drains no prefetch
        T bv[4];
        load_rhs(bv, cI);
        __syncthreads();
        store_diag(reg, cI);
#pragma unroll
        for(int r = 0; r < 4; ++r)
            Cs[(16 * w + Mfma<T>::row(lane, r)) * PNC + (lane & 15)] = fold_acc<T>(acc, r);
        __syncthreads();
        if(tid < PNB)
            rdiag[tid] = T(1) / rdiag[tid];
        __syncthreads();
        issue(reg);

        // ---- y_I = L_II^{-1} c_I ----
        T c[4], rd[4];
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const int  gi = cI * PNB + g + 16 * r;
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
        fused_solve<T>(c, rd, Ls, g, min(PNB, K - cI * PNB));
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const int gi = cI * PNB + g + 16 * r;
            if(gi < K && vok)
                gstore(Bp + boff(gi, vec), c[r] * rd[r]);
        }
        ++cI;
        cJ = 0;
        if(cI >= nblk)
            return true;
        // every wavefront's rows of y_I are in B before any is read back
        __syncthreads();
        load_y(yv, 0); // the next tile is (I + 1, 0)
#pragma unroll
        for(int a = 0; a < MCH; ++a)
            acc[a] = typename Mfma<T>::acc_t{0, 0, 0, 0};
        return false;
    };
    while(!step(ra))
    {
    }
    __syncthreads(); // LDS reuse by the caller's next panel
}
} // namespace panel
} // namespace rbamd
