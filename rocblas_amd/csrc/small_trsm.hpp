// small_trsm.hpp -- the k <= 64 triangular-solve kernel (instantiated per KMAX
// in small_k*.hip).
//
// Every side/uplo/trans case is reduced to forward substitution L y = c on a
// canonical lower triangle, one right-hand-side vector per lane:
//   side L:  op(A) X = alpha B, vectors = columns of B, M = op(A)
//   side R:  X op(A) = alpha B <=> op(A)^T X^T = alpha B^T, vectors = rows of B,
//            M = op(A)^T
//   M[i][j] = TA ? A[j][i] : A[i][j];  when M is upper triangular (REV) the
//   canonical index i is the real index KMAX-1-i, so L[i][j] = M[KMAX-1-i][KMAX-1-j].
// The matrix is embedded in a KMAX x KMAX triangle: entries outside the K x K
// problem are 0 (1 on the diagonal) and right-hand-side entries outside it are
// 0, so padded rows never change real rows -- and because the reversal is
// relative to KMAX, every LDS offset in the substitution is a compile-time
// constant for each (UPPER, TA) variant.
//
// A wavefront serves G "slots" (instances) of VP vectors each (G * VP = 64).
// Data movement, designed for one HBM round trip per wavefront:
//   1. issue every global load: right-hand sides (side L: coalesced down the
//      columns, staged through LDS so lane v can read column v; side R: one
//      row per lane, already coalesced across lanes), then the slots' stored
//      triangles in pair order (below);
//   2. write the triangles to LDS in that same stored order (no index
//      transform), diagonal replaced by its reciprocal;
//   3. fully unrolled substitution; L[i][j] is read (LDS broadcast within a
//      slot) from the stored position of A's entry;
//   4. side L: solutions staged through LDS again for coalesced stores.
//
// Pair order: within a KMAX x KMAX triangle, column c and column KMAX-1-c
// together hold exactly KMAX+1 entries, so entry t lies in pair t / (KMAX+1)
// -- constant divisions, no search -- and consecutive t walk down a column.
#pragma once

#include "device_common.hpp"

namespace rbamd
{

// Returns v unchanged through an empty asm, so comparisons against the result
// are recomputed where they are used instead of being hoisted and kept live
// (dozens of live per-row masks otherwise spill the scalar registers).
__device__ __forceinline__ int fresh(int v)
{
    asm volatile("" : "+s"(v));
    return v;
}

// Same for a per-lane value: address arithmetic derived from the result is
// recomputed instead of being shared with (and kept live from) earlier phases.
__device__ __forceinline__ int fresh_lane(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// offset of stored entry (p, q) of a KMAX triangle in pair order
template <int KMAX, bool UPPER>
__host__ __device__ constexpr int pair_off(int p, int q)
{
    const bool first = q < KMAX / 2;
    const int  c     = first ? q : KMAX - 1 - q;
    if(!UPPER) // column c: rows c..KMAX-1; column KMAX-1-c: rows KMAX-1-c..KMAX-1
        return c * (KMAX + 1) + (first ? p - q : (KMAX - c) + (p - q));
    else // column c: rows 0..c; column KMAX-1-c: rows 0..KMAX-1-c
        return c * (KMAX + 1) + (first ? p : (c + 1) + p);
}

// inverse of pair_off: entry t -> (p, q)
template <int KMAX, bool UPPER>
__host__ __device__ inline void pair_decode(int t, int& p, int& q)
{
    const int c = t / (KMAX + 1);
    const int o = t % (KMAX + 1);
    if(!UPPER)
    {
        const bool first = o < KMAX - c;
        q                = first ? c : KMAX - 1 - c;
        p                = first ? c + o : q + (o - (KMAX - c));
    }
    else
    {
        const bool first = o <= c;
        q                = first ? c : KMAX - 1 - c;
        p                = first ? o : o - (c + 1);
    }
}

// LDS offset of canonical L[i][j]
template <int KMAX, bool UPPER, bool TA>
__host__ __device__ constexpr int lofs(int i, int j)
{
    constexpr bool REV = UPPER != TA;
    const int      mi  = REV ? KMAX - 1 - i : i;
    const int      mj  = REV ? KMAX - 1 - j : j;
    return TA ? pair_off<KMAX, UPPER>(mj, mi) : pair_off<KMAX, UPPER>(mi, mj);
}

// ---------------------------------------------------------------------------
// LDS variant (any KMAX, any vectors-per-slot): the slots' triangles live in
// LDS in stored pair order and every lane reads L[i][j] as an LDS broadcast.
// Region layout: [G slots x slot_stride] triangles, then (side L) the
// 64 x bs_stride staging area for the right-hand sides.
// ---------------------------------------------------------------------------
template <typename T, int KMAX, bool UPPER, bool TA, bool COLS>
__global__ void __launch_bounds__(WAVE) small_trsm_lds_kernel(const SmallArgs<T> a)
{
    constexpr bool REV     = UPPER != TA;
    constexpr int  TRI     = KMAX * (KMAX + 1) / 2;
    constexpr int  BATCHLD = 16;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* const R = reinterpret_cast<T*>(smem_raw);
    __shared__ const T* slotA[WAVE];
    __shared__ T*       slotB[WAVE];
    __shared__ int      slotV0[WAVE];

    const int  lane  = threadIdx.x;
    const int  vpl   = a.vp_log2;
    const int  VP    = 1 << vpl;
    const int  G     = WAVE >> vpl;
    const int  g     = lane >> vpl;
    const int  v     = lane & (VP - 1);
    const int  K     = a.K;
    const bool unit  = a.flags & SF_UNIT;
    const bool scale = a.flags & SF_SCALE;
    const int  lda   = a.lda;
    const int  ldb   = a.ldb;
    const int  bs    = a.bs_stride;
    const int  S     = a.slot_stride;
    T* const   RB    = R + G * S; // side-L staging area

    const T alpha = a.alpha_dev ? *a.alpha_dev : a.alpha_host;

    if(lane < G)
    {
        const int64_t u = int64_t(blockIdx.x) * G + lane;
        if(u < a.units)
        {
            const int64_t inst = a.chunks == 1 ? u : u / a.chunks;
            slotA[lane]        = a.A.get(inst);
            slotB[lane]        = a.B.get(inst);
            slotV0[lane]       = int(u - inst * a.chunks) * VP;
        }
        else
        {
            slotA[lane]  = nullptr;
            slotB[lane]  = nullptr;
            slotV0[lane] = 0;
        }
    }
    __syncthreads();

    T* const   myB   = slotB[g];
    const int  vec   = slotV0[g] + v;
    const bool valid = myB != nullptr && vec < a.nvec;

    if(alpha == T(0))
    {
        // B := 0 and A is not referenced (uniform branch: alpha is global)
        if(valid)
            for(int i = 0; i < K; ++i)
                gstore(myB + (COLS ? (i + int64_t(vec) * ldb) : (vec + int64_t(i) * ldb)), T(0));
        return;
    }

    // Out-of-range lanes load from a valid address (slot 0 always exists) and
    // discard the value, so no load sits behind a branch.
    const T* const safeA = slotA[0];
    T* const       safeB = slotB[0];

    // ---- 1. right-hand sides -> staging area (lane v later reads vector v) ----
    // side L: coalesced down the columns; side R: lane v loads row v (already
    // coalesced across lanes).  Parking the values in LDS keeps them out of
    // registers while the triangles are loaded.
    T c[KMAX];
    for(int s0 = 0; s0 < KMAX; s0 += BATCHLD)
    {
        T   bv[BATCHLD];
        int Kc = fresh(K);
#pragma unroll
        for(int s = 0; s < BATCHLD; ++s)
        {
            if constexpr(COLS)
            {
                const int  e  = lane + WAVE * (s0 + s);
                const int  r  = e % KMAX;
                const int  v2 = e / KMAX;
                T* const   Bp = slotB[v2 >> vpl];
                const int  vc = slotV0[v2 >> vpl] + (v2 & (VP - 1));
                const bool ok = s0 + s < KMAX && Bp && r < Kc && vc < a.nvec;
                const T    x  = gload(ok ? Bp + (r + int64_t(vc) * ldb) : safeB);
                bv[s]         = ok ? x : T(0);
            }
            else
            {
                const int  r  = s0 + s;
                const bool ok = valid && r < Kc;
                const T    x  = gload(ok ? myB + (vec + int64_t(r) * ldb) : safeB);
                bv[s]         = ok ? x : T(0);
            }
        }
#pragma unroll
        for(int s = 0; s < BATCHLD; ++s)
        {
            if(s0 + s < KMAX)
            {
                if constexpr(COLS)
                {
                    const int e  = lane + WAVE * (s0 + s);
                    const int r  = e % KMAX;
                    const int v2 = e / KMAX;
                    RB[v2 * bs + (REV ? KMAX - 1 - r : r)] = bv[s];
                }
                else
                {
                    const int r = s0 + s;
                    RB[lane * bs + (REV ? KMAX - 1 - r : r)] = bv[s];
                }
            }
        }
    }

    // ---- 2. triangles -> LDS in stored pair order ----
    const int totalA = G * TRI;
    for(int t0 = 0; t0 < totalA; t0 += WAVE * BATCHLD)
    {
        T av[BATCHLD];
#pragma unroll
        for(int s = 0; s < BATCHLD; ++s)
        {
            const int e  = t0 + lane + WAVE * s;
            const int g2 = e / TRI;
            int       p, q;
            pair_decode<KMAX, UPPER>(e % TRI, p, q);
            const T*   Ap   = e < totalA ? slotA[g2] : nullptr;
            const bool load = Ap && p < K && q < K && !(unit && p == q);
            const T    x    = gload(load ? Ap + (p + int64_t(q) * lda) : safeA);
            av[s]           = load ? x : (p == q ? T(1) : T(0));
        }
#pragma unroll
        for(int s = 0; s < BATCHLD; ++s)
        {
            const int e = t0 + lane + WAVE * s;
            if(e < totalA)
                R[(e / TRI) * S + e % TRI] = av[s];
        }
    }
    __syncthreads();
#pragma unroll
    for(int i = 0; i < KMAX; ++i)
        c[i] = RB[lane * bs + i];
    if(!unit)
    {
        for(int idx = lane; idx < G * KMAX; idx += WAVE)
        {
            const int j = idx % KMAX;
            T&        d = R[(idx / KMAX) * S + pair_off<KMAX, UPPER>(j, j)];
            d           = T(1) / d;
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
    const T* L = R + g * S;
#pragma unroll
    for(int j = 0; j < KMAX; ++j)
    {
        const T x = c[j] * L[lofs<KMAX, UPPER, TA>(j, j)];
        c[j]      = x;
#pragma unroll
        for(int i = j + 1; i < KMAX; ++i)
            c[i] = fma(-L[lofs<KMAX, UPPER, TA>(i, j)], x, c[i]);
        // bound how far ahead the scheduler hoists LDS reads (registers)
        if(KMAX > 32 || j % 2 == 1)
            asm volatile("" ::: "memory");
    }

    // ---- 4. store: solutions staged through LDS again ----
    // side L: coalesced down the columns; side R: lane v stores row v.
#pragma unroll
    for(int i = 0; i < KMAX; ++i)
        RB[lane * bs + i] = c[i];
    __syncthreads();
    int       Kc = fresh(K);
    const int ln = fresh_lane(lane);
#pragma unroll
    for(int s = 0; s < KMAX; ++s)
    {
        if(s % 8 == 0)
            Kc = fresh(Kc);
        if constexpr(COLS)
        {
            const int e  = ln + WAVE * s;
            const int r  = e % KMAX;
            const int v2 = e / KMAX;
            T* const  Bp = slotB[v2 >> vpl];
            const int vc = slotV0[v2 >> vpl] + (v2 & (VP - 1));
            if(Bp && r < Kc && vc < a.nvec)
                gstore(Bp + (r + int64_t(vc) * ldb), RB[v2 * bs + (REV ? KMAX - 1 - r : r)]);
        }
        else
        {
            if(valid && s < Kc)
                gstore(myB + (vec + int64_t(s) * ldb), RB[lane * bs + (REV ? KMAX - 1 - s : s)]);
        }
    }
}

// ---------------------------------------------------------------------------
// DPP variant (KMAX <= 32; slots of 16 vectors, four slots per wavefront):
// each 16-lane row holds its slot's triangle in registers -- entry t of the
// pair order in lane t % 16, register t / 16, stored negated -- and every
// update c[i] -= L[i][j] x_j is a single v_fmac_{f64,f32}_dpp whose first
// operand is broadcast from the owning lane of the row (row_newbcast), so the
// substitution reads no LDS and spends no separate broadcast instruction.
// Lane n also holds the reciprocal diagonal entries n + 16 d.  An instance
// with more than 16 vectors occupies several slots, each loading the
// triangle (the duplicate loads are the same addresses in one instruction).
//
// Slot bookkeeping is wave-uniform: slot s is unit blockIdx.x * 4 + s, and
// its instance pointers and first vector are read into scalar registers.
// Side L loads B coalesced down the columns -- load step s covers
// 64 / KMAX whole columns of one slot -- so every address is a per-lane
// constant plus a uniform base.  FULL launches (K == KMAX, every slot holds
// 16 vectors) carry no masks at all.
// ---------------------------------------------------------------------------
constexpr int DPP_VP = 16; // vectors per slot
constexpr int DPP_G  = 4;  // slots per wavefront

template <int N>
__device__ __forceinline__ double row_bcast(double x)
{
    return __builtin_amdgcn_mov_dpp(x, 0x150 + N, 0xf, 0xf, true);
}

template <int N>
__device__ __forceinline__ float row_bcast(float x)
{
    return __builtin_amdgcn_mov_dpp(x, 0x150 + N, 0xf, 0xf, true);
}

template <typename T, int KMAX>
using DppTri = T[(KMAX * (KMAX + 1) / 2 + 15) / 16];

// column J of the substitution: c[I] -= L[I][J] x for I = J+1 .. KMAX-1
template <typename T, int KMAX, bool UPPER, bool TA, int J, int I>
struct DppColumn
{
    static __device__ __forceinline__ void run(T (&c)[KMAX], const DppTri<T, KMAX>& Lr, T x)
    {
        if constexpr(I < KMAX)
        {
            constexpr int t = lofs<KMAX, UPPER, TA>(I, J);
            // x was written just before the column's first update; the very
            // first update of the solve also covers a possible EXEC write
            constexpr int W = I == J + 1 ? (J == 0 ? 5 : 2) : 0;
            fmac_bcast<t % 16, W>(c[I], Lr[t / 16], x);
            DppColumn<T, KMAX, UPPER, TA, J, I + 1>::run(c, Lr, x);
        }
    }
};

template <typename T, int KMAX, bool UPPER, bool TA, int J>
struct DppSolve
{
    static __device__ __forceinline__ void run(T (&c)[KMAX], const DppTri<T, KMAX>& Lr,
                                               const T (&Dr)[(KMAX + 15) / 16])
    {
        if constexpr(J < KMAX)
        {
            // lane n holds the reciprocal of the stored diagonal entry n + 16 d,
            // i.e. of real index n + 16 d; canonical J is real index JR
            constexpr int JR = (UPPER != TA) ? KMAX - 1 - J : J;
            const T       x  = c[J] * row_bcast_ordered<JR % 16>(Dr[JR / 16]);
            c[J]             = x;
            DppColumn<T, KMAX, UPPER, TA, J, J + 1>::run(c, Lr, x);
            DppSolve<T, KMAX, UPPER, TA, J + 1>::run(c, Lr, Dr);
        }
    }
};

template <typename T, int KMAX, bool UPPER, bool TA, bool COLS, bool FULL>
__global__ void __launch_bounds__(WAVE) small_trsm_dpp_kernel(const SmallArgs<T> a)
{
    static_assert(KMAX <= 32 && WAVE % KMAX == 0, "register-resident triangles up to 32");
    constexpr bool REV = UPPER != TA;
    constexpr int  TRI = KMAX * (KMAX + 1) / 2;
    constexpr int  NR  = (TRI + 15) / 16;
    constexpr int  ND  = (KMAX + 15) / 16;
    constexpr int  VP  = DPP_VP;
    constexpr int  G   = DPP_G;
    // side L: one load step covers SPS whole columns; STEPS steps per slot
    constexpr int SPS   = WAVE / KMAX;
    constexpr int STEPS = VP / SPS;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* const RB = reinterpret_cast<T*>(smem_raw); // side-L staging, 64 x bs_stride
    __shared__ const T* slotA[G];
    __shared__ T*       slotB[G];
    __shared__ int      slotV0[G];

    const unsigned lane = threadIdx.x;
    const unsigned n    = lane & 15;
    const unsigned g    = lane >> 4;
    const int      K    = FULL ? KMAX : a.K;
    const bool     unit = a.flags & SF_UNIT;
    const int      lda  = a.lda;
    const int      ldb  = a.ldb;
    const int      bs   = a.bs_stride;
    const int      nvec = a.nvec;

    const T alpha = a.alpha_dev ? *a.alpha_dev : a.alpha_host;

    if(lane < G)
    {
        const int64_t u = int64_t(blockIdx.x) * G + lane;
        if(u < a.units)
        {
            const int64_t inst = a.chunks == 1 ? u : u / a.chunks;
            slotA[lane]        = a.A.get(inst);
            slotB[lane]        = a.B.get(inst);
            slotV0[lane]       = int(u - inst * a.chunks) * VP;
        }
        else
        {
            slotA[lane]  = nullptr;
            slotB[lane]  = nullptr;
            slotV0[lane] = 0;
        }
    }
    __syncthreads();

    // this lane's slot (row g) and vector
    const T*   myA   = slotA[g];
    T* const   myB   = slotB[g];
    const int  vec   = slotV0[g] + int(n);
    const bool valid = FULL || (myB != nullptr && vec < nvec);

    if(alpha == T(0))
    {
        if(valid)
            for(int i = 0; i < K; ++i)
                gstore(myB + (COLS ? (i + int64_t(vec) * ldb) : (vec + int64_t(i) * ldb)), T(0));
        return;
    }

    // Out-of-range lanes load from a valid address (slot 0 always exists) and
    // discard the value, so no load sits behind a branch.
    const T* const safeA = uniform_ptr(slotA[0]);
    T* const       safeB = uniform_ptr(slotB[0]);

    // ---- 1a. right-hand sides ----
    T c[KMAX];
    T bv[COLS ? KMAX : 1];
    // side L: lane = (vq, rl) loads row rl of column vq + SPS * s of its step
    const unsigned rl      = lane % KMAX;
    const unsigned vq      = lane / KMAX;
    const int64_t  laneoff = int64_t(rl) + int64_t(vq) * ldb;
    if constexpr(COLS)
    {
#pragma unroll
        for(int s = 0; s < KMAX; ++s)
        {
            const int slot = s / STEPS;
            const int k    = s % STEPS;
            T* const  sB   = uniform_ptr(slotB[slot]);
            const int sv   = uniform_int(slotV0[slot]) + SPS * k;
            const T*  p    = sB + (int64_t(sv) * ldb + laneoff);
            if constexpr(FULL)
                bv[s] = gload(p);
            else
            {
                const bool ok = sB != nullptr && int(rl) < K && sv + int(vq) < nvec;
                const T    x  = gload(ok ? p : safeB);
                bv[s]         = ok ? x : T(0);
            }
        }
    }
    else
    {
        const T* rowB = valid ? myB + vec : safeB;
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
        {
            const int r = REV ? KMAX - 1 - i : i;
            if constexpr(FULL)
                c[i] = gload(rowB + int64_t(r) * ldb);
            else
            {
                const bool ok = valid && r < K;
                const T    x  = gload(ok ? rowB + int64_t(r) * ldb : safeB);
                c[i]          = ok ? x : T(0);
            }
        }
    }

    // ---- 1b. this lane's share of the triangle (negated) and of the diagonal ----
    DppTri<T, KMAX> Lr;
    const T* const  rowA   = FULL || myA ? myA : safeA;
    const bool      slotok = FULL || myA != nullptr;
#pragma unroll
    for(int r = 0; r < NR; ++r)
    {
        const unsigned t = n + 16 * r;
        int            p, q;
        pair_decode<KMAX, UPPER>(int(t < TRI ? t : 0), p, q);
        bool ok = p != q;
        if(TRI % 16 != 0 && r == NR - 1)
            ok = ok && t < TRI;
        if constexpr(!FULL)
            ok = ok && slotok && p < K && q < K;
        const T x = gload(ok ? rowA + (p + int64_t(q) * lda) : rowA);
        Lr[r]     = ok ? -x : T(0);
    }
    T Dr[ND];
#pragma unroll
    for(int d = 0; d < ND; ++d)
    {
        const int  j  = n + 16 * d;
        const bool ok = slotok && j < K && !unit;
        const T    x  = gload(rowA + (ok ? j + int64_t(j) * lda : 0));
        Dr[d]         = ok ? T(1) / x : T(1);
    }

    // ---- 1c. side L: stage the columns so lane v can read column v ----
    if constexpr(COLS)
    {
        const unsigned rr = REV ? KMAX - 1 - rl : rl;
#pragma unroll
        for(int s = 0; s < KMAX; ++s)
            RB[(vq + SPS * s) * bs + rr] = bv[s];
        __syncthreads();
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            c[i] = RB[lane * bs + i];
    }
    if(a.flags & SF_SCALE)
    {
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            c[i] *= alpha;
    }

    // ---- 2. forward substitution, one DPP multiply-add per update ----
    // Every operand the DPP instructions read is final here (an empty asm
    // that redefines it), so the compiler cannot sink the last VALU
    // instruction producing one of them (a negation, a select, the alpha
    // scaling) in front of its first DPP reader, where the required wait
    // states would be missing.  The first update of the solve pads five.
#pragma unroll
    for(int r = 0; r < NR; ++r)
        asm volatile("" : "+v"(Lr[r]));
#pragma unroll
    for(int d = 0; d < ND; ++d)
        asm volatile("" : "+v"(Dr[d]));
#pragma unroll
    for(int i = 0; i < KMAX; ++i)
        asm volatile("" : "+v"(c[i]));
    DppSolve<T, KMAX, UPPER, TA, 0>::run(c, Lr, Dr);

    // ---- 3. store ----
    if constexpr(COLS)
    {
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            RB[lane * bs + i] = c[i];
        __syncthreads();
        const unsigned rr = REV ? KMAX - 1 - rl : rl;
#pragma unroll
        for(int s = 0; s < KMAX; ++s)
        {
            const int slot = s / STEPS;
            const int k    = s % STEPS;
            T* const  sB   = uniform_ptr(slotB[slot]);
            const int sv   = uniform_int(slotV0[slot]) + SPS * k;
            const T   y    = RB[(vq + SPS * s) * bs + rr];
            T* const  p    = sB + (int64_t(sv) * ldb + laneoff);
            if constexpr(FULL)
                gstore(p, y);
            else if(sB != nullptr && int(rl) < K && sv + int(vq) < nvec)
                gstore(p, y);
        }
    }
    else if(valid)
    {
        T* const rowB = myB + vec;
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
        {
            const int r = REV ? KMAX - 1 - i : i;
            if(FULL || r < K)
                gstore(rowB + int64_t(r) * ldb, c[i]);
        }
    }
}

template <typename T, int KMAX, bool UPPER, bool TA, bool COLS>
static void launch_dpp(rocblas_handle h, const SmallArgs<T>& a, dim3 grid, size_t lds)
{
    if(a.flags & SF_FULL)
        hipLaunchKernelGGL((small_trsm_dpp_kernel<T, KMAX, UPPER, TA, COLS, true>), grid,
                           dim3(WAVE), lds, h->stream, a);
    else
        hipLaunchKernelGGL((small_trsm_dpp_kernel<T, KMAX, UPPER, TA, COLS, false>), grid,
                           dim3(WAVE), lds, h->stream, a);
}

template <typename T, int KMAX, bool UPPER, bool TA>
static rocblas_status launch_small_side(rocblas_handle h, const SmallArgs<T>& a, dim3 grid,
                                        size_t lds)
{
    const bool cols = a.flags & SF_COLS;
    if constexpr(KMAX <= 32)
    {
        if(a.flags & SF_DPP)
        {
            if(cols)
                launch_dpp<T, KMAX, UPPER, TA, true>(h, a, grid, lds);
            else
                launch_dpp<T, KMAX, UPPER, TA, false>(h, a, grid, lds);
            return launch_status();
        }
    }
    if(cols)
        hipLaunchKernelGGL((small_trsm_lds_kernel<T, KMAX, UPPER, TA, true>), grid, dim3(WAVE),
                           lds, h->stream, a);
    else
        hipLaunchKernelGGL((small_trsm_lds_kernel<T, KMAX, UPPER, TA, false>), grid, dim3(WAVE),
                           lds, h->stream, a);
    return launch_status();
}

template <typename T, int KMAX>
rocblas_status launch_small_kmax(rocblas_handle h, const SmallArgs<T>& a, dim3 grid, size_t lds)
{
    const bool upper = a.flags & SF_UPPER;
    const bool ta    = a.flags & SF_TA;
    if(!upper && !ta)
        return launch_small_side<T, KMAX, false, false>(h, a, grid, lds);
    if(!upper && ta)
        return launch_small_side<T, KMAX, false, true>(h, a, grid, lds);
    if(upper && !ta)
        return launch_small_side<T, KMAX, true, false>(h, a, grid, lds);
    return launch_small_side<T, KMAX, true, true>(h, a, grid, lds);
}

} // namespace rbamd

#define RBAMD_INSTANTIATE_SMALL(KMAX)                                                     \
    namespace rbamd                                                                       \
    {                                                                                     \
    template rocblas_status launch_small_kmax<float, KMAX>(rocblas_handle,                \
                                                           const SmallArgs<float>&, dim3, \
                                                           size_t);                       \
    template rocblas_status launch_small_kmax<double, KMAX>(rocblas_handle,               \
                                                            const SmallArgs<double>&,     \
                                                            dim3, size_t);                \
    }
