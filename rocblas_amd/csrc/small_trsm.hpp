// small_trsm.hpp -- the k <= 64 triangular-solve kernel (instantiated per KMAX
// in small_k*.hip).
//
// Every side/uplo/trans case is reduced to forward substitution L y = c on the
// canonical lower triangle L = P M P (M = op(A) for side L, op(A)^T for side
// R; P reverses indices when M is upper triangular), one vector per lane.
// A wavefront serves G "slots" (instances) of VP vectors each (G * VP = 64).
//
// Data movement per wavefront, designed for one HBM round trip:
//   1. every global load of the wavefront's work is issued before the first
//      wait: the slots' referenced triangles (dense enumeration, below) and the
//      right-hand sides (side L: coalesced down columns; side R: one row per
//      lane, coalesced across lanes);
//   2. side L: the right-hand sides go through the LDS region so that lane v
//      can read column v, then the same region receives the triangles in the
//      canonical packed layout; reciprocal diagonals are formed once per entry;
//   3. fully unrolled substitution with broadcast LDS reads of L;
//   4. side L: solutions staged through the region again for coalesced stores.
//
// Dense triangle enumeration: within a KMAX x KMAX triangle, column c and
// column KMAX-1-c together hold exactly KMAX+1 entries, so entry t belongs to
// pair t / (KMAX+1) -- constant divisions, no search, consecutive t walk down
// a column (coalesced).
#pragma once

#include "device_common.hpp"

namespace rbamd
{

template <typename T, int KMAX, bool UPPER, bool TA>
__global__ void __launch_bounds__(WAVE) small_trsm_kernel(const SmallArgs<T> a)
{
    constexpr bool REV  = UPPER != TA;
    constexpr int  TRI  = KMAX * (KMAX + 1) / 2;
    constexpr int  GMAX = small_gmax(KMAX);
    constexpr int  NA   = (GMAX * TRI + WAVE - 1) / WAVE;

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
    const bool cols  = a.flags & SF_COLS;
    const bool scale = a.flags & SF_SCALE;
    const int  lda   = a.lda;
    const int  ldb   = a.ldb;
    const int  bs    = a.bs_stride;
    const int  S     = a.slot_stride;

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
                myB[cols ? (i + int64_t(vec) * ldb) : (vec + int64_t(i) * ldb)] = T(0);
        return;
    }

    // ---- 1a. issue the triangle loads ----
    const int nA = (G * TRI + WAVE - 1) / WAVE;
    T         av[NA];
    int       ad[NA];
#pragma unroll
    for(int s = 0; s < NA; ++s)
    {
        ad[s] = -1;
        av[s] = T(0);
        if(s < nA)
        {
            const int e  = lane + WAVE * s;
            const int g2 = e / TRI;
            const int t  = e % TRI;
            const int c  = t / (KMAX + 1);
            const int o  = t % (KMAX + 1);
            int       p, q;
            if(!UPPER)
            {
                // lower: column c has rows c..KMAX-1, column KMAX-1-c rows KMAX-1-c..
                if(o < KMAX - c)
                {
                    q = c;
                    p = c + o;
                }
                else
                {
                    q = KMAX - 1 - c;
                    p = q + (o - (KMAX - c));
                }
            }
            else
            {
                // upper: column c has rows 0..c, column KMAX-1-c rows 0..KMAX-1-c
                if(o <= c)
                {
                    q = c;
                    p = o;
                }
                else
                {
                    q = KMAX - 1 - c;
                    p = o - (c + 1);
                }
            }
            const T* Ap = g2 < G ? slotA[g2] : nullptr;
            if(Ap && p < K && q < K)
            {
                const int mi = TA ? q : p;
                const int mj = TA ? p : q;
                const int i  = REV ? K - 1 - mi : mi;
                const int j  = REV ? K - 1 - mj : mj;
                ad[s]        = g2 * S + colstart<KMAX>(j) + (i - j);
                if(!(unit && p == q))
                    av[s] = Ap[p + int64_t(q) * lda];
                else
                    av[s] = T(1);
            }
        }
    }

    // ---- 1b. issue the right-hand-side loads ----
    T c[KMAX];
    if(cols)
    {
        // element e = lane + 64 s: row e % KMAX of staged vector e / KMAX
        T bv[KMAX];
#pragma unroll
        for(int s = 0; s < KMAX; ++s)
        {
            const int e  = lane + WAVE * s;
            const int r  = e % KMAX;
            const int v2 = e / KMAX;
            T* const  Bp = slotB[v2 >> vpl];
            const int vc = slotV0[v2 >> vpl] + (v2 & (VP - 1));
            bv[s]        = (Bp && r < K && vc < a.nvec) ? Bp[r + int64_t(vc) * ldb] : T(0);
        }
#pragma unroll
        for(int s = 0; s < KMAX; ++s)
        {
            const int e  = lane + WAVE * s;
            const int r  = e % KMAX;
            const int v2 = e / KMAX;
            if(r < K)
                R[v2 * bs + (REV ? K - 1 - r : r)] = bv[s];
        }
        __syncthreads();
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            c[i] = R[lane * bs + i];
        __syncthreads(); // the region now receives the triangles
    }
    else
    {
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            c[i] = (valid && i < K) ? myB[vec + int64_t(REV ? K - 1 - i : i) * ldb] : T(0);
    }

    // ---- 2. triangles -> LDS (canonical packed lower), reciprocal diagonal ----
#pragma unroll
    for(int s = 0; s < NA; ++s)
        if(ad[s] >= 0)
            R[ad[s]] = av[s];
    __syncthreads();
    if(!unit)
    {
        for(int idx = lane; idx < G * KMAX; idx += WAVE)
        {
            const int g2 = idx / KMAX;
            const int j  = idx % KMAX;
            if(j < K)
            {
                T& d = R[g2 * S + colstart<KMAX>(j)];
                d    = T(1) / d;
            }
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
    const T* L = R + g * S;
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
        __syncthreads(); // every lane is done reading the triangles
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            R[lane * bs + i] = c[i];
        __syncthreads();
#pragma unroll
        for(int s = 0; s < KMAX; ++s)
        {
            const int e  = lane + WAVE * s;
            const int r  = e % KMAX;
            const int v2 = e / KMAX;
            T* const  Bp = slotB[v2 >> vpl];
            const int vc = slotV0[v2 >> vpl] + (v2 & (VP - 1));
            if(Bp && r < K && vc < a.nvec)
                Bp[r + int64_t(vc) * ldb] = R[v2 * bs + (REV ? K - 1 - r : r)];
        }
    }
    else if(valid)
    {
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            if(i < K)
                myB[vec + int64_t(REV ? K - 1 - i : i) * ldb] = c[i];
    }
}

template <typename T, int KMAX>
rocblas_status launch_small_kmax(rocblas_handle h, const SmallArgs<T>& a, dim3 grid, size_t lds)
{
    const bool upper = a.flags & SF_UPPER;
    const bool ta    = a.flags & SF_TA;
    if(!upper && !ta)
        hipLaunchKernelGGL((small_trsm_kernel<T, KMAX, false, false>), grid, dim3(WAVE), lds,
                           h->stream, a);
    else if(!upper && ta)
        hipLaunchKernelGGL((small_trsm_kernel<T, KMAX, false, true>), grid, dim3(WAVE), lds,
                           h->stream, a);
    else if(upper && !ta)
        hipLaunchKernelGGL((small_trsm_kernel<T, KMAX, true, false>), grid, dim3(WAVE), lds,
                           h->stream, a);
    else
        hipLaunchKernelGGL((small_trsm_kernel<T, KMAX, true, true>), grid, dim3(WAVE), lds,
                           h->stream, a);
    return launch_status();
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
