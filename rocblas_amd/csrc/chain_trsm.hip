// chain_trsm.hip -- single-launch solve of one instance whose triangle is long
// and whose right-hand side is narrow: one workgroup per (64-row block I,
// 16-vector panel p), chained through a completion counter per panel.
//
// Same canonical form and per-block arithmetic as the panel solve
// (panel_trsm.hpp): forward substitution L y = c on the canonical lower
// triangle, left-looking,
//
//   c_I = scale * b_I - sum_{J<I} L_IJ y_J    (FP64/FP32 MFMA)
//   y_I = L_II^{-1} c_I                       (substitution, one DPP
//                                              multiply-add per update)
//
// but block row I of panel p is its own workgroup: it streams only its own
// tiles L_IJ (they do not depend on the solution, so they are loaded while
// the blocks above are still being solved), waits for y_J through the
// panel's counter, reads y_J back from B, and publishes y_I by raising the
// counter to I + 1.  The dependency chain is one block solve per 64 rows
// instead of one workgroup streaming the whole triangle, and the tile traffic
// is spread over R * P workgroups.
//
// Ordering: workgroup (p, I) has linear index I * P + p and only waits for
// smaller linear indices, which are dispatched first, so the chain cannot
// deadlock.  Every wait is bounded (about a second of polls): a wait that runs
// out raises the handle's error word (host-visible memory, read by the next
// call on the handle and by rocblas_amd_get_async_status) and proceeds, so
// every wave reaches its exit.
//
// Counter word of panel p: the low 32 bits count the block rows that have
// published, the high 32 bits the workgroups that have finished.  Every
// workgroup ends with one atomic add of (1 << 32) + 1; the one that sees the
// other R - 1 finishes in the value it replaced is the last user of the word
// and stores zero.  So the counters are zero after every launch -- after a
// wait that gave up, too -- and a graph replay starts from zero.
//
// Visibility across the XCDs' separate L2 caches: each wave ends its stores
// of y_I with an agent-scope release fence (L2 write-back, wait for the
// stores) before the workgroup barrier that precedes the counter store; a
// waiter polls with agent-scope loads and issues an agent-scope acquire fence
// (cache invalidate) before reading y_J.
#include "panel_trsm.hpp"

namespace rbamd
{
namespace
{
using namespace panel;

// Polls of one wait before it gives up (about a second with s_sleep 2).
constexpr int kChainSpinLimit = 1 << 20;
constexpr unsigned long long kRowsMask = 0xffffffffull;
constexpr unsigned long long kFinishOne = 1ull << 32;

// Wait controls: the poll limit and an offset added to every wait's target.
// Normal launches use kChainSpinLimit and 0; a test hook of the handle sets a
// small limit and an unreachable offset so that every wait gives up.
struct ChainWait
{
    int      spin_limit;
    unsigned need_bias;
};

template <typename T>
__global__ void __launch_bounds__(PNT)
    chain_trsm_kernel(const SmallArgs<T> a, unsigned long long* cnt, unsigned int* err, int P,
                      int R, ChainWait cw)
{
    extern __shared__ __align__(16) unsigned char chain_smem[];
    T* const Ls = reinterpret_cast<T*>(chain_smem); // PNB x PLDA staged tile L_IJ
    T* const Ln = Ls + PNB * PLDA;                  // PNB x PLDA -L_II, strictly lower
    T* const Cs = Ln + PNB * PLDA;                  // PNB x PNC accumulated update
    T* const dg = Cs + PNB * PNC;                   // PNB diagonal of L_II
    T* const Yj = dg + PNB;                         // PNB x PLDY solved block y_J
    __shared__ unsigned long long seen;             // rows seen published (low word)

    const int  I    = blockIdx.x / P;
    const int  p    = blockIdx.x % P;
    const int  tid  = threadIdx.x;
    const int  lane = tid & (WAVE - 1);
    const int  w    = tid / WAVE;
    const int  K    = a.K;
    const int  lda  = a.lda;
    const int  ldb  = a.ldb;
    const bool tA   = a.flags & SF_TA;
    const bool rev  = a.flags & SF_REV;
    const bool unit = a.flags & SF_UNIT;
    const bool cols = a.flags & SF_COLS;
    const T*   Ap   = a.A.get(0);
    T* const   Bp   = a.B.get(0);

    const T alpha = a.alpha_dev ? *a.alpha_dev : a.alpha_host;
    const T scale = (a.flags & SF_SCALE) ? alpha : T(1);

    // substitution layout: vector vl = 4w + lane/16, rows g + 16r of a block
    const int  g   = lane & 15;
    const int  vl  = 4 * w + (lane >> 4);
    const int  vec = p * PNC + vl;
    const bool vok = vec < a.nvec;

    auto boff = [&](int i, int v) -> int64_t {
        const int r = rev ? K - 1 - i : i;
        return cols ? int64_t(r) + int64_t(v) * ldb : int64_t(v) + int64_t(r) * ldb;
    };

    if(alpha == T(0))
    {
        // B := 0 for this block and panel; nobody waits (alpha is uniform)
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const int gi = I * PNB + g + 16 * r;
            if(vok && gi < K)
                gstore(Bp + boff(gi, vec), T(0));
        }
        return;
    }

    // Tile (I, J) of canonical L, element s of this thread: canonical (i, j)
    // chosen so that consecutive threads read consecutive addresses.
    auto tile_ij = [&](int s, int& i, int& j) {
        const int e = tid + PNT * s;
        if(!tA)
        {
            i = e % PNB;
            j = e / PNB;
        }
        else
        {
            j = e % PNB;
            i = e / PNB;
        }
    };
    T    reg[PEPT];
    int  tJ        = 0;
    auto load_tile = [&](int J) {
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
            const int  gi = I * PNB + i, gj = tJ * PNB + j;
            const bool in = gi < K && gj <= gi && !(unit && gi == gj);
            Ls[i * PLDA + j] = in ? reg[s] : T(0);
        }
    };
    // the diagonal block: -L strictly below the diagonal, the diagonal apart
    auto store_diag = [&]() {
#pragma unroll
        for(int s = 0; s < PEPT; ++s)
        {
            int i, j;
            tile_ij(s, i, j);
            const int gi = I * PNB + i;
            Ln[i * PLDA + j] = (gi < K && j < i) ? -reg[s] : T(0);
            if(i == j)
                dg[i] = (gi < K && !unit) ? reg[s] : T(1);
        }
    };

    // counter[p] >= need: bounded poll by one thread, then every wave acquires
    if(tid == 0)
        seen = 0;
    auto wait_rows = [&](int need_rows) {
        const unsigned long long need = (unsigned long long)need_rows + cw.need_bias;
        __syncthreads();
        if(seen >= need)
            return;
        __syncthreads(); // every thread has read `seen` before it changes
        if(tid == 0)
        {
            unsigned long long v = 0;
            for(int it = 0;; ++it)
            {
                v = __hip_atomic_load(cnt + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                    & kRowsMask;
                if(v >= need)
                    break;
                if(it >= cw.spin_limit)
                {
                    __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    v = kRowsMask; // give up: every later wait of this workgroup returns
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            seen = v;
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    };

    // right-hand sides of this block, consumed after the accumulation
    T bv[4];
#pragma unroll
    for(int r = 0; r < 4; ++r)
    {
        const int gi = I * PNB + g + 16 * r;
        bv[r]        = gload(Bp + ((vok && gi < K) ? boff(gi, vec) : 0));
    }

    // ---- the diagonal block first: it does not depend on the solution ----
    load_tile(I);
    store_diag();
    __syncthreads();
    T rd[4];
#pragma unroll
    for(int r = 0; r < 4; ++r)
        rd[r] = T(1) / dg[g + 16 * r];

    // ---- c_I accumulation: sum_{J<I} L_IJ y_J on MFMA ----
    typename Mfma<T>::acc_t acc[MCH];
#pragma unroll
    for(int a = 0; a < MCH; ++a)
        acc[a] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    const T* lrow = Ls + (16 * w + (lane & 15)) * PLDA + (lane >> 4);
    const T* ycol = Yj + (lane >> 4) * PLDY + (lane & 15);
    if(I > 0)
        load_tile(0);
    for(int J = 0; J < I; ++J)
    {
        wait_rows(J + 1);
        // y_J from B (written by workgroup (p, J)) into LDS
        T y[4];
#pragma unroll
        for(int r = 0; r < 4; ++r)
        {
            const int gi = J * PNB + g + 16 * r;
            y[r]         = vok ? gload(Bp + boff(gi, vec)) : T(0);
        }
        __syncthreads(); // previous readers of Ls and Yj are done
#pragma unroll
        for(int r = 0; r < 4; ++r)
            Yj[(g + 16 * r) * PLDY + vl] = y[r];
        store_tile();
        __syncthreads();
        if(J + 1 < I)
            load_tile(J + 1);
        mfma_tile<T>(acc, lrow, ycol);
    }
#pragma unroll
    for(int r = 0; r < 4; ++r)
        Cs[(16 * w + Mfma<T>::row(lane, r)) * PNC + (lane & 15)] = fold_acc<T>(acc, r);
    __syncthreads();

    // ---- y_I = L_II^{-1} c_I ----
    T c[4];
#pragma unroll
    for(int r = 0; r < 4; ++r)
    {
        const int  gi = I * PNB + g + 16 * r;
        const bool ok = vok && gi < K;
        c[r]          = ok ? scale * bv[r] - Cs[(g + 16 * r) * PNC + vl] : T(0);
    }
    // every operand of the DPP updates is final here (an empty asm that
    // redefines it), so no VALU instruction producing one of them is sunk in
    // front of its first DPP reader
#pragma unroll
    for(int r = 0; r < 4; ++r)
    {
        asm volatile("" : "+v"(c[r]));
        asm volatile("" : "+v"(rd[r]));
    }
    fused_solve<T>(c, rd, Ln, g, min(PNB, K - I * PNB));
#pragma unroll
    for(int r = 0; r < 4; ++r)
    {
        const int gi = I * PNB + g + 16 * r;
        if(vok && gi < K)
            gstore(Bp + boff(gi, vec), c[r] * rd[r]);
    }

    // ---- publish y_I and count this workgroup finished ----
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); // this wave's stores, written back
    __syncthreads();
    if(tid == 0)
    {
        const unsigned long long old = __hip_atomic_fetch_add(
            cnt + p, kFinishOne + 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        // the last of the panel's R workgroups: nobody reads the word any more
        if((old >> 32) == (unsigned long long)(R - 1))
            __hip_atomic_store(cnt + p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
} // namespace

size_t chain_lds_bytes(size_t elt)
{
    return elt * (2 * size_t(PNB) * PLDA + size_t(PNB) * PNC + PNB + size_t(PNB) * PLDY);
}

template <typename T>
bool chain_fits(rocblas_handle h, int K, int nvec, int batch)
{
    // Each block row streams its own tiles, so the tile traffic grows with
    // P * R^2 while the dependency chain grows with R: the chain pays off for
    // small grids and for triangles long relative to the panel count (measured
    // at 8x CUs: 4096x256 567 vs 937 us, 2048x512 336 vs 438 us recursive;
    // 1024x1024, P = 4R, 219 vs 183 us).
    const int64_t R = (K + PNB - 1) / PNB, P = (nvec + PNC - 1) / PNC;
    const int64_t cus = h->num_cus;
    return h->chain_flags != nullptr && batch == 1 && K > 64 && R >= 2 && P <= kChainMaxPanels
           && (R * P <= 2 * cus || (P <= R && R * P <= 8 * cus))
           && chain_lds_bytes(sizeof(T)) <= h->max_lds;
}

template <typename T>
rocblas_status launch_chain(rocblas_handle h, const SmallArgs<T>& a)
{
    if(!chain_fits<T>(h, a.K, a.nvec, 1))
        return rocblas_status_internal_error;
    const int R = (a.K + PNB - 1) / PNB, P = (a.nvec + PNC - 1) / PNC;
    ChainWait cw{kChainSpinLimit, 0u};
    if(h->chain_debug_timeout)
        cw = ChainWait{64, 1u << 30};
    hipLaunchKernelGGL(chain_trsm_kernel<T>, dim3(unsigned(R * P)), dim3(PNT),
                       chain_lds_bytes(sizeof(T)), h->stream, a,
                       static_cast<unsigned long long*>(h->chain_flags), h->async_error, P, R,
                       cw);
    return launch_status();
}

template bool           chain_fits<float>(rocblas_handle, int, int, int);
template bool           chain_fits<double>(rocblas_handle, int, int, int);
template rocblas_status launch_chain<float>(rocblas_handle, const SmallArgs<float>&);
template rocblas_status launch_chain<double>(rocblas_handle, const SmallArgs<double>&);

} // namespace rbamd
