// gemm.hip -- MFMA GEMM for the trailing updates of the blocked solve.
//
//   C := alpha * op(A) * op(B) + beta * C   (column-major, batched)
//
// FP64 uses v_mfma_f64_16x16x4_f64, FP32 v_mfma_f32_16x16x4_f32; operands
// stored in a wider type than the MFMA's are rounded on their way into LDS
// (the mixed-precision solve's FP32 updates of FP64 data).  Tiles are
// staged through LDS (double buffered), with the next K step's global loads in
// flight during the current step's MFMAs.  Workgroup ids are remapped so that
// each XCD's L2 sees contiguous runs of tiles, grouped by row panel.
// A triangular operand (trtri's doubling products, the mixed form's
// residuals) limits each tile's K range to the triangle and is masked in its
// diagonal block; such products with few tiles split their tiles' K ranges or
// balance their K steps over the workgroups.
#include "device_common.hpp"

#include <cstdlib>
#include <type_traits>

namespace rbamd
{
// ---------------------------------------------------------------------------
// MFMA GEMM:  C := alpha * op(A) * op(B) + beta * C   (column-major, batched)
//   op(A) is M x K, op(B) is K x N.  Workgroup tile BM x BN, K step BK,
//   WGM x WGN wavefronts each owning a (BM/WGM) x (BN/WGN) block of 16x16
//   MFMA tiles.  Global -> register prefetch of the next K step overlaps the
//   MFMAs of the current one; LDS is double buffered (one barrier per step).
// ---------------------------------------------------------------------------

template <typename TA, typename TB, typename TC, typename T, bool TRI, int OPA, int OPB, int BM,
          int BN, int BK, int WGM, int WGN>
__global__ void __launch_bounds__(WAVE * WGM * WGN)
    gemm_kernel(const GemmArgsT<TA, TB, TC> a)
{
    using acc_t         = typename Mfma<T>::acc_t;
    constexpr int NT    = WAVE * WGM * WGN;
    constexpr int WTM   = BM / WGM;
    constexpr int WTN   = BN / WGN;
    constexpr int TM    = WTM / 16;
    constexpr int TN    = WTN / 16;
    constexpr int LDA_S = BM + 2;
    constexpr int LDB_S = BN + 2;
    constexpr int EA    = BM * BK / NT;
    constexpr int EB    = BK * BN / NT;
    static_assert(EA * NT == BM * BK && EB * NT == BK * BN, "tile/threads mismatch");
    static_assert(TM >= 1 && TN >= 1 && BK % 4 == 0, "bad tile");
    static_assert(NT % BM == 0 && NT % BN == 0 && NT % BK == 0, "load geometry");

    __shared__ T As[2][BK * LDA_S];
    __shared__ T Bs[2][BK * LDB_S];

    if(a.skip_if_zero && *a.skip_if_zero == TC(0))
        return;

    const int tid  = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w    = tid / WAVE;
    const int wm   = w % WGM;
    const int wn   = w / WGM;

    // tile index: spread consecutive tiles over the 8 XCDs' L2s in contiguous
    // runs, then group row-panels so neighbouring workgroups share A panels.
    // With a triangular operand the work per row panel varies, and contiguous
    // runs would hand one XCD all the long panels: keep dispatch order there
    // (round-robin over XCDs, each XCD one row panel of every group).
    const int total = a.tiles_m * a.tiles_n;
    int       bid   = blockIdx.x;
    if constexpr(!TRI)
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

    // K range holding non-zero products (triangular operands)
    int kbeg = 0, kend = a.K;
    if constexpr(TRI)
    {
        if(a.tri & GT_A_LOWER)
            kend = min(kend, m0 + BM);
        if(a.tri & GT_A_UPPER)
            kbeg = max(kbeg, m0);
        if(a.tri & GT_B_LOWER)
            kbeg = max(kbeg, n0);
        if(a.tri & GT_B_UPPER)
            kend = min(kend, n0 + BN);
        kbeg = kbeg / BK * BK;
    }
    const int ktiles = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;

    // Per-thread load geometry, fixed over K: element s of the A tile is
    // (row ia(s), k offset ka(s)), of the B tile (k offset kb(s), column jb(s)),
    // chosen so that consecutive threads read consecutive addresses.
    auto ia = [&](int s) { return OPA == 0 ? tid % BM : tid / BK + s * (NT / BK); };
    auto ka = [&](int s) { return OPA == 0 ? tid / BM + s * (NT / BM) : tid % BK; };
    auto kb = [&](int s) { return OPB == 0 ? tid % BK : tid / BN + s * (NT / BN); };
    auto jb = [&](int s) { return OPB == 0 ? tid / BK + s * (NT / BK) : tid % BN; };
    const bool unitA = TRI && (a.tri & GT_UNIT) && (a.tri & (GT_A_LOWER | GT_A_UPPER));
    const bool unitB = TRI && (a.tri & GT_UNIT) && (a.tri & (GT_B_LOWER | GT_B_UPPER));

    for(int b = blockIdx.y; b < a.batch; b += gridDim.y)
    {
        const TA* Ap = a.A.get(b);
        const TB* Bp = a.B.get(b);
        TC*       Cp = a.C.get(b);

        acc_t acc[TM][TN];
#pragma unroll
        for(int i = 0; i < TM; ++i)
#pragma unroll
            for(int j = 0; j < TN; ++j)
                acc[i][j] = acc_t{0, 0, 0, 0};

        // raw operand values of the tile in flight; bounds and triangular masks
        // are applied when the tile is written to LDS, one loop iteration
        // later, so no load is predicated on them and all of a tile's loads
        // are in flight together
        TA  ra[EA];
        TB  rb[EB];
        int k_ld = 0;

        // Branch-free tile loads: out-of-range indices are clamped to a valid element.
        auto load_tiles = [&](int k0) {
            k_ld = k0;
#pragma unroll
            for(int s = 0; s < EA; ++s)
            {
                const int gic = min(m0 + ia(s), a.M - 1);
                const int gkc = min(k0 + ka(s), kend - 1);
                ra[s] = gload(Ap + (OPA == 0 ? gic + int64_t(gkc) * a.lda
                                             : gkc + int64_t(gic) * a.lda));
            }
#pragma unroll
            for(int s = 0; s < EB; ++s)
            {
                const int gjc = min(n0 + jb(s), a.N - 1);
                const int gkc = min(k0 + kb(s), kend - 1);
                rb[s] = gload(Bp + (OPB == 0 ? gkc + int64_t(gjc) * a.ldb
                                             : gjc + int64_t(gkc) * a.ldb));
            }
        };
        auto store_tiles = [&](int buf) {
#pragma unroll
            for(int s = 0; s < EA; ++s)
            {
                const int gi = m0 + ia(s), gk = k_ld + ka(s);
                TA        v  = ra[s];
                if constexpr(TRI)
                {
                    const bool zero = ((a.tri & GT_A_LOWER) && gk > gi)
                                      || ((a.tri & GT_A_UPPER) && gk < gi);
                    v = zero ? TA(0) : (unitA && gk == gi ? TA(1) : v);
                }
                As[buf][ka(s) * LDA_S + ia(s)] = (gi < a.M && gk < kend) ? T(v) : T(0);
            }
#pragma unroll
            for(int s = 0; s < EB; ++s)
            {
                const int gj = n0 + jb(s), gk = k_ld + kb(s);
                TB        v  = rb[s];
                if constexpr(TRI)
                {
                    const bool zero = ((a.tri & GT_B_LOWER) && gk < gj)
                                      || ((a.tri & GT_B_UPPER) && gk > gj);
                    v = zero ? TB(0) : (unitB && gk == gj ? TB(1) : v);
                }
                Bs[buf][kb(s) * LDB_S + jb(s)] = (gj < a.N && gk < kend) ? T(v) : T(0);
            }
        };

        if(ktiles > 0)
            load_tiles(kbeg);
        for(int kt = 0; kt < ktiles; ++kt)
        {
            const int buf = kt & 1;
            store_tiles(buf);
            __syncthreads();
            // the load past the last tile reads clamped addresses and is discarded
            load_tiles(kbeg + (kt + 1) * BK);
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
                // operands swapped: the MFMA computes the transposed tile, so a
                // lane's results run down a column of C (coalesced stores)
#pragma unroll
                for(int i = 0; i < TM; ++i)
#pragma unroll
                    for(int j = 0; j < TN; ++j)
                        acc[i][j] = Mfma<T>::run(bf[j], af[i], acc[i][j]);
            }
        }

        // epilogue: lane holds row lane%16 and columns Mfma<T>::row(lane, r) of each tile
#pragma unroll
        for(int i = 0; i < TM; ++i)
        {
            const int row = m0 + wm * WTM + i * 16 + (lane & 15);
#pragma unroll
            for(int j = 0; j < TN; ++j)
#pragma unroll
                for(int r = 0; r < 4; ++r)
                {
                    const int col = n0 + wn * WTN + j * 16 + Mfma<T>::row(lane, r);
                    if(row < a.M && col < a.N)
                    {
                        TC& cref = Cp[row + int64_t(col) * a.ldc];
                        TC  v    = a.alpha * TC(acc[i][j][r]);
                        if(a.beta != TC(0))
                            v = fma(a.beta, cref, v);
                        cref = v;
                    }
                }
        }
        __syncthreads(); // LDS reuse across batch iterations
    }
}

// ---------------------------------------------------------------------------
// FP64 GEMM with the operand tiles loaded straight into LDS
// (global_load_lds_dwordx4): C := alpha * op(A) * op(B) + beta * C, dense,
// 128 x 128 tiles, K steps of BK (8 or 16) and NS LDS stages with NS - 1
// tiles in flight.  No staging registers, so the deeper pipeline costs no
// occupancy.  Each wavefront's load instruction writes 64 x 16 contiguous LDS
// bytes, lane l at byte 16 l.  An operand tile (x = row of op(A) or column of
// op(B), 128 of them, by BK k) takes one of two LDS layouts, by which index
// is contiguous in memory:
//   x contiguous (A, or B transposed): BK rows (k) of 128 doubles, pitch 144
//            doubles, so the fragment reads of two k rows fall in disjoint
//            bank halves; one instruction is one row, lane l holding
//            x = 2l, 2l + 1.
//   k contiguous (B, or A transposed): 128 rows (x) of BK doubles, 128 / BK
//            rows per instruction; lane l holds row l / (BK/2) and the k
//            pair (l % (BK/2)) ^ key(x), an XOR swizzle by x / (256 / (8 BK))
//            that puts the 16 rows a fragment read touches in distinct banks.
// Eligible when lda, ldb and every operand address are even / 16-byte
// aligned, K is a multiple of BK and the extent of an x-contiguous operand
// (M for A, N for transposed B) is even (launch_gemm checks); rows / columns
// beyond M / N are read clamped and only feed parts of C that are not stored.
// ---------------------------------------------------------------------------
constexpr int DMA_LDA = 144;

// One global_load_lds_dwordx4, written as inline assembly so the compiler's
// wait-count pass does not treat every later LDS read as depending on it (it
// cannot tell the stages apart and would wait for all loads in flight); the
// kernel's explicit vmcnt waits order it instead.  M0 holds the wavefront's
// LDS destination; one wait state separates its write from the load.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm" // m0 is reserved; the clobber is still honoured
__device__ __forceinline__ void dma16(const double* src, const double* lds)
{
    // wave-uniform (derived from the wavefront index); readfirstlane puts it
    // in the scalar register the "s" constraint needs
    const uint32_t dst = __builtin_amdgcn_readfirstlane(uint32_t(
        reinterpret_cast<uintptr_t>((const __attribute__((address_space(3))) double*)(lds))));
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
                 :
                 : "v"(src), "s"(dst)
                 : "memory", "m0");
}
#pragma clang diagnostic pop

// wait until at most N of this wavefront's vector-memory loads are in flight
// (gfx9 s_waitcnt encoding; expcnt and lgkmcnt left at their maxima)
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 16, "vmcnt");
    __builtin_amdgcn_s_waitcnt(0x0F70 | N);
}

// Workgroup barrier without __syncthreads' memory fence, whose vmcnt(0)
// would also wait for the loads of the tiles still in flight; the "memory"
// clobber keeps LDS accesses on their side of it.  The LDS reads before it
// are complete (their values feed MFMAs already issued) and wait_vmcnt has
// retired this wavefront's loads of the tile the next reads use.
__device__ __forceinline__ void barrier_no_fence()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// One operand's tiles in the DMA kernel: XC = x contiguous in memory.
template <int BK, bool XC>
struct DmaOperand
{
    static constexpr int ELS   = XC ? BK * DMA_LDA : 128 * BK;
    static constexpr int NSLOT = BK / 2;         // k pairs per row (k-contiguous layout)
    static constexpr int JPC   = 256 / (8 * BK); // rows per 256-byte bank cycle
    static constexpr int PER   = BK / 4;         // load instructions per wavefront per tile

    // element (k, x) of a stage
    static __device__ __forceinline__ int at(int k, int x)
    {
        if constexpr(XC)
            return k * DMA_LDA + x;
        else
            return x * BK + (((k >> 1) ^ ((x / JPC) % NSLOT)) << 1) + (k & 1);
    }

    // Per-lane load state, a few 32-bit values so it stays cheap in
    // registers: the x-contiguous layout keeps this lane's x pair (each
    // instruction adds a wave-uniform row), the k-contiguous one the row of
    // its first instruction and its k-pair slot.  Row xl of the tile (for
    // instruction g of wavefront w: (w PER + g) 128/BK + l / (BK/2)) has the
    // swizzle key (xl / JPC) % (BK/2) = (4 (w PER + g) + l / 16) % (BK/2).
    struct Lane
    {
        int x;
        int row0;
        int slot;
        int quad;
    };

    static __device__ __forceinline__ Lane lane_state(int w, int lane, int x0, int X)
    {
        Lane l{};
        if constexpr(XC)
            l.x = min(x0 + 2 * lane, X - 2);
        else
        {
            l.row0 = x0 + w * PER * (128 / BK) + lane / NSLOT;
            l.slot = lane % NSLOT;
            l.quad = lane / 16;
        }
        return l;
    }

    // source of load instruction g of wavefront w at k offset k0
    static __device__ __forceinline__ const double*
        src(const double* base, const Lane& l, int g, int w, int64_t k0, int ld, int X)
    {
        if constexpr(XC)
            return base + l.x + (k0 + (w * PER + g)) * ld;
        else
        {
            const int row = min(l.row0 + g * (128 / BK), X - 1);
            const int key = (4 * (w * PER + g) + l.quad) % NSLOT;
            return base + (k0 + 2 * (l.slot ^ key)) + int64_t(row) * ld;
        }
    }

    // LDS destination of load instruction g of wavefront w
    static __device__ __forceinline__ int dst(int g, int w)
    {
        return XC ? (w * PER + g) * DMA_LDA : (w * PER + g) * 128;
    }
};

// Wait of a split piece: thread 0 polls the tile's counter (agent scope)
// until it reaches `need`, at most spin_limit times; a wait that gives up
// raises the error word (the sum is then wrong, and the host reports it) and
// goes on.  Every wavefront then acquires what the earlier pieces wrote.
__device__ __forceinline__ void split_wait(unsigned* cnt, unsigned need, int spin_limit,
                                           unsigned* err)
{
    if(threadIdx.x == 0)
    {
        for(int it = 0;; ++it)
        {
            if(__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need)
                break;
            if(it >= spin_limit)
            {
                __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// A split piece has stored its part of C: every wavefront releases its
// stores, then one thread counts the piece finished; whichever of the tile's
// `parts` pieces finishes last (normally the last piece) clears the counter
// for the next launch, after every other piece has counted itself.
__device__ __forceinline__ void split_signal(unsigned* cnt, unsigned parts)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if(threadIdx.x == 0)
    {
        const unsigned old
            = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if(old == parts - 1)
            __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One tile job of the DMA kernels: the 128 x 128 tile at (m0, n0) of
// instance b, over the K range [kbeg, kend), through LDS stages As / Bs.
// TRI: 0 for dense operands, else the one GemmTri mask bit of the
// triangular operand (GT_A_LOWER ... GT_B_UPPER), fixed at compile time so
// that no mask test branches at run time; GT_UNIT is read from a.tri.
// [dlo, dhi) is the triangular operand's diagonal block, whose K steps are
// masked in LDS.  The result goes
//   to `part` when it is given: the raw sums, element v = (4 i + j) 4 + r of
//     thread t at part[v * 256 + t] (coalesced), for a later in-order sum;
//   else to C: alpha * sum + beta * C, or, for piece `piece` > 0 of `parts`
//     pieces of the tile, added to C after the earlier pieces, in piece
//     order through the tile's counter cnt.
template <int BK, int NS, int OPA, int OPB, int TRI>
__device__ __forceinline__ void dma_job(const GemmArgsT<double, double, double>& a,
                                        double (*As)[DmaOperand<BK, OPA == 0>::ELS],
                                        double (*Bs)[DmaOperand<BK, OPB == 1>::ELS], int b,
                                        int m0, int n0, int kbeg, int kend, int dlo, int dhi,
                                        int piece, int parts, unsigned* cnt, double* part)
{
    constexpr bool TRI_A  = TRI == GT_A_LOWER || TRI == GT_A_UPPER;
    constexpr bool TO_END = TRI == GT_A_LOWER || TRI == GT_B_UPPER; // range ends at the diagonal
    using OA              = DmaOperand<BK, OPA == 0>;
    using OB              = DmaOperand<BK, OPB == 1>;
    constexpr int PER     = BK / 4;  // load instructions per operand per wavefront
    constexpr int LOADS   = 2 * PER; // loads per wavefront per tile

    const int tid    = threadIdx.x;
    const int lane   = tid & (WAVE - 1);
    const int w      = tid / WAVE;
    const int wm     = w % 2;
    const int wn     = w / 2;
    const int ktiles = kend > kbeg ? (kend - kbeg) / BK : 0;

    const typename OA::Lane la = OA::lane_state(w, lane, m0, a.M);
    const typename OB::Lane lb = OB::lane_state(w, lane, n0, a.N);

    const double* Ap = a.A.get(b);
    const double* Bp = a.B.get(b);
    double*       Cp = a.C.get(b);

    // tiles past the last one re-read the last tile (in bounds) into a
    // stage nobody reads any more, so every step waits the same way
    auto issue = [&](int kt, int st) {
        const int64_t k0 = kbeg + int64_t(min(kt, ktiles - 1)) * BK;
#pragma unroll
        for(int g = 0; g < PER; ++g)
            dma16(OA::src(Ap, la, g, w, k0, a.lda, a.M), &As[st][OA::dst(g, w)]);
#pragma unroll
        for(int g = 0; g < PER; ++g)
            dma16(OB::src(Bp, lb, g, w, k0, a.ldb, a.N), &Bs[st][OB::dst(g, w)]);
    };

    d4_t acc[4][4];
#pragma unroll
    for(int i = 0; i < 4; ++i)
#pragma unroll
        for(int j = 0; j < 4; ++j)
            acc[i][j] = d4_t{0, 0, 0, 0};

    // A K step of the diagonal block: the triangular operand's stage is
    // cleared outside its triangle (whatever the other triangle holds never
    // enters) and set to one on a unit diagonal, 4 elements per thread,
    // before any wavefront reads it.
    auto mask_stage = [&](int st, int k0) {
        double* stage = TRI_A ? As[st] : Bs[st];
        // opaque copy of the thread index: keeps the compiler from holding
        // these addresses over the whole K loop (registers the kernel lacks)
        int t = tid;
        asm volatile("" : "+v"(t));
        const int x    = t & 127;
        const int unit = a.tri & GT_UNIT;
#pragma unroll
        for(int q = 0; q < 4; ++q)
        {
            const int k = (t >> 7) * 4 + q;
            const int e = k0 + k - ((TRI_A ? m0 : n0) + x); // k - row (A) / column (B)
            const int o = TRI_A ? OA::at(k, x) : OB::at(k, x);
            if(TO_END ? e > 0 : e < 0)
                stage[o] = 0.0;
            else if(unit && e == 0)
                stage[o] = 1.0;
        }
    };
    auto compute = [&](int st) {
        const double* as = As[st];
        const double* bs = Bs[st];
#pragma unroll
        for(int k4 = 0; k4 < BK / 4; ++k4)
        {
            const int kr = k4 * 4 + lane / 16;
            double    af[4], bf[4];
#pragma unroll
            for(int i = 0; i < 4; ++i)
                af[i] = as[OA::at(kr, wm * 64 + i * 16 + (lane & 15))];
#pragma unroll
            for(int j = 0; j < 4; ++j)
                bf[j] = bs[OB::at(kr, wn * 64 + j * 16 + (lane & 15))];
#pragma unroll
            for(int i = 0; i < 4; ++i)
#pragma unroll
                for(int j = 0; j < 4; ++j)
                    acc[i][j] = Mfma<double>::run(bf[j], af[i], acc[i][j]);
        }
    };
    // one K step: tiles kt .. kt + NS - 2 are in flight (LOADS each), so
    // waiting down to (NS - 2) * LOADS retires tile kt; the barrier
    // publishes it to every wavefront and retires every read of the stage
    // tile kt + NS - 1 reuses
    auto step = [&](int kt, int st) {
        wait_vmcnt<(NS - 2) * LOADS>();
        barrier_no_fence();
        issue(kt + NS - 1, (st + NS - 1) % NS);
        if constexpr(TRI != 0)
        {
            const int k0 = kbeg + kt * BK;
            if(k0 >= dlo && k0 < dhi)
            {
                mask_stage(st, k0);
                barrier_no_fence();
            }
        }
        compute(st);
    };

    if(ktiles > 0)
    {
#pragma unroll
        for(int t = 0; t < NS - 1; ++t)
            issue(t, t);
        for(int kt = 0; kt < ktiles; kt += NS)
        {
#pragma unroll
            for(int t = 0; t < NS; ++t)
                if(t == 0 || kt + t < ktiles)
                    step(kt + t, t);
        }
        wait_vmcnt<0>(); // the tiles past the end
    }

    if(part)
    {
#pragma unroll
        for(int i = 0; i < 4; ++i)
#pragma unroll
            for(int j = 0; j < 4; ++j)
#pragma unroll
                for(int r = 0; r < 4; ++r)
                    part[((i * 4 + j) * 4 + r) * 256 + tid] = acc[i][j][r];
        __syncthreads(); // LDS reuse by the next job
        return;
    }

    // Pieces of a tile add into C one after another, in piece order, so the
    // sum is the same whatever order they finish in: piece p waits until p
    // pieces have finished (a bounded poll by one thread, then every
    // wavefront acquires), piece 0 applies beta.  Earlier pieces have lower
    // workgroup ids and were dispatched first, so every wait ends.
    if constexpr(TRI != 0)
    {
        if(parts > 1 && piece > 0)
            split_wait(cnt, unsigned(piece) + a.need_bias, a.spin_limit, a.split_err);
    }

    // epilogue: as gemm_kernel's
#pragma unroll
    for(int i = 0; i < 4; ++i)
    {
        const int row = m0 + wm * 64 + i * 16 + (lane & 15);
#pragma unroll
        for(int j = 0; j < 4; ++j)
#pragma unroll
            for(int r = 0; r < 4; ++r)
            {
                const int col = n0 + wn * 64 + j * 16 + Mfma<double>::row(lane, r);
                if(row < a.M && col < a.N)
                {
                    double& cref = Cp[row + int64_t(col) * a.ldc];
                    double  v    = a.alpha * acc[i][j][r];
                    if(piece > 0)
                        v = cref + v;
                    else if(a.beta != 0.0)
                        v = fma(a.beta, cref, v);
                    cref = v;
                }
            }
    }
    if constexpr(TRI != 0)
    {
        if(parts > 1)
            split_signal(cnt, unsigned(parts));
    }
    __syncthreads(); // LDS reuse by the next job
}

// DMA GEMM, one workgroup per 128 x 128 tile (and instance, in y).  With a
// triangular operand (TRI != 0) a tile takes the K range in which the
// operand is non-zero for its rows (A) or columns (B), up to and including
// their diagonal block; with a split (kc > 0) that range is cut into
// P = ceil(len / kc) near-equal pieces, whole K steps each, piece p of tile t
// being workgroup t + p * tiles (none past P).
template <int BK, int NS, int OPA, int OPB, int TRI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    gemm_dma_kernel(const GemmArgsT<double, double, double> a)
{
    static_assert((BK == 8 || BK == 16) && NS >= 2 && NS <= 3, "configuration");
    constexpr bool TRI_A  = TRI == GT_A_LOWER || TRI == GT_A_UPPER;
    constexpr bool TO_END = TRI == GT_A_LOWER || TRI == GT_B_UPPER;
    __shared__ __align__(16) double As[NS][DmaOperand<BK, OPA == 0>::ELS];
    __shared__ __align__(16) double Bs[NS][DmaOperand<BK, OPB == 1>::ELS];

    if(a.skip_if_zero && *a.skip_if_zero == 0.0)
        return;

    // tile index: the same order as gemm_kernel -- XCD-contiguous runs of
    // row-panel groups, except with a triangular operand, whose row (column)
    // panels carry unequal work: contiguous runs would give one XCD all the
    // long panels (1.9x the mean at 16384^2), so dispatch order is kept there
    // (round-robin over the XCDs)
    const int total = a.tiles_m * a.tiles_n;
    int       bid   = blockIdx.x;
    int       piece = 0;
    if constexpr(TRI != 0)
    {
        piece = bid / total;
        bid -= piece * total;
    }
    else
    {
        const int q = total / 8, r = total % 8;
        const int x = bid % 8, l = bid / 8;
        bid         = x * q + min(x, r) + l;
    }
    const int     tile  = bid;
    constexpr int GROUP = 8;
    const int     grp   = bid / (GROUP * a.tiles_n);
    const int     fm    = grp * GROUP;
    const int     gsz   = min(a.tiles_m - fm, GROUP);
    const int     m0    = (fm + (bid % gsz)) * 128;
    const int     n0    = ((bid % (GROUP * a.tiles_n)) / gsz) * 128;
    int           kbeg = 0, kend = a.K, parts = 1, dlo = 0, dhi = 0;
    if constexpr(TRI != 0)
    {
        dlo = TRI_A ? m0 : n0;
        dhi = min(a.K, dlo + 128);
        if constexpr(TO_END)
            kend = dhi;
        else
            kbeg = dlo;
        const int len = max(kend - kbeg, 0);
        if(a.kc > 0 && len > a.kc)
        {
            parts        = (len + a.kc - 1) / a.kc;
            const int pl = ((len + parts - 1) / parts + BK - 1) / BK * BK;
            kbeg         = min(kend, kbeg + piece * pl);
            kend         = min(kend, kbeg + pl);
        }
        if(piece >= parts)
            return;
    }
    for(int b = blockIdx.y; b < a.batch; b += gridDim.y)
        dma_job<BK, NS, OPA, OPB, TRI>(
            a, As, Bs, b, m0, n0, kbeg, kend, dlo, dhi, piece, parts,
            parts > 1 ? a.split_cnt + (int64_t(b) * total + tile) : nullptr, nullptr);
}

// Balanced triangular products.  The tiles of every instance, ordered
// (instance, panel, tile in panel) with a panel the row (A) or column (B)
// block of the triangle, lay their K steps end to end, N in all; workgroup w
// of W (W <= N) takes steps [w N / W, (w + 1) N / W).  A tile wholly inside
// one workgroup's range is written to C; the pieces of a tile cut between
// workgroups go to the workspace, two slots of 128 x 128 per workgroup (its
// first tile, its last tile), and gemm_dma_reduce_kernel adds them in
// workgroup order.  Nothing waits and nothing is shared within the launch.
struct BalancedSplit
{
    int     panels, other, total;
    int64_t per_inst, n_all, W;

    template <int TRI>
    __device__ __forceinline__ static BalancedSplit of(const GemmArgsT<double, double, double>& a)
    {
        constexpr bool TRI_A = TRI == GT_A_LOWER || TRI == GT_A_UPPER;
        BalancedSplit  z{};
        z.panels = TRI_A ? a.tiles_m : a.tiles_n;
        z.other  = TRI_A ? a.tiles_n : a.tiles_m;
        z.total  = a.tiles_m * a.tiles_n;
        for(int q = 0; q < z.panels; ++q)
            z.per_inst += int64_t(z.other) * z.steps<TRI>(a, q);
        z.n_all = z.per_inst * a.batch;
        z.W     = a.split_w;
        return z;
    }
    // K steps (of 8) of the tiles in panel q
    template <int TRI>
    __device__ __forceinline__ static int steps(const GemmArgsT<double, double, double>& a, int q)
    {
        constexpr bool TO_END = TRI == GT_A_LOWER || TRI == GT_B_UPPER;
        const int      lo = q * 128, hi = min(a.K, lo + 128);
        return (TO_END ? hi : a.K - lo) / 8;
    }
    __device__ __forceinline__ int64_t first(int64_t w) const { return w * n_all / W; }
    // the workgroup whose range holds step x
    __device__ __forceinline__ int64_t owner(int64_t x) const
    {
        return ((x + 1) * W - 1) / n_all;
    }
};

template <int OPA, int OPB, int TRI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    gemm_dma_balanced_kernel(const GemmArgsT<double, double, double> a)
{
    constexpr int  BK     = 8;
    constexpr int  NS     = 3;
    constexpr bool TRI_A  = TRI == GT_A_LOWER || TRI == GT_A_UPPER;
    constexpr bool TO_END = TRI == GT_A_LOWER || TRI == GT_B_UPPER;
    static_assert(TRI != 0, "triangular operands only");
    __shared__ __align__(16) double As[NS][DmaOperand<BK, OPA == 0>::ELS];
    __shared__ __align__(16) double Bs[NS][DmaOperand<BK, OPB == 1>::ELS];

    const BalancedSplit z  = BalancedSplit::of<TRI>(a);
    const int64_t       wg = blockIdx.x;
    const int64_t       s0 = z.first(wg);
    int64_t             e  = z.first(wg + 1);
    while(e > s0)
    {
        // the tile holding step e - 1: instance b, panel q, tile i of the panel
        const int64_t x  = e - 1;
        const int     b  = int(x / z.per_inst);
        int64_t       r  = x - int64_t(b) * z.per_inst;
        int           q  = 0;
        int           ns = BalancedSplit::steps<TRI>(a, 0);
        while(r >= int64_t(z.other) * ns)
        {
            r -= int64_t(z.other) * ns;
            ++q;
            ns = BalancedSplit::steps<TRI>(a, q);
        }
        const int     i     = int(r / ns);
        const int64_t ts    = x - r % ns; // the tile's first step
        const int64_t js    = max(s0, ts);
        const bool    whole = js == ts && e == ts + ns;
        const int     dlo   = q * 128;
        const int     kb    = TO_END ? 0 : dlo;
        double*       part  = whole ? nullptr
                                    : a.split_ws + (2 * wg + (js == s0 ? 0 : 1)) * int64_t(16384);
        dma_job<BK, NS, OPA, OPB, TRI>(a, As, Bs, b, (TRI_A ? q : i) * 128, (TRI_A ? i : q) * 128,
                                       kb + int(js - ts) * BK, kb + int(e - ts) * BK, dlo,
                                       min(a.K, dlo + 128), 0, 1, nullptr, part);
        e = js;
    }
}

// The split tiles of a balanced product: C = alpha * (the pieces' sums, in
// workgroup order) + beta * C.  Four workgroups per tile and instance, each
// for a quarter of every thread's 64 values (more loads in flight than one
// workgroup per tile); tiles that one workgroup computed whole return at once.
template <int TRI>
__global__ void __launch_bounds__(256)
    gemm_dma_reduce_kernel(const GemmArgsT<double, double, double> a)
{
    constexpr bool      TRI_A = TRI == GT_A_LOWER || TRI == GT_A_UPPER;
    constexpr int       NV    = 16; // values per thread per workgroup
    const BalancedSplit z     = BalancedSplit::of<TRI>(a);
    const int           qtr   = blockIdx.x % 4;
    const int           tb    = blockIdx.x / 4;
    const int           b     = tb / z.total;
    const int           t     = tb - b * z.total;
    const int           q     = t / z.other;
    const int           i     = t - q * z.other;
    int64_t             ts    = int64_t(b) * z.per_inst;
    for(int p = 0; p < q; ++p)
        ts += int64_t(z.other) * BalancedSplit::steps<TRI>(a, p);
    const int     ns = BalancedSplit::steps<TRI>(a, q);
    ts += int64_t(i) * ns;
    const int64_t w0 = z.owner(ts), wl = z.owner(ts + ns - 1);
    if(w0 == wl)
        return;

    const int tid  = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w    = tid / WAVE;
    const int m0   = (TRI_A ? q : i) * 128;
    const int n0   = (TRI_A ? i : q) * 128;
    double*   Cp   = a.C.get(b);
    // the piece of workgroup g sits in its first-tile slot unless g's range
    // starts before this tile; a piece's 64 values load together
    auto piece = [&](int64_t g) {
        return a.split_ws + (2 * g + (z.first(g) >= ts ? 0 : 1)) * int64_t(16384) + tid;
    };
    double        sum[NV];
    const double* p0 = piece(w0) + qtr * NV * 256;
#pragma unroll
    for(int u = 0; u < NV; ++u)
        sum[u] = p0[u * 256];
    for(int64_t g = w0 + 1; g <= wl; ++g)
    {
        const double* pg = piece(g) + qtr * NV * 256;
#pragma unroll
        for(int u = 0; u < NV; ++u)
            sum[u] += pg[u * 256];
    }
#pragma unroll
    for(int u = 0; u < NV; ++u)
    {
        const int v  = qtr * NV + u;
        const int i4 = v / 16, j4 = (v / 4) % 4, r = v % 4;
        const int row = m0 + (w % 2) * 64 + i4 * 16 + (lane & 15);
        const int col = n0 + (w / 2) * 64 + j4 * 16 + Mfma<double>::row(lane, r);
        if(row < a.M && col < a.N)
        {
            double& cref = Cp[row + int64_t(col) * a.ldc];
            double  out  = a.alpha * sum[u];
            if(a.beta != 0.0)
                out = fma(a.beta, cref, out);
            cref = out;
        }
    }
}

// ---------------------------------------------------------------------------
// The FP32 form: four-float loads, so the layouts change.
//   x contiguous: BK rows (k) of 128 floats, two rows per instruction (lanes
//            0-31, 32-63); lane p of a row holds x quad p ^ 4 (k % 4), which
//            puts the four k rows a fragment read touches in distinct banks.
//   k contiguous: 128 rows (x) of BK floats, 64 / (BK/4) rows per
//            instruction; lane l holds row l / (BK/4) and the k quad
//            (l % (BK/4)) ^ ((x / (64/BK)) % (BK/4)), distinct banks for the
//            16 rows of a fragment read.
// Eligible when lda, ldb and every operand address are multiples of 4 floats
// / 16 bytes, K is a multiple of BK and the extent of an x-contiguous operand
// is a multiple of 4.
// ---------------------------------------------------------------------------
template <int BK, bool XC>
struct DmaOperandF
{
    static constexpr int ELS = BK * 128;
    static constexpr int NQ  = BK / 4;                      // k quads per row (k contiguous)
    static constexpr int RPI = XC ? 2 : 64 / NQ;            // rows per load instruction
    static constexpr int PER = (XC ? BK : 128) / RPI / 4;   // instructions per wavefront per tile
    static constexpr int G   = 64 / BK;                     // rows sharing a bank base (k contiguous)

    static __device__ __forceinline__ int at(int k, int x)
    {
        if constexpr(XC)
            return k * 128 + 4 * ((x >> 2) ^ (4 * (k & 3))) + (x & 3);
        else
            return x * BK + 4 * ((k >> 2) ^ ((x / G) % NQ)) + (k & 3);
    }

    struct Lane
    {
        int lane;
        int x0;
    };

    static __device__ __forceinline__ Lane lane_state(int, int lane, int x0, int)
    {
        return Lane{lane, x0};
    }

    static __device__ __forceinline__ const float*
        src(const float* base, const Lane& l, int g, int w, int64_t k0, int ld, int X)
    {
        const int r0 = (w * PER + g) * RPI; // first row (k for XC, x for k contiguous)
        if constexpr(XC)
        {
            const int r = r0 + l.lane / 32;
            const int q = (l.lane % 32) ^ (4 * (r & 3));
            const int x = min(l.x0 + 4 * q, X - 4);
            return base + x + (k0 + r) * ld;
        }
        else
        {
            const int xl = r0 + l.lane / NQ;
            const int kq = (l.lane % NQ) ^ ((xl / G) % NQ);
            return base + (k0 + 4 * kq) + int64_t(min(l.x0 + xl, X - 1)) * ld;
        }
    }

    static __device__ __forceinline__ int dst(int g, int w)
    {
        return (w * PER + g) * RPI * (XC ? 128 : BK);
    }
};

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm" // m0 is reserved; the clobber is still honoured
__device__ __forceinline__ void dma16(const float* src, const float* lds)
{
    const uint32_t dst = __builtin_amdgcn_readfirstlane(uint32_t(
        reinterpret_cast<uintptr_t>((const __attribute__((address_space(3))) float*)(lds))));
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off"
                 :
                 : "v"(src), "s"(dst)
                 : "memory", "m0");
}
#pragma clang diagnostic pop

template <int BK, int NS, int OPA, int OPB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    gemm_dma_f32_kernel(const GemmArgsT<float, float, float> a)
{
    static_assert((BK == 16 || BK == 32) && NS >= 2 && NS <= 3, "configuration");
    using OA            = DmaOperandF<BK, OPA == 0>;
    using OB            = DmaOperandF<BK, OPB == 1>;
    constexpr int LOADS = OA::PER + OB::PER; // loads per wavefront per tile
    __shared__ __align__(16) float As[NS][OA::ELS];
    __shared__ __align__(16) float Bs[NS][OB::ELS];

    if(a.skip_if_zero && *a.skip_if_zero == 0.0f)
        return;

    const int tid  = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w    = tid / WAVE;
    const int wm   = w % 2;
    const int wn   = w / 2;

    const int total = a.tiles_m * a.tiles_n;
    int       bid   = blockIdx.x;
    {
        const int q = total / 8, r = total % 8;
        const int x = bid % 8, l = bid / 8;
        bid         = x * q + min(x, r) + l;
    }
    constexpr int GROUP  = 8;
    const int     grp    = bid / (GROUP * a.tiles_n);
    const int     fm     = grp * GROUP;
    const int     gsz    = min(a.tiles_m - fm, GROUP);
    const int     m0     = (fm + (bid % gsz)) * 128;
    const int     n0     = ((bid % (GROUP * a.tiles_n)) / gsz) * 128;
    const int     ktiles = a.K / BK;

    const typename OA::Lane la = OA::lane_state(w, lane, m0, a.M);
    const typename OB::Lane lb = OB::lane_state(w, lane, n0, a.N);

    for(int b = blockIdx.y; b < a.batch; b += gridDim.y)
    {
        const float* Ap = a.A.get(b);
        const float* Bp = a.B.get(b);
        float*       Cp = a.C.get(b);

        auto issue = [&](int kt, int st) {
            const int64_t k0 = int64_t(min(kt, ktiles - 1)) * BK;
#pragma unroll
            for(int g = 0; g < OA::PER; ++g)
                dma16(OA::src(Ap, la, g, w, k0, a.lda, a.M), &As[st][OA::dst(g, w)]);
#pragma unroll
            for(int g = 0; g < OB::PER; ++g)
                dma16(OB::src(Bp, lb, g, w, k0, a.ldb, a.N), &Bs[st][OB::dst(g, w)]);
        };

        f4_t acc[4][4];
#pragma unroll
        for(int i = 0; i < 4; ++i)
#pragma unroll
            for(int j = 0; j < 4; ++j)
                acc[i][j] = f4_t{0, 0, 0, 0};

        auto compute = [&](int st) {
            const float* as = As[st];
            const float* bs = Bs[st];
#pragma unroll
            for(int k4 = 0; k4 < BK / 4; ++k4)
            {
                const int kr = k4 * 4 + lane / 16;
                float     af[4], bf[4];
#pragma unroll
                for(int i = 0; i < 4; ++i)
                    af[i] = as[OA::at(kr, wm * 64 + i * 16 + (lane & 15))];
#pragma unroll
                for(int j = 0; j < 4; ++j)
                    bf[j] = bs[OB::at(kr, wn * 64 + j * 16 + (lane & 15))];
#pragma unroll
                for(int i = 0; i < 4; ++i)
#pragma unroll
                    for(int j = 0; j < 4; ++j)
                        acc[i][j] = Mfma<float>::run(bf[j], af[i], acc[i][j]);
            }
        };
        auto step = [&](int kt, int st) {
            wait_vmcnt<(NS - 2) * LOADS>();
            barrier_no_fence();
            issue(kt + NS - 1, (st + NS - 1) % NS);
            compute(st);
        };

        if(ktiles > 0)
        {
#pragma unroll
            for(int t = 0; t < NS - 1; ++t)
                issue(t, t);
            for(int kt = 0; kt < ktiles; kt += NS)
            {
#pragma unroll
                for(int t = 0; t < NS; ++t)
                    if(t == 0 || kt + t < ktiles)
                        step(kt + t, t);
            }
            wait_vmcnt<0>();
        }

#pragma unroll
        for(int i = 0; i < 4; ++i)
        {
            const int row = m0 + wm * 64 + i * 16 + (lane & 15);
#pragma unroll
            for(int j = 0; j < 4; ++j)
#pragma unroll
                for(int r = 0; r < 4; ++r)
                {
                    const int col = n0 + wn * 64 + j * 16 + Mfma<float>::row(lane, r);
                    if(row < a.M && col < a.N)
                    {
                        float& cref = Cp[row + int64_t(col) * a.ldc];
                        float  v    = a.alpha * acc[i][j][r];
                        if(a.beta != 0.0f)
                            v = fmaf(a.beta, cref, v);
                        cref = v;
                    }
                }
        }
        __syncthreads(); // LDS reuse across batch iterations
    }
}

static bool gemm_dma_eligible(const GemmArgsT<float, float, float>& a, int bk)
{
    auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    // an x-contiguous operand is loaded four rows of op() at a time
    const bool a_quads = a.opA == 1 || (a.M >= 4 && a.M % 4 == 0);
    const bool b_quads = a.opB == 0 || (a.N >= 4 && a.N % 4 == 0);
    return !a.tri && a_quads && b_quads && a.K >= bk && a.K % bk == 0 && a.lda % 4 == 0
           && a.ldb % 4 == 0 && !a.A.arr && !a.B.arr && al16(a.A.base + a.A.offset)
           && al16(a.B.base + a.B.offset) && a.A.stride % 4 == 0 && a.B.stride % 4 == 0;
}

// The DMA kernel's conditions: plain FP64 operands, 16-byte aligned rows; at
// most one triangular operand, square (K x K), and then at most one of the
// two operands transposed.
static bool gemm_dma_eligible(const GemmArgsT<double, double, double>& a, int bk)
{
    auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    // an x-contiguous operand is loaded two rows of op() at a time
    const bool a_pairs = a.opA == 1 || (a.M >= 2 && a.M % 2 == 0);
    const bool b_pairs = a.opB == 0 || (a.N >= 2 && a.N % 2 == 0);
    const int  kind    = a.tri & (GT_A_LOWER | GT_A_UPPER | GT_B_LOWER | GT_B_UPPER);
    const bool tri_ok  = !a.tri
                        || ((kind == GT_A_LOWER || kind == GT_A_UPPER || kind == GT_B_LOWER
                             || kind == GT_B_UPPER)
                            && !(a.opA == 1 && a.opB == 1)
                            && (!(a.tri & (GT_A_LOWER | GT_A_UPPER)) || a.M == a.K)
                            && (!(a.tri & (GT_B_LOWER | GT_B_UPPER)) || a.N == a.K));
    return tri_ok && a_pairs && b_pairs && a.K >= bk && a.K % bk == 0 && a.lda % 2 == 0
           && a.ldb % 2 == 0 && !a.A.arr && !a.B.arr
           && al16(a.A.base + a.A.offset) && al16(a.B.base + a.B.offset) && a.A.stride % 2 == 0
           && a.B.stride % 2 == 0;
}

// MFMA type of an operand pair: FP64 only when both operands are FP64.
template <typename TA, typename TB>
using mfma_t = std::conditional_t<std::is_same_v<TA, double> && std::is_same_v<TB, double>,
                                  double, float>;

template <typename TA, typename TB, typename TC, bool TRI, int BM, int BN, int BK, int WGM,
          int WGN>
static void launch_gemm_ops(rocblas_handle h, const GemmArgsT<TA, TB, TC>& a, dim3 grid)
{
    using T = mfma_t<TA, TB>;
    // operand layouts the callers use: FP32-copy updates are op(A) x B
    // (side left) or X x op(A) (side right)
    constexpr bool any_ops = std::is_same_v<TA, TB>;
    constexpr bool ta_ok   = any_ops || std::is_same_v<TA, float>;
    constexpr bool tb_ok   = any_ops || std::is_same_v<TB, float>;
    const dim3     block(WAVE * WGM * WGN);
    if(a.opA == 0 && a.opB == 0)
        hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, T, TRI, 0, 0, BM, BN, BK, WGM, WGN>), grid,
                           block, 0, h->stream, a);
    else if(a.opA == 1 && a.opB == 0)
    {
        if constexpr(ta_ok)
            hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, T, TRI, 1, 0, BM, BN, BK, WGM, WGN>), grid,
                               block, 0, h->stream, a);
    }
    else if(a.opA == 0 && a.opB == 1)
    {
        if constexpr(tb_ok)
            hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, T, TRI, 0, 1, BM, BN, BK, WGM, WGN>), grid,
                               block, 0, h->stream, a);
    }
    else
    {
        if constexpr(any_ops)
            hipLaunchKernelGGL((gemm_kernel<TA, TB, TC, T, TRI, 1, 1, BM, BN, BK, WGM, WGN>), grid,
                               block, 0, h->stream, a);
    }
}

template <typename TA, typename TB, typename TC>
static bool gemm_ops_supported(const GemmArgsT<TA, TB, TC>& a)
{
    if constexpr(std::is_same_v<TA, TB>)
        return true; // every transpose pair, masked or not (the inversion's doubling steps
                     // take triangular op(A) = X^T with op(B) = A^T)
    else if constexpr(std::is_same_v<TA, float>)
        return a.opB == 0 && !a.tri;
    else
        return a.opA == 0 && !a.tri;
}

template <typename TA, typename TB, typename TC, int BM, int BN, int BK, int WGM, int WGN>
static rocblas_status launch_gemm_cfg(rocblas_handle h, GemmArgsT<TA, TB, TC> a)
{
    a.tiles_m        = (a.M + BM - 1) / BM;
    a.tiles_n        = (a.N + BN - 1) / BN;
    const int64_t nt = int64_t(a.tiles_m) * a.tiles_n;
    if(nt > int64_t(0x7fffffff))
        return rocblas_status_invalid_size;
    if(!gemm_ops_supported(a))
        return rocblas_status_not_implemented;
    const dim3 grid(unsigned(nt), grid_batch(a.batch));
    if(a.tri)
    {
        if constexpr(std::is_same_v<TA, TB> && std::is_same_v<TA, TC>)
            launch_gemm_ops<TA, TB, TC, true, BM, BN, BK, WGM, WGN>(h, a, grid);
    }
    else
        launch_gemm_ops<TA, TB, TC, false, BM, BN, BK, WGM, WGN>(h, a, grid);
    return launch_status();
}

template <int BK, int NS, int TRI>
static void launch_gemm_dma_ops(rocblas_handle h, const GemmArgsT<double, double, double>& d,
                                dim3 grid)
{
    if(d.opA == 0 && d.opB == 0)
        hipLaunchKernelGGL((gemm_dma_kernel<BK, NS, 0, 0, TRI>), grid, dim3(256), 0, h->stream,
                           d);
    else if(d.opA == 1 && d.opB == 0)
        hipLaunchKernelGGL((gemm_dma_kernel<BK, NS, 1, 0, TRI>), grid, dim3(256), 0, h->stream,
                           d);
    else if(d.opA == 0 && d.opB == 1)
        hipLaunchKernelGGL((gemm_dma_kernel<BK, NS, 0, 1, TRI>), grid, dim3(256), 0, h->stream,
                           d);
    else if constexpr(TRI == 0)
        hipLaunchKernelGGL((gemm_dma_kernel<BK, NS, 1, 1, TRI>), grid, dim3(256), 0, h->stream,
                           d);
}

template <int BK, int NS>
static void launch_gemm_dma_f32_ops(rocblas_handle h, const GemmArgsT<float, float, float>& d,
                                    dim3 grid)
{
    if(d.opA == 0 && d.opB == 0)
        hipLaunchKernelGGL((gemm_dma_f32_kernel<BK, NS, 0, 0>), grid, dim3(256), 0, h->stream, d);
    else if(d.opA == 1 && d.opB == 0)
        hipLaunchKernelGGL((gemm_dma_f32_kernel<BK, NS, 1, 0>), grid, dim3(256), 0, h->stream, d);
    else if(d.opA == 0 && d.opB == 1)
        hipLaunchKernelGGL((gemm_dma_f32_kernel<BK, NS, 0, 1>), grid, dim3(256), 0, h->stream, d);
    else
        hipLaunchKernelGGL((gemm_dma_f32_kernel<BK, NS, 1, 1>), grid, dim3(256), 0, h->stream, d);
}

static void launch_gemm_dma(rocblas_handle h, const GemmArgsT<float, float, float>& d, dim3 grid,
                            int bk)
{
    if(bk == 32)
        launch_gemm_dma_f32_ops<32, 2>(h, d, grid);
    else
        launch_gemm_dma_f32_ops<16, 3>(h, d, grid);
}

static void launch_gemm_dma(rocblas_handle h, const GemmArgsT<double, double, double>& d,
                            dim3 grid, int bk)
{
    // triangular operands: the 8-deep kernel only (the 16-deep one needs
    // more registers than its two waves per SIMD allow once the per-tile K
    // range is added, and would spill)
    switch(d.tri & (GT_A_LOWER | GT_A_UPPER | GT_B_LOWER | GT_B_UPPER))
    {
    case GT_A_LOWER: launch_gemm_dma_ops<8, 3, GT_A_LOWER>(h, d, grid); break;
    case GT_A_UPPER: launch_gemm_dma_ops<8, 3, GT_A_UPPER>(h, d, grid); break;
    case GT_B_LOWER: launch_gemm_dma_ops<8, 3, GT_B_LOWER>(h, d, grid); break;
    case GT_B_UPPER: launch_gemm_dma_ops<8, 3, GT_B_UPPER>(h, d, grid); break;
    default:
        if(bk == 16)
            launch_gemm_dma_ops<16, 2, 0>(h, d, grid);
        else
            launch_gemm_dma_ops<8, 3, 0>(h, d, grid);
    }
}

// ROCBLAS_AMD_GEMM_DMA (for A/B comparisons): 0 keeps every GEMM on
// gemm_kernel, 8 uses only the shallower DMA kernels (FP64 8-deep with three
// stages, FP32 16-deep with three stages); by default the deeper two-stage
// ones run where K allows (8192 x 16384 x 8192 on MI355X: FP64 69.4 TF/s,
// 8-deep 68.8, gemm_kernel 61.5; FP32 132.4, 16-deep 133, gemm_kernel 106.5).
static int gemm_dma_mode()
{
    static const int mode = [] {
        const char* e = std::getenv("ROCBLAS_AMD_GEMM_DMA");
        if(!e)
            return 16;
        const int v = std::atoi(e);
        return v == 0 ? 0 : v == 8 ? 8 : 16;
    }();
    return mode;
}

// ROCBLAS_AMD_GEMM_SPLIT (for A/B comparisons): 0 never splits K; otherwise
// the shortest K piece a split may make (a multiple of 128, default 128).
static int gemm_split_min()
{
    static const int v = [] {
        const char* e = std::getenv("ROCBLAS_AMD_GEMM_SPLIT");
        if(!e)
            return 128;
        const int x = std::atoi(e);
        return x <= 0 ? 0 : std::max(128, x / 128 * 128);
    }();
    return v;
}

// A triangular-operand product with few 128 x 128 tiles on the DMA kernel,
// each tile's K range cut into pieces so that the launch fills the chip's two
// workgroups per CU: the longest piece kc is the shortest multiple of 128
// (not below the minimum) for which the pieces of every instance fit in
// those slots; without such a kc the launch is not split.  The pieces of a
// tile add into C in piece order (gemm_dma_kernel), so the result depends
// only on the shapes and the CU count, not on timing.
static rocblas_status launch_gemm_dma_split(rocblas_handle h, const GemmArgsT<double, double, double>& a)
{
    GemmArgsT<double, double, double> d = a;
    d.tiles_m                           = (a.M + 127) / 128;
    d.tiles_n                           = (a.N + 127) / 128;
    const int64_t tiles                 = int64_t(d.tiles_m) * d.tiles_n;
    const int64_t slots                 = 2 * int64_t(h->num_cus);
    const bool    onA                   = a.tri & (GT_A_LOWER | GT_A_UPPER);
    const bool    toK                   = a.tri & (GT_A_LOWER | GT_B_UPPER); // range ends at the diagonal
    // the K range of the tiles in row (A) / column (B) panel t, as the kernel takes it
    auto len_of = [&](int t) {
        const int lo = t * 128, hi = std::min(a.K, lo + 128);
        return toK ? hi : a.K - lo;
    };
    const int panels = onA ? d.tiles_m : d.tiles_n;
    const int other  = onA ? d.tiles_n : d.tiles_m;
    auto parts = [&](int len, int kc) { return len > kc ? (len + kc - 1) / kc : 1; };
    d.kc      = 0;
    int pmax  = 1;
    const int kmin = gemm_split_min();
    if(kmin && split_counters(h) && h->async_error && tiles * a.batch <= kSplitCounters
       && a.batch <= 65535)
        for(int kc = kmin; kc < a.K; kc += 128)
        {
            int64_t n  = 0;
            int     pm = 1;
            for(int t = 0; t < panels; ++t)
            {
                const int p = parts(len_of(t), kc);
                n += int64_t(p) * other;
                pm = std::max(pm, p);
            }
            if(n * a.batch <= slots)
            {
                if(pm > 1)
                    d.kc = kc, pmax = pm;
                break;
            }
        }
    d.split_cnt  = split_counters(h);
    d.split_err  = h->async_error;
    d.spin_limit = h->chain_debug_timeout ? 64 : (1 << 20);
    d.need_bias  = h->chain_debug_timeout ? (1u << 30) : 0u;
    const dim3 grid(unsigned(tiles * pmax), grid_batch(a.batch));
    launch_gemm_dma(h, d, grid, 8);
    return launch_status();
}

// Shortest K of a triangular product balanced over the workgroups whatever
// its tile count (half of it with 64 tiles or more); shorter ones run on the
// masked register-staged kernels, which measured as fast or faster there.
// ROCBLAS_AMD_BALANCED_MIN_K overrides it for A/B comparisons.
static int balanced_min_k()
{
    static const int v = [] {
        const char* e = std::getenv("ROCBLAS_AMD_BALANCED_MIN_K");
        return e ? std::max(128, std::atoi(e)) : 1024;
    }();
    return v;
}

// Workgroups of a balanced launch of a triangular product by shape (0: the
// shape does not qualify): every K step of its tiles, over
// min(2 x CUs, steps / the shortest piece) workgroups.
static int64_t balanced_shape_workgroups(rocblas_handle h, int M, int N, int K, int tri, int batch)
{
    const int kind = tri & (GT_A_LOWER | GT_A_UPPER | GT_B_LOWER | GT_B_UPPER);
    if(!kind || !gemm_dma_mode() || !gemm_split_min() || M < 128 || N < 128
       || K < balanced_min_k() / 2 || K % 8)
        return 0;
    const int64_t tm = (M + 127) / 128, tn = (N + 127) / 128;
    if(tm * tn * batch >= 2 * int64_t(h->num_cus) || tm * tn * batch > kSplitCounters)
        return 0;
    // between half the minimum K and the minimum, only with 64 tiles or more
    // (n = 3000 lower, K = 952 over 128 tiles: 0.72 -> 0.62 ms; K = 512 over
    // 16 or 32 tiles ran slower than the masked kernels)
    if(K < balanced_min_k() && tm * tn * batch < 64)
        return 0;
    const bool onA   = kind == GT_A_LOWER || kind == GT_A_UPPER;
    const bool toEnd = kind == GT_A_LOWER || kind == GT_B_UPPER;
    int64_t    steps = 0;
    for(int64_t q = 0; q < (onA ? tm : tn); ++q)
    {
        const int64_t lo = q * 128, hi = std::min<int64_t>(K, lo + 128);
        steps += (onA ? tn : tm) * ((toEnd ? hi : K - lo) / 8);
    }
    steps *= batch;
    const int64_t W = std::min<int64_t>(2 * int64_t(h->num_cus),
                                        std::max<int64_t>(1, steps / (gemm_split_min() / 8)));
    return W <= steps ? W : 0;
}

// The balanced launch of `a`, if its shape qualifies, the caller allows a
// split and the workspace holds its 2 W partial tiles.
static int64_t balanced_workgroups(rocblas_handle h, const GemmArgsT<double, double, double>& a)
{
    if(!a.tri || !a.split_ok || !a.split_ws || !gemm_dma_eligible(a, 8))
        return 0;
    const int64_t W = balanced_shape_workgroups(h, a.M, a.N, a.K, a.tri, a.batch);
    return 2 * W <= a.split_ws_tiles ? W : 0;
}

size_t gemm_tri_split_bytes(rocblas_handle h, int M, int N, int K, int tri, int batch)
{
    return size_t(2 * balanced_shape_workgroups(h, M, N, K, tri, batch)) * 128 * 128
           * sizeof(double);
}

template <int OPA, int OPB, int TRI>
static void launch_balanced_ops(rocblas_handle h, const GemmArgsT<double, double, double>& d)
{
    hipLaunchKernelGGL((gemm_dma_balanced_kernel<OPA, OPB, TRI>), dim3(unsigned(d.split_w)),
                       dim3(256), 0, h->stream, d);
    hipLaunchKernelGGL((gemm_dma_reduce_kernel<TRI>),
                       dim3(unsigned(4 * int64_t(d.tiles_m) * d.tiles_n * d.batch)), dim3(256), 0,
                       h->stream, d);
}

template <int TRI>
static void launch_balanced_kind(rocblas_handle h, const GemmArgsT<double, double, double>& d)
{
    if(d.opA == 0 && d.opB == 0)
        launch_balanced_ops<0, 0, TRI>(h, d);
    else if(d.opA == 1)
        launch_balanced_ops<1, 0, TRI>(h, d);
    else
        launch_balanced_ops<0, 1, TRI>(h, d);
}

// A triangular product with few tiles, its K steps balanced over W
// workgroups (gemm_dma_balanced_kernel) and the split tiles summed in
// workgroup order (gemm_dma_reduce_kernel): the result depends only on the
// shapes and the CU count.
static rocblas_status launch_gemm_dma_balanced(rocblas_handle h,
                                               const GemmArgsT<double, double, double>& a,
                                               int64_t W)
{
    GemmArgsT<double, double, double> d = a;
    d.tiles_m                           = (a.M + 127) / 128;
    d.tiles_n                           = (a.N + 127) / 128;
    d.split_w                           = int(W);
    switch(a.tri & (GT_A_LOWER | GT_A_UPPER | GT_B_LOWER | GT_B_UPPER))
    {
    case GT_A_LOWER: launch_balanced_kind<GT_A_LOWER>(h, d); break;
    case GT_A_UPPER: launch_balanced_kind<GT_A_UPPER>(h, d); break;
    case GT_B_LOWER: launch_balanced_kind<GT_B_LOWER>(h, d); break;
    default: launch_balanced_kind<GT_B_UPPER>(h, d); break;
    }
    return launch_status();
}

// Pick the largest tile that still gives every CU work.
template <typename TA, typename TB, typename TC>
static rocblas_status launch_gemm_sel(rocblas_handle h, const GemmArgsT<TA, TB, TC>& a)
{
    if(a.M <= 0 || a.N <= 0)
        return rocblas_status_success;
    const int64_t cus   = h->num_cus;
    auto          tiles = [&](int bm, int bn) {
        return int64_t((a.M + bm - 1) / bm) * ((a.N + bn - 1) / bn) * a.batch;
    };
    if(a.N <= 32)
        return tiles(128, 32) >= cus ? launch_gemm_cfg<TA, TB, TC, 128, 32, 16, 4, 1>(h, a)
                                     : launch_gemm_cfg<TA, TB, TC, 32, 32, 16, 2, 2>(h, a);
    if(a.M <= 32)
        return tiles(32, 128) >= cus ? launch_gemm_cfg<TA, TB, TC, 32, 128, 16, 1, 4>(h, a)
                                     : launch_gemm_cfg<TA, TB, TC, 32, 32, 16, 2, 2>(h, a);
    // triangular products whose tiles alone would leave workgroup slots idle,
    // where the caller allows a split: with a workspace, the K steps balanced
    // over the workgroups; without one, a single product with 128-511 tiles
    // on the DMA kernel with its tiles' K ranges split (fewer tiles: the
    // pieces would be too short for the DMA kernel's per-step latency, and
    // the masked kernels below run faster)
    if constexpr(std::is_same_v<TA, double> && std::is_same_v<TB, double>
                 && std::is_same_v<TC, double>)
    {
        if(const int64_t W = balanced_workgroups(h, a))
            return launch_gemm_dma_balanced(h, a, W);
        if(a.tri && a.split_ok && a.batch == 1 && gemm_dma_mode() && tiles(128, 128) * 2 >= cus
           && tiles(128, 128) < 2 * cus && gemm_dma_eligible(a, 8))
            return launch_gemm_dma_split(h, a);
    }
    if(tiles(128, 128) >= cus)
    {
        // FP64 runs best with K steps of 8 (two waves per SIMD), FP32 with 16
        // (measured at 8192x16384x8192 on MI355X: 65 vs 50 TF/s, 96 vs 94 TF/s)
        if constexpr(std::is_same_v<TA, float> && std::is_same_v<TB, float>
                     && std::is_same_v<TC, float>)
        {
            const int mode = gemm_dma_mode();
            const int bk   = mode != 8 && a.K % 32 == 0 ? 32 : 16;
            if(mode && gemm_dma_eligible(a, bk))
            {
                GemmArgsT<TA, TB, TC> d = a;
                d.tiles_m               = (a.M + 127) / 128;
                d.tiles_n               = (a.N + 127) / 128;
                const dim3 grid(unsigned(int64_t(d.tiles_m) * d.tiles_n), grid_batch(a.batch));
                launch_gemm_dma(h, d, grid, bk);
                return launch_status();
            }
        }
        if constexpr(std::is_same_v<TA, double> && std::is_same_v<TB, double>
                     && std::is_same_v<TC, double>)
        {
            const int mode = gemm_dma_mode();
            const int bk   = mode == 16 && a.K % 16 == 0 && !a.tri ? 16 : 8;
            if(mode && gemm_dma_eligible(a, bk))
            {
                GemmArgsT<TA, TB, TC> d = a;
                d.tiles_m               = (a.M + 127) / 128;
                d.tiles_n               = (a.N + 127) / 128;
                const dim3 grid(unsigned(int64_t(d.tiles_m) * d.tiles_n), grid_batch(a.batch));
                d.kc                    = 0;
                launch_gemm_dma(h, d, grid, bk);
                return launch_status();
            }
        }
        if constexpr(std::is_same_v<mfma_t<TA, TB>, double>)
            return launch_gemm_cfg<TA, TB, TC, 128, 128, 8, 2, 2>(h, a);
        else
            return launch_gemm_cfg<TA, TB, TC, 128, 128, 16, 2, 2>(h, a);
    }
    if(tiles(64, 64) >= cus)
        return launch_gemm_cfg<TA, TB, TC, 64, 64, 16, 2, 2>(h, a);
    return launch_gemm_cfg<TA, TB, TC, 32, 32, 16, 2, 2>(h, a);
}

template <typename TA, typename TB, typename TC>
rocblas_status launch_gemm(rocblas_handle h, const GemmArgsT<TA, TB, TC>& a)
{
    return launch_gemm_sel(h, a);
}

template rocblas_status launch_gemm(rocblas_handle, const GemmArgsT<float, float, float>&);
template rocblas_status launch_gemm(rocblas_handle, const GemmArgsT<double, double, double>&);
template rocblas_status launch_gemm(rocblas_handle, const GemmArgsT<float, double, double>&);
template rocblas_status launch_gemm(rocblas_handle, const GemmArgsT<double, float, double>&);

} // namespace rbamd
