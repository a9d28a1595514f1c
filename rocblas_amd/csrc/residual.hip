// residual.hip -- normwise backward error of many solved instances at once:
//
//   eta_i = ||alpha_i B0_i - op(A_i) X_i||_1
//           / (||op(A_i)||_1 ||X_i||_1 + |alpha_i| ||B0_i||_1)
//
// (side right: X_i op(A_i)), the acceptance measure of the solve, in FP64
// whatever the storage type.  Instances have their own sizes, cases and alpha;
// the work is split into items (instance, 16-vector panel p, 64-row block I)
// that a grid of 256-thread workgroups takes one each:
//
//   r_I = alpha b0_I - sum_{J<=I} L_IJ x_J      (FP64 MFMA, canonical form)
//
// with the canonical form of the solve (forward substitution on a canonical
// lower triangle L; side L: vectors are columns of B, L = op(A) up to index
// reversal; side R: vectors are rows of B, L = op(A)^T up to reversal).  Each
// item adds |r|, |x_I| and |b0_I| into per-column sums of its instance (the
// columns of B: the vectors for side L, the canonical rows for side R; a
// reversal only permutes columns, which the max does not see), and the items
// of panel 0 add |L_IJ| into the column sums of op(A) (side L: columns of L;
// side R: rows of L).  A last launch takes the maxima and forms eta.  The sums
// are FP64 atomic adds, so their order -- and eta's last bits -- can vary from
// run to run; eta is a measurement, not a result.
#include "device_common.hpp"
#include "panel_trsm.hpp"
#include "residual.hpp"

namespace rbamd
{
namespace
{
using namespace panel;

__global__ void __launch_bounds__(256) eta_zero_kernel(double* p, int64_t n)
{
    for(int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        p[i] = 0.0;
}

template <typename T>
__global__ void __launch_bounds__(PNT) eta_items_kernel(const EtaArgs<T> a)
{
    __shared__ double Ls[PNB * PLDA]; // tile L_IJ (masked, diagonal included); then |b0_I|
    __shared__ double Yj[PNB * PLDY]; // x_J, 64 rows x 16 vectors; x_I after the loop
    __shared__ double Cs[PNB * PLDY]; // sum_J L_IJ x_J, then |r|
    double* const     Xa = Yj;
    double* const     Ba = Ls;

    const int item = blockIdx.x;
    const int inst = a.items[3 * item];
    const int p    = a.items[3 * item + 1];
    const int I    = a.items[3 * item + 2];

    const int  K     = a.K[inst];
    const int  nvec  = a.nvec[inst];
    const int  flags = a.flags[inst];
    const int  lda   = a.lda[inst];
    const int  ldx   = a.ldx[inst];
    const int  ldb   = a.ldb[inst];
    const bool tA    = flags & SF_TA;
    const bool rev   = flags & SF_REV;
    const bool unit  = flags & SF_UNIT;
    const bool cols  = flags & SF_COLS;
    const T*   Ap    = a.A[inst];
    const T*   Xp    = a.X[inst];
    const T*   Bp    = a.B0[inst];
    const double alpha = double(a.alpha[inst]);
    double* const acc  = a.acc + a.acc_off[inst];
    const int     ncol = cols ? nvec : K;
    double* const accR = acc;
    double* const accX = acc + ncol;
    double* const accB = acc + 2 * ncol;
    double* const accA = acc + 3 * ncol;

    const int tid  = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int w    = tid / WAVE;
    const int v0   = p * PNC;

    // element (canonical row i, vector v) of an m x n operand of the solve
    auto boff = [&](int i, int v, int ld) -> int64_t {
        const int r = rev ? K - 1 - i : i;
        return cols ? int64_t(r) + int64_t(v) * ld : int64_t(v) + int64_t(r) * ld;
    };
    // this thread's (row, vector) pairs of a 64 x 16 block, ordered so that
    // consecutive threads read consecutive addresses
    auto blk_iv = [&](int s, int& i, int& v) {
        const int e = tid + PNT * s;
        if(cols)
        {
            i = e % PNB;
            v = e / PNB;
        }
        else
        {
            v = e % PNC;
            i = e / PNC;
        }
    };

    typename Mfma<double>::acc_t acc0{0, 0, 0, 0}, acc1{0, 0, 0, 0};
    for(int J = 0; J <= I; ++J)
    {
        // tile L_IJ, canonical (i, j) walked in A's storage order
#pragma unroll
        for(int s = 0; s < PEPT; ++s)
        {
            const int e = tid + PNT * s;
            const int i = tA ? e / PNB : e % PNB;
            const int j = tA ? e % PNB : e / PNB;
            const int gi = I * PNB + i, gj = J * PNB + j;
            double    v  = 0.0;
            if(gi < K && gj <= gi)
            {
                if(gi == gj && unit)
                    v = 1.0;
                else
                {
                    const int mi = rev ? K - 1 - gi : gi, mj = rev ? K - 1 - gj : gj;
                    v = double(gload(Ap + (tA ? int64_t(mj) + int64_t(mi) * lda
                                              : int64_t(mi) + int64_t(mj) * lda)));
                }
            }
            Ls[i * PLDA + j] = v;
        }
        // x_J
#pragma unroll
        for(int s = 0; s < 4; ++s)
        {
            int i, v;
            blk_iv(s, i, v);
            const int gi = J * PNB + i, gv = v0 + v;
            Yj[i * PLDY + v] = (gi < K && gv < nvec) ? double(gload(Xp + boff(gi, gv, ldx)))
                                                     : 0.0;
        }
        __syncthreads();
        if(p == 0 && tid < PNB)
        {
            // |L_IJ| into the column sums of op(A)
            double sum = 0.0;
            if(cols)
            {
                for(int i = 0; i < PNB; ++i)
                    sum += fabs(Ls[i * PLDA + tid]);
                if(J * PNB + tid < K)
                    atomicAdd(accA + J * PNB + tid, sum);
            }
            else
            {
                for(int j = 0; j < PNB; ++j)
                    sum += fabs(Ls[tid * PLDA + j]);
                if(I * PNB + tid < K)
                    atomicAdd(accA + I * PNB + tid, sum);
            }
        }
#pragma unroll
        for(int k4 = 0; k4 < PNB / 4; k4 += 2)
        {
            const int kk0 = 4 * k4 + (lane >> 4);
            const int kk1 = kk0 + 4;
            acc0 = Mfma<double>::run(Ls[(16 * w + (lane & 15)) * PLDA + kk0],
                                     Yj[kk0 * PLDY + (lane & 15)], acc0);
            acc1 = Mfma<double>::run(Ls[(16 * w + (lane & 15)) * PLDA + kk1],
                                     Yj[kk1 * PLDY + (lane & 15)], acc1);
        }
        __syncthreads();
    }
#pragma unroll
    for(int r = 0; r < 4; ++r)
        Cs[(16 * w + Mfma<double>::row(lane, r)) * PLDY + (lane & 15)] = acc0[r] + acc1[r];
    __syncthreads();
    // |alpha b0 - L x|, |x| (x_I is the last block staged), |b0| of this block
#pragma unroll
    for(int s = 0; s < 4; ++s)
    {
        int i, v;
        blk_iv(s, i, v);
        const int  gi = I * PNB + i, gv = v0 + v;
        const bool ok = gi < K && gv < nvec;
        const double b = ok ? double(gload(Bp + boff(gi, gv, ldb))) : 0.0;
        Cs[i * PLDY + v] = ok ? fabs(alpha * b - Cs[i * PLDY + v]) : 0.0;
        Xa[i * PLDY + v] = fabs(Xa[i * PLDY + v]);
        Ba[i * PLDY + v] = fabs(b);
    }
    __syncthreads();
    if(cols)
    {
        // one column per vector: sums over the block's 64 rows
        if(tid < 3 * PNC)
        {
            const int     v   = tid % PNC;
            const double* src = tid < PNC ? Cs : tid < 2 * PNC ? Xa : Ba;
            double*       dst = tid < PNC ? accR : tid < 2 * PNC ? accX : accB;
            double        sum = 0.0;
            for(int i = 0; i < PNB; ++i)
                sum += src[i * PLDY + v];
            if(v0 + v < nvec)
                atomicAdd(dst + v0 + v, sum);
        }
    }
    else if(tid < 3 * PNB)
    {
        // one column per canonical row: sums over the panel's 16 vectors
        const int     i   = tid % PNB;
        const double* src = tid < PNB ? Cs : tid < 2 * PNB ? Xa : Ba;
        double*       dst = tid < PNB ? accR : tid < 2 * PNB ? accX : accB;
        double        sum = 0.0;
        for(int v = 0; v < PNC; ++v)
            sum += src[i * PLDY + v];
        if(I * PNB + i < K)
            atomicAdd(dst + I * PNB + i, sum);
    }
}

template <typename T>
__global__ void __launch_bounds__(256) eta_final_kernel(const EtaArgs<T> a)
{
    __shared__ double red[4][256];
    const int inst = blockIdx.x;
    const int K    = a.K[inst];
    const int nvec = a.nvec[inst];
    const int ncol = (a.flags[inst] & SF_COLS) ? nvec : K;
    const double* acc = a.acc + a.acc_off[inst];
    double mr = 0.0, mx = 0.0, mb = 0.0, ma = 0.0;
    for(int c = threadIdx.x; c < ncol; c += 256)
    {
        mr = fmax(mr, acc[c]);
        mx = fmax(mx, acc[ncol + c]);
        mb = fmax(mb, acc[2 * ncol + c]);
    }
    for(int c = threadIdx.x; c < K; c += 256)
        ma = fmax(ma, acc[3 * ncol + c]);
    red[0][threadIdx.x] = mr;
    red[1][threadIdx.x] = mx;
    red[2][threadIdx.x] = mb;
    red[3][threadIdx.x] = ma;
    __syncthreads();
    for(int s = 128; s > 0; s >>= 1)
    {
        if(threadIdx.x < s)
            for(int q = 0; q < 4; ++q)
                red[q][threadIdx.x] = fmax(red[q][threadIdx.x], red[q][threadIdx.x + s]);
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        const double alpha = fabs(double(a.alpha[inst]));
        const double den   = red[3][0] * red[1][0] + alpha * red[2][0];
        const double num   = red[0][0];
        a.eta[inst]        = den > 0.0 ? num / den : (num > 0.0 ? INFINITY : 0.0);
    }
}
} // namespace

int64_t eta_acc_doubles(int K, int nvec, bool side_left)
{
    return 3 * int64_t(side_left ? nvec : K) + K;
}

int64_t eta_items(int K, int nvec)
{
    return int64_t((K + PNB - 1) / PNB) * ((nvec + PNC - 1) / PNC);
}

void eta_fill_items(int inst, int K, int nvec, int* out)
{
    const int R = (K + PNB - 1) / PNB, P = (nvec + PNC - 1) / PNC;
    for(int p = 0; p < P; ++p)
        for(int I = 0; I < R; ++I)
        {
            *out++ = inst;
            *out++ = p;
            *out++ = I;
        }
}

template <typename T>
rocblas_status eta_launch(hipStream_t stream, const EtaArgs<T>& a)
{
    if(a.batch <= 0)
        return rocblas_status_success;
    const unsigned zb = unsigned(std::min<int64_t>((a.acc_total + 255) / 256, 4096));
    hipLaunchKernelGGL(eta_zero_kernel, dim3(std::max(zb, 1u)), dim3(256), 0, stream, a.acc,
                       a.acc_total);
    rocblas_status st = launch_status();
    if(st != rocblas_status_success)
        return st;
    if(a.nitems > 0)
    {
        hipLaunchKernelGGL(eta_items_kernel<T>, dim3(unsigned(a.nitems)), dim3(PNT), 0, stream,
                           a);
        if((st = launch_status()) != rocblas_status_success)
            return st;
    }
    hipLaunchKernelGGL(eta_final_kernel<T>, dim3(unsigned(a.batch)), dim3(256), 0, stream, a);
    return launch_status();
}

template rocblas_status eta_launch<float>(hipStream_t, const EtaArgs<float>&);
template rocblas_status eta_launch<double>(hipStream_t, const EtaArgs<double>&);

} // namespace rbamd
