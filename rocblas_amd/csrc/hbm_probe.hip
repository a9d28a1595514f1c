// hbm_probe.hip -- streaming kernel that measures the HBM rate a memory-bound
// kernel can reach for a given read/write mix on this device.
//
//   rocblas_amd_hbm_stream(handle, src, read_bytes, dst, write_bytes, iters, &seconds)
//
// The triangular solves of the strided-batched workload are bound by HBM
// bandwidth; their roofline fraction is quoted against the 8 TB/s peak, and
// this kernel gives the attainable rate for the same mix of reads and writes:
// every lane moves 16 bytes per access (dwordx4), each wavefront keeps
// UNROLL x (reads per step) loads in flight, and the reported time is the
// best over grids of 4, 8 and 16 wavefronts per SIMD with plain and with
// non-temporal loads and stores, so the rate is set by the memory system,
// not by latency or one cache policy.
//
// Layout of one step (G = threads in the grid): thread g reads
// src[(step*RQ + q)*G + g] for q < RQ and writes dst[(step*WQ + q)*G + g]
// for q < WQ, where (RQ, WQ) is the smallest integer ratio of the requested
// byte counts among the instantiated forms.  Written values depend on the
// values read, so no load is dead.
#include "internal.hpp"
#include "device_common.hpp"

#include "../../include/rocblas_amd_ext.h"

#include <algorithm>

namespace rbamd
{
namespace
{
using u32x4 = unsigned __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ u32x4 ld16(const u32x4* p)
{
    if constexpr(NT)
        return __builtin_nontemporal_load((const __attribute__((address_space(1))) u32x4*)(p));
    else
        return gload(p);
}

template <bool NT>
__device__ __forceinline__ void st16(u32x4* p, u32x4 v)
{
    if constexpr(NT)
        __builtin_nontemporal_store(v, (__attribute__((address_space(1))) u32x4*)(p));
    else
        gstore(p, v);
}

template <int RQ, int WQ, int UNROLL, bool NT>
__global__ void __launch_bounds__(256) hbm_stream_kernel(const u32x4* __restrict__ src,
                                                         u32x4* __restrict__ dst,
                                                         long long steps,
                                                         u32x4* __restrict__ sink)
{
    const long long G = (long long)gridDim.x * blockDim.x;
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    u32x4           acc = {0u, 0u, 0u, 0u};
    for(long long s0 = 0; s0 < steps; s0 += UNROLL)
    {
        u32x4 v[UNROLL][RQ > 0 ? RQ : 1];
#pragma unroll
        for(int u = 0; u < UNROLL; ++u)
        {
            const long long s = s0 + u < steps ? s0 + u : steps - 1;
#pragma unroll
            for(int q = 0; q < RQ; ++q)
                v[u][q] = ld16<NT>(src + ((s * RQ + q) * G + g));
        }
#pragma unroll
        for(int u = 0; u < UNROLL; ++u)
        {
            u32x4 w = acc;
#pragma unroll
            for(int q = 0; q < RQ; ++q)
                w ^= v[u][q];
            if(s0 + u < steps)
            {
                acc = w;
#pragma unroll
                for(int q = 0; q < WQ; ++q)
                {
                    u32x4 o = w;
                    o.x += unsigned(q);
                    st16<NT>(dst + (((s0 + u) * WQ + q) * G + g), o);
                }
            }
        }
    }
    // read-only form: keep every load live without writing per element
    if(WQ == 0 && (acc.x & acc.y & acc.z & acc.w) == 0xffffffffu)
        gstore(sink + g, acc);
}

// one launch shape: G threads, NT cache policy; average seconds of a pass
template <int RQ, int WQ, bool NT>
rocblas_status run_shape(rocblas_handle h, const u32x4* src, u32x4* dst, long long steps,
                         long long G, int iters, double* seconds)
{
    constexpr int UNROLL = (RQ + WQ) <= 2 ? 8 : 4;
    const dim3    grid(unsigned(G / 256));
    hipEvent_t    e0, e1;
    if(hipEventCreate(&e0) != hipSuccess)
        return rocblas_status_internal_error;
    if(hipEventCreate(&e1) != hipSuccess)
    {
        (void)hipEventDestroy(e0);
        return rocblas_status_internal_error;
    }
    // one untimed launch brings the code and the TLB entries in
    hipLaunchKernelGGL((hbm_stream_kernel<RQ, WQ, UNROLL, NT>), grid, dim3(256), 0, h->stream,
                       src, dst, steps, dst);
    rocblas_status st = launch_status();
    if(st == rocblas_status_success)
    {
        (void)hipEventRecord(e0, h->stream);
        for(int i = 0; i < iters; ++i)
            hipLaunchKernelGGL((hbm_stream_kernel<RQ, WQ, UNROLL, NT>), grid, dim3(256), 0,
                               h->stream, src, dst, steps, dst);
        st = launch_status();
        (void)hipEventRecord(e1, h->stream);
    }
    if(st == rocblas_status_success)
    {
        float ms = 0.f;
        if(hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            st = rocblas_status_internal_error;
        else
            *seconds = double(ms) * 1e-3 / iters;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

// the best over 4, 8 and 16 wavefronts per SIMD, plain and non-temporal;
// `unit` is the bytes of one access step at 16 wavefronts per SIMD
template <int RQ, int WQ>
rocblas_status run_form(rocblas_handle h, const u32x4* src, u32x4* dst, long long units,
                        int iters, double* seconds)
{
    const long long G16  = (long long)h->num_cus * 4 * 16 * WAVE;
    double          best = 0.0;
    for(int wps : {4, 8, 16})
    {
        const long long G     = G16 / 16 * wps;
        const long long steps = units * (16 / wps);
        for(int nt = 0; nt < 2; ++nt)
        {
            double         t  = 0.0;
            rocblas_status st = nt ? run_shape<RQ, WQ, true>(h, src, dst, steps, G, iters, &t)
                                   : run_shape<RQ, WQ, false>(h, src, dst, steps, G, iters, &t);
            if(st != rocblas_status_success)
                return st;
            if(best == 0.0 || t < best)
                best = t;
        }
    }
    *seconds = best;
    return rocblas_status_success;
}

} // namespace
} // namespace rbamd

using namespace rbamd;

extern "C" rocblas_status rocblas_amd_hbm_stream(rocblas_handle handle,
                                                 const void*    src,
                                                 size_t         read_bytes,
                                                 void*          dst,
                                                 size_t         write_bytes,
                                                 rocblas_int    iters,
                                                 double*        seconds)
try
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(!seconds || (read_bytes && !src) || !dst)
        return rocblas_status_invalid_pointer;
    if(iters < 1)
        return rocblas_status_invalid_size;
    // supported mixes (reads : writes) and the grid that streams them
    struct Form
    {
        int rq, wq;
    };
    static const Form forms[] = {{1, 0}, {0, 1}, {1, 1}, {2, 1}, {5, 2}, {3, 1}};
    int               f       = -1;
    for(int i = 0; i < int(sizeof(forms) / sizeof(forms[0])); ++i)
        if(read_bytes * size_t(forms[i].wq) == write_bytes * size_t(forms[i].rq))
            f = i;
    if(f < 0)
        return rocblas_status_invalid_size;
    const int       rq = forms[f].rq, wq = forms[f].wq;
    const long long per_unit = (long long)handle->num_cus * 4 * 16 * WAVE * 16; // 16 waves/SIMD
    const long long units
        = rq ? (long long)(read_bytes / (per_unit * rq)) : (long long)(write_bytes / (per_unit * wq));
    if(units < 1 || (rq && size_t(units * per_unit * rq) != read_bytes)
       || (wq && size_t(units * per_unit * wq) != write_bytes))
        return rocblas_status_invalid_size; // byte counts must be whole units
    DeviceGuard  dg(handle->device);
    const u32x4* s = static_cast<const u32x4*>(src);
    u32x4*       d = static_cast<u32x4*>(dst);
    switch(rq * 10 + wq)
    {
    case 10: return run_form<1, 0>(handle, s, d, units, iters, seconds);
    case 1: return run_form<0, 1>(handle, s, d, units, iters, seconds);
    case 11: return run_form<1, 1>(handle, s, d, units, iters, seconds);
    case 21: return run_form<2, 1>(handle, s, d, units, iters, seconds);
    case 52: return run_form<5, 2>(handle, s, d, units, iters, seconds);
    case 31: return run_form<3, 1>(handle, s, d, units, iters, seconds);
    default: return rocblas_status_invalid_size;
    }
}
catch(...)
{
    return rocblas_status_internal_error;
}

extern "C" rocblas_status rocblas_amd_hbm_stream_step_bytes(rocblas_handle handle, size_t* bytes)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(!bytes)
        return rocblas_status_invalid_pointer;
    *bytes = size_t(handle->num_cus) * 4 * 16 * WAVE * 16;
    return rocblas_status_success;
}
