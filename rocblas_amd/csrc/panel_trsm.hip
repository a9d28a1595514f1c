// panel_trsm.hip -- single-launch solve for 64 < k <= 512: one workgroup per
// panel of 16 right-hand-side vectors of one instance (panel_trsm.hpp), the
// panel's solved rows kept in LDS.  Panels are independent, so there is no
// inter-workgroup communication and a whole k <= 512 solve is one launch.
#include "panel_trsm.hpp"

namespace rbamd
{
namespace
{
template <typename T>
__global__ void __launch_bounds__(panel::PNT) __attribute__((amdgpu_waves_per_eu(2)))
    panel_trsm_kernel(const SmallArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char panel_smem[];
    const T   alpha = a.alpha_dev ? *a.alpha_dev : a.alpha_host;
    const T   scale = (a.flags & SF_SCALE) ? alpha : T(1);
    const int v0    = blockIdx.x * panel::PNC;
    for(int64_t b = blockIdx.y; b < a.units; b += gridDim.y)
    {
        T* Bp = a.B.get(b);
        if(alpha == T(0))
            panel::panel_zero<T>(Bp, a.K, a.nvec, a.ldb, a.flags, v0);
        else
            panel::panel_solve<T>(a.A.get(b), Bp, a.K, a.nvec, a.lda, a.ldb, a.flags, scale, v0,
                                  panel_smem);
    }
}
} // namespace

template <typename T>
bool panel_fits(rocblas_handle h, int K)
{
    return K <= kPanelMaxK && panel::lds_bytes<T>(K) <= h->max_lds;
}

template <typename T>
rocblas_status launch_panel(rocblas_handle h, const SmallArgs<T>& a, int batch)
{
    if(!panel_fits<T>(h, a.K))
        return rocblas_status_internal_error;
    SmallArgs<T> s = a;
    s.units        = batch;
    const unsigned panels = unsigned((a.nvec + panel::PNC - 1) / panel::PNC);
    const dim3     grid(panels, grid_batch(batch));
    // dynamic LDS beyond 64 KiB is opted into once per process
    static const hipError_t opt = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&panel_trsm_kernel<T>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(panel::lds_bytes<T>(kPanelMaxK)));
    (void)opt;
    (void)hipGetLastError();
    hipLaunchKernelGGL(panel_trsm_kernel<T>, grid, dim3(panel::PNT), panel::lds_bytes<T>(a.K),
                       h->stream, s);
    return launch_status();
}

template rocblas_status launch_panel<float>(rocblas_handle, const SmallArgs<float>&, int);
template rocblas_status launch_panel<double>(rocblas_handle, const SmallArgs<double>&, int);
template bool panel_fits<float>(rocblas_handle, int);
template bool panel_fits<double>(rocblas_handle, int);

} // namespace rbamd
