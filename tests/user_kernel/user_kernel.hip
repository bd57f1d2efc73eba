// A user HIP kernel that walks neighbours through the device-side header
// (include/nlhip_device.h) without materialising a pair list -- the
// counterpart of the reference's "SortedCellList in user kernels" test
// (test/test_gpu.jl:94-122), which counts neighbours per atom inside a
// custom kernel.  Besides the count this kernel sums R and S per atom, so
// the test also checks the vectors the callback receives.
//
// Built by the package build into libnlh_user_kernel.so next to this file;
// only the GPU tests load it.
#include <hip/hip_runtime.h>

#include "nlhip.h"
#include "nlhip_device.h"

template <typename TI>
__global__ void count_and_sum(nlh_celllist_dev<TI> c, int32_t *count, double *Rsum, int64_t *Ssum)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;   // 1-based
    if (i > c.nat) return;
    int32_t n = 0;
    double r[3] = {0.0, 0.0, 0.0};
    int64_t s[3] = {0, 0, 0};
    nlh_for_each_neighbour(c, i, [&](TI j, const double *R, const TI *S) {
        (void)j;
        ++n;
        for (int k = 0; k < 3; ++k) {
            r[k] += R[k];
            s[k] += (int64_t)S[k];
        }
    });
    count[i - 1] = n;
    for (int k = 0; k < 3; ++k) {
        Rsum[3 * (i - 1) + k] = r[k];
        Ssum[3 * (i - 1) + k] = s[k];
    }
}

extern "C" int user_count_neighbours(const nlh_celllist_view *v, const nlh_geometry *g, int32_t *count,
                                     double *Rsum, int64_t *Ssum)
{
    if (!v || !g || !count || !Rsum || !Ssum) return NLH_EINVAL;
    const int64_t n = v->nat;
    if (n == 0) return NLH_OK;
    int32_t *dc = nullptr;
    double *dr = nullptr;
    int64_t *ds = nullptr;
    if (hipMalloc(&dc, n * 4) != hipSuccess || hipMalloc(&dr, n * 24) != hipSuccess ||
        hipMalloc(&ds, n * 24) != hipSuccess) {
        (void)hipFree(dc);
        (void)hipFree(dr);
        return NLH_ENOMEM;
    }
    const unsigned blocks = (unsigned)((n + 127) / 128);
    if (v->ti_bytes == 8)
        hipLaunchKernelGGL(count_and_sum<int64_t>, dim3(blocks), dim3(128), 0, 0,
                           nlh_make_celllist_dev<int64_t>(*v, *g), dc, dr, ds);
    else
        hipLaunchKernelGGL(count_and_sum<int32_t>, dim3(blocks), dim3(128), 0, 0,
                           nlh_make_celllist_dev<int32_t>(*v, *g), dc, dr, ds);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(count, dc, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(Rsum, dr, n * 24, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(Ssum, ds, n * 24, hipMemcpyDeviceToHost);
    (void)hipFree(dc);
    (void)hipFree(dr);
    (void)hipFree(ds);
    return e == hipSuccess ? NLH_OK : NLH_EDEVICE;
}
