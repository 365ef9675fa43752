// cdc_noalloc.hip -- the one entry point that allocates nothing:
// cdc_max_signal_speed_noalloc, a stateless variant of
// cdc_solver_max_signal_speed for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).
//
// The per-workgroup partial maxima live in `scratch`, a file-scope device
// array of kNoAllocParts doubles (8 KiB, held by every GPU the library runs
// on until the process ends).  Host code never passes `scratch` itself to a
// kernel (in host code the name is a host-side stand-in, not a device
// address): on a device's first call its address is looked up once with
// hipGetSymbolAddress, with that device current, and kept per device.
// Every partial a call's reduction reads is written by that call's first
// kernel, so nothing needs clearing between calls; calls that overlap on one
// GPU share the array and race.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "cdc_device.h"
#include "cdc_host.h"

namespace cdc {

namespace {

constexpr int kNoAllocParts = 1024;          // partial maxima (one per workgroup)
constexpr int kNoAllocThreads = 256;
constexpr int64_t kNoAllocMaxCells = (int64_t)1 << 40;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kNoAllocParts];

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};         // device address of `scratch`, per device

// Max signal speed of cells [c][v] (the boundary layout, any cell order),
// grid-stride; one partial per workgroup into part[blockIdx.x].
template <int NV>
__global__ void __launch_bounds__(kNoAllocThreads)
aos_max_speed_kernel(const double *__restrict__ u, int64_t ncell, double gm, double *__restrict__ part) {
    __shared__ double red[kNoAllocThreads];
    double m = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * kNoAllocThreads + threadIdx.x; c < ncell;
         c += (int64_t)gridDim.x * kNoAllocThreads) {
        double w[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) w[v] = u[c * NV + v];
        m = fmax(m, dev::max_signal_speed<NV>(w, gm));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kNoAllocThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// *out = max of part[0 .. n-1], one workgroup
__global__ void __launch_bounds__(kNoAllocThreads)
partial_max_kernel(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double red[kNoAllocThreads];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += kNoAllocThreads) m = fmax(m, part[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kNoAllocThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// The current device's address of `scratch`, looked up on its first call.
int scratch_for_current_device(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return host::set_msg(CDC_ERR_INVALID_ARGUMENT, "device index beyond the supported range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return host::set_msg(CDC_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return CDC_OK;
}

}  // namespace

}  // namespace cdc

extern "C" int32_t cdc_max_signal_speed_noalloc(int32_t kind, double gamma_law_index, const double *conserved,
                                                int64_t ncell, double *out, void *stream) {
    using namespace cdc;
    if (!conserved || !out) return host::set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    if (kind != CDC_ISOTHERMAL && kind != CDC_EULER)
        return host::set_msg(CDC_ERR_INVALID_ARGUMENT, "hydro kind must be CDC_ISOTHERMAL or CDC_EULER");
    if (kind == CDC_EULER && !(gamma_law_index > 1.0))
        return host::set_msg(CDC_ERR_INVALID_ARGUMENT, "gamma_law_index must exceed 1");
    if (ncell < 1 || ncell > kNoAllocMaxCells)
        return host::set_msg(CDC_ERR_INVALID_ARGUMENT, "ncell must be in [1, 2^40] (the no-allocation "
                                                       "entry point's limit)");
    double *part = nullptr;
    const int rc = scratch_for_current_device(&part);
    if (rc) return rc;
    const hipStream_t s = (hipStream_t)stream;
    const int64_t want = (ncell + kNoAllocThreads - 1) / kNoAllocThreads;
    const int nparts = (int)(want < kNoAllocParts ? want : kNoAllocParts);
    if (kind == CDC_ISOTHERMAL)
        hipLaunchKernelGGL(aos_max_speed_kernel<3>, dim3(nparts), dim3(kNoAllocThreads), 0, s, conserved, ncell, 1.0,
                           part);
    else
        hipLaunchKernelGGL(aos_max_speed_kernel<4>, dim3(nparts), dim3(kNoAllocThreads), 0, s, conserved, ncell,
                           gamma_law_index, part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(partial_max_kernel, dim3(1), dim3(kNoAllocThreads), 0, s, (const double *)part, nparts, out);
    HIP_TRY(hipGetLastError());
    return CDC_OK;
}
