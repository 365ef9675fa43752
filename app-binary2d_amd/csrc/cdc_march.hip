// cdc_march.hip -- the column-marching stage kernel for strips of 32 rows,
// its launch for every strip height, and the strip-edge array builder.
#include "cdc_march.h"

namespace cdc {

// The strip-edge columns of a state (see StageArgs::ec), one thread per entry.
__global__ void __launch_bounds__(256) edge_build_kernel(const double *__restrict__ u, double *__restrict__ E, int N,
                                                         int nv) {
    const int64_t total = (int64_t)nv * N * (N / 64) * 4;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < total) E[e] = edge_entry(u, e, N, nv);
}

hipError_t launch_edge_build(int nv, int N, const double *u, double *E, hipStream_t s) {
    const int64_t total = (int64_t)nv * N * (N / 64) * 4;
    hipLaunchKernelGGL(edge_build_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, u, E, N, nv);
    return hipGetLastError();
}

// StageArgs::stat, one thread per cell
__global__ void __launch_bounds__(256) stat_build_kernel(const double *__restrict__ u0, const double *__restrict__ aux,
                                                         double *__restrict__ stat, int64_t nn, int nv, int npair) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= nn) return;
    stat[2 * c] = u0[sidx(nv, 0, nn, c)];
    stat[2 * c + 1] = u0[sidx(nv, 1, nn, c)];
    stat[2 * nn + 2 * c] = u0[sidx(nv, 2, nn, c)];
    stat[2 * nn + 2 * c + 1] = nv == 4 ? u0[sidx(nv, 3, nn, c)] : aux[c];
    if (npair == 3) {
        stat[4 * nn + 2 * c] = aux[c];
        stat[4 * nn + 2 * c + 1] = 0.0;
    }
}

hipError_t launch_stat_build(int nv, int npair, const double *u0, const double *aux, double *stat, int64_t nn,
                             hipStream_t s) {
    hipLaunchKernelGGL(stat_build_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, u0, aux, stat, nn,
                       nv, npair);
    return hipGetLastError();
}

hipError_t launch_march_batch(int nv, int rows, int cool, int visc, int read_base, const StageArgs *args,
                              const int32_t *begin, int nruns, int ntiles, const StageArgs *fin,
                              const int32_t *fin_begin, int fin_nruns, int nblk, hipEvent_t ev0, hipEvent_t ev1,
                              hipStream_t s) {
    if (rows == 32)
        return launch_march_h<32>(nv, cool, visc, read_base, args, begin, nruns, ntiles, fin, fin_begin, fin_nruns,
                                  nblk, ev0, ev1, s);
    if (rows == 16)
        return launch_march_rows16(nv, cool, visc, read_base, args, begin, nruns, ntiles, fin, fin_begin, fin_nruns,
                                   nblk, ev0, ev1, s);
    if (rows == 8)
        return launch_march_rows8(nv, cool, visc, read_base, args, begin, nruns, ntiles, fin, fin_begin, fin_nruns,
                                  nblk, ev0, ev1, s);
    return hipErrorInvalidValue;
}

}  // namespace cdc
