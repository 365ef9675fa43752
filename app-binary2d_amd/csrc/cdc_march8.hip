// cdc_march8.hip -- the column-marching stage kernel for strips of 8 rows
// (two strip waves per SIMD on grids with one 16-row strip per SIMD, such
// as 1024^2 in blocks of 64^2); a translation unit of its own, so that the
// strip heights compile in parallel.
#include "cdc_march.h"

namespace cdc {

hipError_t launch_march_rows8(int nv, int cool, int visc, int read_base, const StageArgs *args,
                              const int32_t *begin, int nruns, int ntiles, const StageArgs *fin,
                              const int32_t *fin_begin, int fin_nruns, int nblk, hipEvent_t ev0, hipEvent_t ev1,
                              hipStream_t s) {
    return launch_march_h<8>(nv, cool, visc, read_base, args, begin, nruns, ntiles, fin, fin_begin, fin_nruns, nblk,
                             ev0, ev1, s);
}

}  // namespace cdc
