// cdc_march16.hip -- the column-marching stage kernel for strips of 16 rows
// (a translation unit of its own, so that the two strip heights compile in
// parallel).
#include "cdc_march.h"

namespace cdc {

hipError_t launch_march_rows16(int nv, int cool, int visc, int read_base, const StageArgs *args,
                               const int32_t *begin, int nruns, int ntiles, const StageArgs *fin,
                               const int32_t *fin_begin, int fin_nruns, int nblk, hipEvent_t ev0, hipEvent_t ev1,
                               hipStream_t s) {
    return launch_march_h<16>(nv, cool, visc, read_base, args, begin, nruns, ntiles, fin, fin_begin, fin_nruns, nblk,
                              ev0, ev1, s);
}

}  // namespace cdc
