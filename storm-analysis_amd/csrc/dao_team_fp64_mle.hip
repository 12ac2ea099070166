// k_fit_team instantiations in FP64 with the MLE error model fixed at compile
// time (as dao_fit_fp64_mle.hip).
#include "dao_ops.h"

namespace dao {

void launch_fit_team_fp64_mle(const FitArgs& a, int nblocks, hipStream_t st) {
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_fit_team_t<PPL, SHP, NW, false, EM_MLE>(a, nblocks, st);
                 })));
}

}  // namespace dao
