// k_fit_team instantiations in FP64, the arithmetic of dao_fit_fp64.hip, for
// the error models other than MLE; MLE has its own instances
// (dao_team_fp64_mle.hip).  Same wave counts.
#include "dao_ops.h"

namespace dao {

void launch_fit_team_fp64(const FitArgs& a, int nblocks, hipStream_t st) {
  if (a.m.errm == EM_MLE) {
    launch_fit_team_fp64_mle(a, nblocks, st);
    return;
  }
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_fit_team_t<PPL, SHP, NW, false>(a, nblocks, st);
                 })));
}

}  // namespace dao
