// k_fit_frames instantiations in FP64 (the reference's arithmetic) for the
// error models other than MLE (read at run time); MLE has its own instances
// (dao_fit_fp64_mle.hip).  8 waves per workgroup, 4 for the widest ROIs,
// whose LDS slices are larger.
#include "dao_ops.h"

namespace dao {

void launch_fit_frames_fp64(const FitArgs& a, int nframes, hipStream_t st) {
  if (a.m.errm == EM_MLE) {
    launch_fit_frames_fp64_mle(a, nframes, st);
    return;
  }
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_fit_frames_t<PPL, SHP, NW, false>(a, nframes, st);
                 })));
}

}  // namespace dao
