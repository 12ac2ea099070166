// k_fit_frames instantiations in mixed precision (MLE): FP32 deviance terms,
// Jacobian and Hessian; FP64 shape, accumulators, solve and accept/reject.
// Same wave counts as the FP64 fitter.
#include "dao_ops.h"

namespace dao {

void launch_fit_frames_mixed(const FitArgs& a, int nframes, hipStream_t st) {
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_fit_frames_t<PPL, SHP, NW, true>(a, nframes, st);
                 })));
}

}  // namespace dao
