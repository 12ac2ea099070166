// k_fit_frames instantiations in FP64 with the MLE error model fixed at
// compile time: the turn's deviance and Jacobian carry no branches or
// registers for the other error models.
#include "dao_ops.h"

namespace dao {

void launch_fit_frames_fp64_mle(const FitArgs& a, int nframes, hipStream_t st) {
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_fit_frames_t<PPL, SHP, NW, false, EM_MLE>(a, nframes, st);
                 })));
}

}  // namespace dao
