// k_resident instantiations in mixed precision (MLE): 8 waves per
// workgroup, 4 for the widest ROIs.
#include "dao_resident_ops.h"

namespace dao {

void launch_resident_mixed(const FitArgs& a, const ResArgs& r, int blocks, hipStream_t st) {
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_resident_t<PPL, SHP, NW, true, EM_MLE, false>(a, r, blocks, st);
                 })));
}

}  // namespace dao
