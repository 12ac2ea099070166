// k_resident instantiations in FP64 (MLE) whose requests are fitted by teams
// of workgroups: 8 waves per workgroup, 4 for the widest ROIs.
#include "dao_resident_ops.h"

namespace dao {

void launch_resident_team_fp64(const FitArgs& a, const ResArgs& r, int blocks, hipStream_t st) {
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_resident_t<PPL, SHP, NW, false, EM_MLE, true>(a, r, blocks, st);
                 })));
}

}  // namespace dao
