// k_fit_team instantiations in FP64 (every error model), the arithmetic of
// dao_fit_fp64.hip; same wave counts.
#include "dao_ops.h"

namespace dao {

void launch_fit_team_fp64(const FitArgs& a, int nblocks, hipStream_t st) {
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_fit_team_t<PPL, SHP, NW, false>(a, nblocks, st);
                 })));
}

}  // namespace dao
