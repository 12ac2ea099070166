// k_fit_team instantiations in mixed precision (MLE), the arithmetic of
// dao_fit_mixed.hip; same wave counts.
#include "dao_ops.h"

namespace dao {

void launch_fit_team_mixed(const FitArgs& a, int nblocks, hipStream_t st) {
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_fit_team_t<PPL, SHP, NW, true>(a, nblocks, st);
                 })));
}

}  // namespace dao
