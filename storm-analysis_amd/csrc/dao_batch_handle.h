// The daoBatch handle (include/dao_batch_amd.h), shared by the batch and the
// resident-fitter entry points.
#pragma once
#include <memory>
#include <vector>

#include "dao_engine.h"
#include "../../include/dao_batch_amd.h"

struct daoBatch {
  std::unique_ptr<dao::Engine> eng;
  std::vector<int> counters;  // [F][8], host totals not yet handed out
  double* d_tmp = nullptr;
  size_t tmp_bytes = 0;
};

// daoNewPeaks' peak types: PF_FINDER, PF_TESTING, PF_TEXT ("text" and
// "hdf5"), or -1 (after the reference's message) for anything else.
int dao_peak_format(const char* p_type);
