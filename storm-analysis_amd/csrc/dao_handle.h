// The fitData handle of the C-ABI: the reference's leading layout (read and
// written by the Python binding), then the device state it views -- frame
// `frame` of an Engine.  A single-frame handle (daoInitialize) owns its
// engine; the channels of a multiplane fitter are handles viewing the frames
// of one engine that the multiplane handle owns.
#pragma once
#include <memory>
#include <vector>

#include "dao_engine.h"
#include "../../include/dao_fit_amd.h"

namespace dao {

struct FitHandle {
  fitData pub;  // must stay first
  std::unique_ptr<Engine> own;
  Engine* eng = nullptr;
  int frame = 0;
  Model model;
  std::vector<int> roi_x, roi_y;
};

inline FitHandle* handle_of(fitData* fd) { return reinterpret_cast<FitHandle*>(fd); }

// Adds frame f's counter increments (rows of C_COUNT in `c`, as
// Engine::read_counters returns them) to the handle's fitData counters.
inline void add_frame_counters(fitData& p, const std::vector<int>& c, int f) {
  const int* q = c.data() + (size_t)f * C_COUNT;
  p.n_dposv += q[C_DPOSV];
  p.n_iterations += q[C_ITER];
  p.n_lost += q[C_LOST];
  p.n_margin += q[C_MARGIN];
  p.n_neg_fi += q[C_NEG_FI];
  p.n_neg_height += q[C_NEG_HEIGHT];
  p.n_non_converged += q[C_NON_CONV];
  p.n_non_decr += q[C_NON_DECR];
}

// daoNewPeaks on the handle's frame (throws dao::Error).
void handle_new_peaks(FitHandle* h, const double* peak_params, const char* p_type, int n_peaks);

}  // namespace dao
