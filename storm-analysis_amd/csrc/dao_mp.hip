// Multiplane DAO fitter on the device: mpIterateLM
// (storm_analysis/multi_plane/mp_fit.c:180-487) with the 2D Gaussian
// channels of mp_fit_dao.c (mpDaoUpdate :173-209, mpUpdate mp_fit.c:604-690).
//
// A peak is a group of copies, one per channel (image plane).  Its turn sums
// the channels' MLE deviances, solves each channel's damped normal equations,
// and moves the group together: heights, widths and backgrounds per channel,
// x and y as the height- and z-weighted mean of the channels' steps mapped to
// channel 0, then mapped out to the other channels by their affine
// transforms.  One wavefront fits one frame-set (its channels are frames of
// the engine) and runs its groups strictly in the reference's index order, so
// no scheduling is needed and every group sees every earlier group's update
// in every channel; a launch runs many frame-sets side by side (frames are
// independent).  The per-channel pixel work reuses the single-channel ROI
// cache, error, Jacobian / Hessian, Cholesky and recentering code; the
// wave's LDS slice holds the ROI cache and, after it, the group's state
// (MpWave).  FP64 throughout, as the reference.
#include "dao_ops.h"

namespace dao {

namespace {

constexpr int kMpN = 5;                       // 2D: h, x, y, width, bg (dao_fit.c:409-467)
constexpr int kMpH = kMpN * (kMpN + 1) / 2;   // packed upper triangle

// The group's state, every lane holding (and writing) the same values.
struct MpWave {
  PeakState st[kMpMaxChannels];  // working peaks
  double J[kMpMaxChannels][kMpN];
  double H[kMpMaxChannels][kMpH];
  double dl[kMpMaxChannels][kMpN];  // solved steps
  double hts[kMpMaxChannels];       // heights weighting the x / y steps
  int cnt[kMpMaxChannels][C_COUNT];
};

template <int PPL>
__host__ __device__ constexpr int mp_slice_doubles() {
  return RoiLds<PPL>::kDoubles + (int)((sizeof(MpWave) + 7) / 8);
}

struct Cpl {
  const double *xt0N, *yt0N, *xtN0, *ytN0, *wmin, *wmax, *wx, *wy;
};

__device__ inline Cpl coupling(const MpArgs& a) {
  const double* t = a.cpl;
  const int c = a.nch;
  Cpl k;
  k.xt0N = t;
  k.yt0N = t + 3 * c;
  k.xtN0 = t + 6 * c;
  k.ytN0 = t + 9 * c;
  k.wmin = t + 12 * c;
  k.wmax = t + 13 * c;
  k.wx = t + 14 * c;
  k.wy = k.wx + a.n_weights * c;
  return k;
}

// The channel's model: the shared one with the channel's width range.
__device__ inline Model channel_model(const MpArgs& a, const Cpl& k, int c) {
  Model m = a.m;
  m.wmin = k.wmin[c];
  m.wmax = k.wmax[c];
  return m;
}

__device__ inline void add_counts(MpWave& s, int c, const int* d) {
#pragma unroll
  for (int q = 0; q < C_COUNT; q++) s.cnt[c][q] += d[q];
}

// mFitAddPeak / mFitSubtractPeak of state w in frame fr.
template <int PPL>
__device__ void mp_stamp(const MpArgs& a, const Model& m, const Frame& fr, double* slice, const PeakState& w,
                         int sign) {
  Roi<PPL> r;
  r.init(m, a.roi_tab, fr, slice);
  r.load(a.px, fr, w.xi, w.yi);
  r.shape(m, w);
  r.stamp(w, sign);
  r.store(a.px);
}

// mFitCalcErr of state w in frame fr: 0 and the error, or 1 (fi <= 0,
// n_neg_fi counted); the pixels' cached values are refreshed either way.
template <int PPL>
__device__ int mp_error(const MpArgs& a, const Model& m, const Frame& fr, double* slice, const PeakState& w,
                        double& err, int* d) {
  Roi<PPL> r;
  r.init(m, a.roi_tab, fr, slice);
  r.load(a.px, fr, w.xi, w.yi);
  const int rc = eval_error<PPL>(m, r, a.px.asx, err, d[C_NEG_FI]);
  r.store(a.px);
  return rc;
}

// mpWeightIndex / mpWeightInterpolate (mp_fit.c:697-741).
__device__ inline void weight_index(const MpArgs& a, double z, int& zi, double& dz) {
  const double s = a.w_z_scale * (z - a.w_z_offset);
  zi = (int)s;
  dz = s - (double)zi;
  if (zi < 0) zi = 0;
  if (zi > a.n_weights - 2) zi = a.n_weights - 2;
}

__device__ inline double weight_at(const double* w, double dz, int zi, int nch, int c) {
#pragma clang fp contract(off)
  const double w1 = w[nch * zi + c];
  const double w2 = w[nch * (zi + 1) + c];
  return w1 + (w2 - w1) * dz;
}

// mpDaoUpdate + mpUpdate: the channels' steps s.dl applied to s.st.
__device__ void mp_update(const MpArgs& a, const Cpl& k, MpWave& s, int f0) {
#pragma clang fp contract(off)
  const int nch = a.nch;
  // heights float per channel, never below 0.01 (mp_fit_dao.c:180-191)
  for (int c = 0; c < nch; c++) {
    double h = s.st[c].prm[P_HEIGHT] - s.dl[c][0];
    if (h < 0.01) h = 0.01;
    s.st[c].prm[P_HEIGHT] = h;
    s.hts[c] = h;
  }
  // widths (XWIDTH only) per channel, clamped by daoCheck2D at the old ROI
  // origin (mp_fit_dao.c:193-204)
  for (int c = 0; c < nch; c++) {
    PeakState w = s.st[c];
    w.prm[P_XW] -= s.dl[c][3];
    const Model m = channel_model(a, k, c);
    const Frame fr = a.frames[f0 + c];
    int d[C_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
    (void)peak_check<SH_2D>(m, fr, w, d);
    add_counts(s, c, d);
    s.st[c] = w;
  }
  wave_sync_lds();
  // x and y: weighted mean of the channels' steps mapped to channel 0
  int zi;
  double dz;
  weight_index(a, s.st[0].prm[P_ZC], zi, dz);
  for (int axis = 0; axis < 2; axis++) {
    const double* tr = axis == 0 ? k.xtN0 : k.ytN0;
    const double* wt = axis == 0 ? k.wx : k.wy;
    double sum = 0.0, norm = 0.0;
    for (int c = 0; c < nch; c++) {
      const double w = weight_at(wt, dz, zi, nch, c);
      double dd = tr[3 * c + 1] * s.dl[c][1];
      dd += tr[3 * c + 2] * s.dl[c][2];
      sum += dd * w * s.hts[c];
      norm += w * s.hts[c];
    }
    s.st[0].prm[axis == 0 ? P_XC : P_YC] -= sum / norm;
  }
  wave_sync_lds();
  // the other channels follow channel 0 through their transforms
  const double xoff = a.m.xoff, x0 = s.st[0].prm[P_XC], y0 = s.st[0].prm[P_YC];
  for (int c = 1; c < nch; c++) {
    double t = k.xt0N[3 * c];
    t += k.xt0N[3 * c + 1] * (x0 + xoff);
    t += k.xt0N[3 * c + 2] * (y0 + xoff);
    s.st[c].prm[P_XC] = t - xoff;
    t = k.yt0N[3 * c];
    t += k.yt0N[3 * c + 1] * (x0 + xoff);
    t += k.yt0N[3 * c + 2] * (y0 + xoff);
    s.st[c].prm[P_YC] = t - xoff;
  }
  for (int c = 0; c < nch; c++) {
    PeakState w = s.st[c];
    roi_follow(w);
    w.prm[P_BG] -= s.dl[c][4];
    s.st[c] = w;
  }
  wave_sync_lds();
}

// mpResetWorkingPeaks (mp_fit.c:496-516): every channel back to its record,
// keeping `added`, lambda x 4, status ERROR.
__device__ void mp_reset(const MpArgs& a, MpWave& s, int f0, int i) {
  for (int c = 0; c < a.nch; c++) {
    PeakState w = s.st[c];
    reset_from(w, a.pk.rec + a.frames[f0 + c].peak0 + i);
    w.status = ST_ERROR;
    s.st[c] = w;
  }
  wave_sync_lds();
}

__device__ inline void mp_all_status(MpWave& s, int nch, int st) {
  for (int c = 0; c < nch; c++) s.st[c].status = st;
  wave_sync_lds();
}

// The body of mpIterateLM's loop for group i (mp_fit.c:195-456).
template <int PPL>
__device__ void mp_turn(const MpArgs& a, const Cpl& k, MpWave& s, double* slice, int f0, int i) {
  const int nch = a.nch;
  double start = 0.0;
  for (int c = 0; c < nch; c++) {
    const Frame fr = a.frames[f0 + c];
    const Model m = channel_model(a, k, c);
    PeakState w;
    load_state(a.pk.rec + fr.peak0 + i, w);
    Roi<PPL> r;
    r.init(m, a.roi_tab, fr, slice);
    r.load(a.px, fr, w.xi, w.yi);
    r.shape(m, w);
    double e;
    int d[C_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (eval_error<PPL>(m, r, a.px.asx, e, d[C_NEG_FI]) == 0) w.err = e;  // return value ignored by the reference
    start += w.err;
    calc_jh<PPL, SH_2D>(m, r, w, a.px.asx);
    const double* hj = r.hj();
#pragma unroll
    for (int q = 0; q < kMpN; q++) s.J[c][q] = hj[q];
#pragma unroll
    for (int q = 0; q < kMpH; q++) s.H[c][q] = hj[kMpN + q];
    r.stamp(w, -1);
    w.added--;
    r.store(a.px);
    add_counts(s, c, d);
    s.st[c] = w;
    wave_sync_lds();
  }

  // Every retry multiplies lambda by 4, so more than ~60 cycles cannot
  // happen; the guard only makes termination evident.
  for (int guard = 0;; guard++) {
    if (s.st[0].lam > 1.0e20 || guard > 200) {
      for (int c = 0; c < nch; c++) {
        s.cnt[c][C_LOST]++;
        s.st[c].status = ST_ERROR;
      }
      wave_sync_lds();
      break;
    }
    mp_all_status(s, nch, ST_RUNNING);

    // per-channel solves; a failure stops the loop (later channels' cycle
    // counts are not incremented, mp_fit.c:243-269)
    bool bad = false;
    for (int c = 0; c < nch; c++) {
      s.st[c].iter++;
      s.cnt[c][C_ITER]++;
      double b[kMpN];
#pragma unroll
      for (int q = 0; q < kMpN; q++) b[q] = s.J[c][q];
      if (chol_solve<kMpN>(s.H[c], s.st[c].lam, b) != 0) {
        s.cnt[c][C_DPOSV]++;
        bad = true;
        break;
      }
#pragma unroll
      for (int q = 0; q < kMpN; q++) s.dl[c][q] = b[q];
    }
    wave_sync_lds();
    if (bad) {
      for (int c = 0; c < nch; c++) {
        s.st[c].status = ST_ERROR;
        s.st[c].lam = s.st[c].lam * 4.0;
      }
      wave_sync_lds();
      continue;
    }

    mp_update(a, k, s, f0);
    for (int c = 0; c < nch; c++) {
      PeakState w = s.st[c];
      int d[C_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (peak_check<SH_2D>(channel_model(a, k, c), a.frames[f0 + c], w, d)) bad = true;
      add_counts(s, c, d);
      s.st[c] = w;
    }
    wave_sync_lds();
    if (bad) {
      mp_reset(a, s, f0, i);
      continue;
    }

    for (int c = 0; c < nch; c++) {
      mp_stamp<PPL>(a, channel_model(a, k, c), a.frames[f0 + c], slice, s.st[c], +1);
      s.st[c].added++;
    }
    double cur = 0.0;
    for (int c = 0; c < nch; c++) {
      double e;
      int d[C_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (mp_error<PPL>(a, channel_model(a, k, c), a.frames[f0 + c], slice, s.st[c], e, d) != 0) {
        bad = true;
      } else {
        s.st[c].err = e;
        cur += e;
      }
      add_counts(s, c, d);
    }
    wave_sync_lds();
    if (bad) {
      for (int c = 0; c < nch; c++) {
        mp_stamp<PPL>(a, channel_model(a, k, c), a.frames[f0 + c], slice, s.st[c], -1);
        s.st[c].added--;
      }
      mp_reset(a, s, f0, i);
      continue;
    }

    if (cur > start) {
      if (((cur - start) / start) < a.tol) {
        mp_all_status(s, nch, ST_CONVERGED);
        break;
      }
      for (int c = 0; c < nch; c++) {
        s.cnt[c][C_NON_DECR]++;
        mp_stamp<PPL>(a, channel_model(a, k, c), a.frames[f0 + c], slice, s.st[c], -1);
        s.st[c].added--;
      }
      mp_reset(a, s, f0, i);
      continue;
    }
    if (((start - cur) / start) < a.tol) {
      mp_all_status(s, nch, ST_CONVERGED);
    } else if (s.st[0].lam > 1.0e-3) {
      for (int c = 0; c < nch; c++) s.st[c].lam = 0.75 * s.st[c].lam;
      wave_sync_lds();
    }
    break;
  }
  // mpCopyFromWorking (mp_fit.c:121-133): channel 0's status for the group
  const int status = s.st[0].status;
  for (int c = 0; c < nch; c++) {
    PeakState w = s.st[c];
    w.status = status;
    store_state(a.pk.rec + a.frames[f0 + c].peak0 + i, w);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}

// mpCheckError (mp_fit.c:44-67): a group in ERROR in some channel leaves the
// fit in every channel.
template <int PPL>
__device__ void mp_check_group(const MpArgs& a, const Cpl& k, double* slice, int f0, int i) {
  bool bad = false;
  for (int c = 0; c < a.nch; c++) bad |= a.pk.rec[a.frames[f0 + c].peak0 + i].status == ST_ERROR;
  if (!uniform_i(bad ? 1 : 0)) return;
  for (int c = 0; c < a.nch; c++) {
    const Frame fr = a.frames[f0 + c];
    PeakRec* rec = a.pk.rec + fr.peak0 + i;
    PeakState w;
    load_state(rec, w);
    if (w.status != ST_ERROR) {
      const Model m = channel_model(a, k, c);
      if (origin_ok(m, fr, w.xi, w.yi)) mp_stamp<PPL>(a, m, fr, slice, w, -1);
      w.added--;
    }
    w.status = ST_ERROR;
    store_state(rec, w);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  }
}

// Whole doFit (or one sweep when max_more == 0) of one frame-set per wave.
template <int PPL, int NW>
__global__ __launch_bounds__(NW * 64) void k_fit_mp(MpArgs a) {
  extern __shared__ double lds_dyn[];
  constexpr int kSlice = mp_slice_doubles<PPL>();
  const int wv = threadIdx.x >> 6;
  const int set = blockIdx.x * NW + wv;
  if (set >= a.nsets) return;  // the whole wave: waves never wait on each other
  double* const slice = lds_dyn + wv * kSlice;
  MpWave& s = *reinterpret_cast<MpWave*>(slice + RoiLds<PPL>::kDoubles);
  const Cpl k = coupling(a);
  const int f0 = a.set_frame0[set];
  for (int c = 0; c < a.nch; c++)
    for (int q = 0; q < C_COUNT; q++) s.cnt[c][q] = 0;
  wave_sync_lds();
  const int nfit = a.frames[f0].nfit;
  const int lane = lane_id();
  int sweeps = 0;
  for (int it = 0;; it++) {
    // RUNNING groups (channel 0 decides, as MultiFitter.doFit's getUnconverged)
    int nrun = 0;
    for (int i0 = 0; i0 < nfit; i0 += kWave) {
      const int i = i0 + lane;
      const bool r = i < nfit && a.pk.rec[a.frames[f0].peak0 + i].status == ST_RUNNING;
      nrun += __popcll(__ballot(r));
    }
    if (it > 0 && (nrun == 0 || it > a.max_more)) break;
    sweeps++;
    for (int i = 0; i < nfit; i++) {
      if (uniform_i(a.pk.rec[a.frames[f0].peak0 + i].status) != ST_RUNNING) continue;
      mp_turn<PPL>(a, k, s, slice, f0, i);
    }
    // mFitRecenterPeaks per channel (multi_fit.c:1644-1690), then the
    // groups' ERROR states synchronised
    for (int c = 0; c < a.nch; c++) {
      const Frame fr = a.frames[f0 + c];
      const Model m = channel_model(a, k, c);
      for (int i = 0; i < nfit; i++) {
        const PeakRec* rec = a.pk.rec + fr.peak0 + i;
        const int status = uniform_i(rec->status);
        if (status != ST_RUNNING) continue;
        const double dx = rec->prm[P_XC] - rec->xi, dy = rec->prm[P_YC] - rec->yi;
        const bool move = (dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25);
        if (!uniform_i(move ? 1 : 0)) continue;
        int d[C_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
        recenter_turn<PPL>(m, a.px, a.pk, fr, a.roi_tab, slice, fr.peak0 + i, d);
        add_counts(s, c, d);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      }
    }
    for (int i = 0; i < nfit; i++) mp_check_group<PPL>(a, k, slice, f0, i);
    wave_sync_lds();
  }
  if (lane < C_COUNT)
    for (int c = 0; c < a.nch; c++)
      if (s.cnt[c][lane] != 0) atomicAdd(&a.frame_counters[(f0 + c) * C_COUNT + lane], s.cnt[c][lane]);
  if (lane == 0) a.frame_sweeps[f0] = sweeps;
}

template <int PPL, int NW>
void launch_mp_t(const MpArgs& a, hipStream_t st) {
  const size_t bytes = sizeof(double) * (size_t)NW * mp_slice_doubles<PPL>();
  static size_t allowed = 64 * 1024;  // larger dynamic LDS must be enabled per kernel
  if (bytes > allowed) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fit_mp<PPL, NW>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      return;  // the launch below then fails; the caller's error check reports it
    allowed = bytes;
  }
  k_fit_mp<PPL, NW><<<(a.nsets + NW - 1) / NW, NW * 64, bytes, st>>>(a);
}

}  // namespace

void launch_fit_mp(const MpArgs& a, hipStream_t st) {
  if (a.nsets <= 0) return;
  DAO_PPL_SWITCH(a.m.nroi, ({
                   constexpr int NW = PPL <= 8 ? 4 : 2;
                   launch_mp_t<PPL, NW>(a, st);
                 }));
}

}  // namespace dao
