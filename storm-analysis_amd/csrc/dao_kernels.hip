// HIP kernels of the MI355X 3D-DAOSTORM fitter (gfx950).
//
// What the reference does sequentially (storm_analysis/sa_library/multi_fit.c
// mFitIterateLM:1072-1349 visits peaks in index order and each peak
// read-modify-writes the fit image over its ROI) is run here by an ordered
// executor: every image operation ("op") reserves a box of pixels; op j waits
// for every earlier op i < j of its frame whose box overlaps, and all other ops
// run concurrently.  Each op is one 64-lane wavefront; the ROI's pixels are
// spread over the lanes (pixel j of the ROI -> lane j % 64, slot j / 64) and
// kept in registers for the whole turn, so a Levenberg-Marquardt turn touches
// HBM once on entry and once on exit.  Cross-lane sums use wavefront
// shuffles.  The 4..6-parameter damped normal equations are solved in FP64
// registers by every lane (uniform control flow).
//
// An LM turn may move its ROI (multi_fit.c:1933-1946).  If a candidate ROI
// leaves the reserved box the op puts its frame back as it found it, flags the
// frame, and the host re-runs that frame's sweep with a larger box for that
// peak (rollback of the frame's completed ops: k_rollback).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "dao_device.h"

namespace dao {

// ---------------------------------------------------------------------------
// wavefront helpers
// ---------------------------------------------------------------------------

__device__ inline double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ inline int wmin_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int u = __shfl_xor(v, o, kWave);
    v = u < v ? u : v;
  }
  return v;
}

__device__ inline int wsum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ inline int lane_id() { return threadIdx.x & (kWave - 1); }

// ---------------------------------------------------------------------------
// per-peak uniform state (every lane holds the same copy)
// ---------------------------------------------------------------------------

struct PeakState {
  double prm[kNFit];
  double lam;
  double err;
  int xi, yi;
  int status;
  int iter;
  int added;
};

__device__ inline void load_state(const Peaks& pk, int p, PeakState& s) {
#pragma unroll
  for (int q = 0; q < kNFit; q++) s.prm[q] = pk.prm[(long long)q * pk.cap + p];
  s.lam = pk.lam[p];
  s.err = pk.err[p];
  s.xi = pk.xi[p];
  s.yi = pk.yi[p];
  s.status = pk.status[p];
  s.iter = pk.iter[p];
  s.added = pk.added[p];
}

__device__ inline void store_state(const Peaks& pk, int p, const PeakState& s) {
  if (lane_id() == 0) {
#pragma unroll
    for (int q = 0; q < kNFit; q++) pk.prm[(long long)q * pk.cap + p] = s.prm[q];
    pk.lam[p] = s.lam;
    pk.err[p] = s.err;
    pk.xi[p] = s.xi;
    pk.yi[p] = s.yi;
    pk.status[p] = s.status;
    pk.iter[p] = s.iter;
    pk.added[p] = s.added;
  }
}

// ---------------------------------------------------------------------------
// Gaussian model (dao_fit.c)
// ---------------------------------------------------------------------------

// daoCalcWidthsFromZ (dao_fit.c:1009-1035); also returns the squared terms.
__device__ inline void widths_from_z(const Model& m, PeakState& s, double& wx_term, double& wy_term) {
  double u = (s.prm[P_ZC] - m.wxz[1]) / m.wxz[2];
  double u2 = u * u;
  double t = 1.0 + u2 + m.wxz[3] * u2 * u + m.wxz[4] * u2 * u2;
  wx_term = t * t;
  s.prm[P_XW] = 2.0 / (m.wxz[0] * t);
  u = (s.prm[P_ZC] - m.wyz[1]) / m.wyz[2];
  u2 = u * u;
  t = 1.0 + u2 + m.wyz[3] * u2 * u + m.wyz[4] * u2 * u2;
  wy_term = t * t;
  s.prm[P_YW] = 2.0 / (m.wyz[0] * t);
}

// Chain factors d(width)/dz of daoCalcJHZ (dao_fit.c:904-916).
__device__ inline void z_grad(const Model& m, const PeakState& s, double& gx, double& gy) {
  double u = (s.prm[P_ZC] - m.wxz[1]) / m.wxz[2];
  double u2 = u * u;
  double t = 1.0 + u2 + m.wxz[3] * u2 * u + m.wxz[4] * u2 * u2;
  double zt = 2.0 * u + 3.0 * m.wxz[3] * u2 + 4.0 * m.wxz[4] * u2 * u;
  gx = -2.0 * zt / (m.wxz[0] * (t * t));
  u = (s.prm[P_ZC] - m.wyz[1]) / m.wyz[2];
  u2 = u * u;
  t = 1.0 + u2 + m.wyz[3] * u2 * u + m.wyz[4] * u2 * u2;
  zt = 2.0 * u + 3.0 * m.wyz[3] * u2 + 4.0 * m.wyz[4] * u2 * u;
  gy = -2.0 * zt / (m.wyz[0] * (t * t));
}

// One separable shape sample per lane: lanes 0..31 hold column offsets
// xt = i - (X - xi) - xoff and ext = exp(-xt^2 XW); lanes 32..63 the rows
// (dao_fit.c:968-998; the row offset also uses xoff, dao_fit.c:979).
__device__ inline void lane_shape(const Model& m, const PeakState& s, double& t, double& e) {
  int l = lane_id();
  int i = l & 31;
  bool row = l >= 32;
  double o = row ? -(s.prm[P_YC] - (double)s.yi) - m.xoff : -(s.prm[P_XC] - (double)s.xi) - m.xoff;
  double w = row ? s.prm[P_YW] : s.prm[P_XW];
  t = (double)i + o;
  e = exp(-t * t * w);
}

// ---------------------------------------------------------------------------
// ROI pixel cache: pixel j of the ROI lives in lane j % 64, slot j / 64
// ---------------------------------------------------------------------------

template <int PPL>
struct Roi {
  // per slot: column | row << 8 of the ROI pixel, and its offset row*W+col
  int rxy[PPL];
  int koff[PPL];
  double f[PPL], bg[PPL], tfi[PPL], psf[PPL];
  int cnt[PPL];
  unsigned stale;  // bit s: slot s is stale
  unsigned okm;    // bit s: slot s is a pixel of the ROI
  long long base;  // pixel index of the ROI origin the cache holds
  int at_x, at_y;
  bool valid;
  // read-only pixel data, read from memory when needed
  const double* x;
  const double* rqe;
  const double* scm;

  __device__ bool ok(int s) const { return (okm >> s) & 1u; }
  __device__ int rx(int s) const { return rxy[s] & 0xff; }
  __device__ int ry(int s) const { return rxy[s] >> 8; }
  __device__ long long k(int s) const { return base + koff[s]; }

  __device__ void init(const Model& m, const int* roi_tab, const Pix& px, const Frame& fr) {
    int l = lane_id();
    okm = 0;
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      int j = l + s * kWave;
      bool o = j < m.nroi;
      int v = o ? roi_tab[j] : 0;
      int cx = v & 0xffff, cy = v >> 16;
      rxy[s] = cx | (cy << 8);
      koff[s] = cy * fr.W + cx;
      if (o) okm |= 1u << s;
    }
    x = px.x;
    rqe = px.rqe;
    scm = px.scmos;
    valid = false;
  }

  __device__ void load(const Pix& px, const Frame& fr, int xi, int yi) {
    base = fr.pix0 + (long long)yi * fr.W + xi;
    stale = 0;
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      if (ok(s)) {
        long long q = k(s);
        f[s] = px.f[q];
        bg[s] = px.bg[q];
        cnt[s] = px.cnt[q];
        tfi[s] = px.tfi[q];
        if (px.stale[q]) stale |= 1u << s;
      }
    }
    at_x = xi;
    at_y = yi;
    valid = true;
  }

  __device__ void store(const Pix& px, const Frame& fr) {
    if (!valid) return;
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      if (ok(s)) {
        long long q = k(s);
        px.f[q] = f[s];
        px.bg[q] = bg[s];
        px.cnt[q] = cnt[s];
        px.tfi[q] = tfi[s];
        px.stale[q] = (stale >> s) & 1u;
      }
    }
  }

  // Make the cache hold the ROI at (xi, yi).
  __device__ void move_to(const Pix& px, const Frame& fr, int xi, int yi) {
    if (valid && at_x == xi && at_y == yi) return;
    store(px, fr);
    load(px, fr, xi, yi);
  }

  // psf at the cache's pixels for state st (cache must be at st.xi, st.yi).
  __device__ void shape(const Model& m, const PeakState& st) {
    double t, e;
    lane_shape(m, st, t, e);
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      double ex = __shfl(e, rx(s), kWave);
      double ey = __shfl(e, 32 + ry(s), kWave);
      psf[s] = ey * ex;
    }
  }

  // mFitAddPeak / mFitSubtractPeak (multi_fit.c:28-57, 1894-1922).
  __device__ void stamp(const PeakState& st, int sign) {
    double h = st.prm[P_HEIGHT];
    double b = st.prm[P_BG];
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      if (ok(s)) {
        long long q = k(s);
        double r = rqe[q];
        double sc = scm[q];
        if (sign > 0) {
          f[s] += h * psf[s] * r;
          cnt[s] += 1;
          bg[s] += b * r + sc;
        } else {
          f[s] -= h * psf[s] * r;
          cnt[s] -= 1;
          bg[s] -= (b * r + sc);
        }
      }
    }
    stale = okm;
  }
};

__device__ inline double anscombe(double v) {
  if (v < (-0.375 + 1.0e-6)) return 0.0;
  return 2.0 * sqrt(v + 0.375);
}

// mFitCalcErr{,ALS,LS,DWLS,FWLS} (multi_fit.c:111-458) over the cache's ROI.
// Returns 1 when MLE/FWLS meet fi <= 0 at some pixel: pixels before the first
// such pixel (in ROI order) are refreshed, the rest keep their cached fi, and
// the error is not updated.  Counts n_neg_fi into *neg_fi.
template <int PPL>
__device__ int eval_error(const Model& m, Roi<PPL>& c, const double* asx_frame, const Frame& fr,
                          double& err, int& neg_fi) {
  const int l = lane_id();
  double fi[PPL];
  int first_bad = 0x7fffffff;
#pragma unroll
  for (int s = 0; s < PPL; s++) {
    fi[s] = 0.0;
    if (c.ok(s)) {
      fi[s] = c.f[s] + c.bg[s] / ((double)c.cnt[s]);
      if ((m.errm == EM_MLE || m.errm == EM_FWLS) && fi[s] <= 0.0) {
        int j = l + s * kWave;
        first_bad = j < first_bad ? j : first_bad;
      }
    }
  }
  first_bad = wmin_i(first_bad);
  double sum = 0.0;
  int zeros = 0;
#pragma unroll
  for (int s = 0; s < PPL; s++) {
    int j = l + s * kWave;
    if (c.ok(s) && j < first_bad) {
      long long q = c.k(s);
      double xk = c.x[q];
      double e;
      double t = fi[s];
      if (m.errm == EM_MLE) {
        e = 2.0 * ((t - xk) - xk * log(t / xk));
      } else if (m.errm == EM_ALS) {
        double a = anscombe(t);
        if (a == 0.0 && ((c.stale >> s) & 1u)) zeros++;
        t = a;
        double d = a - asx_frame[q];
        e = d * d;
      } else if (m.errm == EM_LS) {
        double d = t - xk;
        e = d * d;
      } else if (m.errm == EM_DWLS) {
        double d = t - xk;
        e = d * d / xk;
      } else {
        double d = t - xk;
        e = d * d / t;
      }
      c.tfi[s] = t;
      c.stale &= ~(1u << s);
      sum += e;
    }
  }
  if (m.errm == EM_ALS) neg_fi += wsum_i(zeros);
  if (first_bad != 0x7fffffff) {
    neg_fi += 1;
    return 1;
  }
  err = wsum(sum);
  return 0;
}

// daoCalcJH* (dao_fit.c:48-958): J (jac) and the upper triangle of H, packed
// row by row (H[0..N-1] = row 0, ...), reduced over the wave.  Reads the cached
// transformed fit tfi.
template <int PPL, int SHAPE>
__device__ void calc_jh(const Model& m, const Roi<PPL>& c, const PeakState& st, const double* asx,
                        const Frame& fr, double* J, double* H) {
  constexpr int N = SHAPE == SH_2DFIXED ? 4 : (SHAPE == SH_3D ? 6 : 5);
  double a1 = st.prm[P_HEIGHT];
  double wx = st.prm[P_XW];
  double wy = (SHAPE == SH_3D || SHAPE == SH_Z) ? st.prm[P_YW] : wx;
  double gx = 0.0, gy = 0.0;
  if (SHAPE == SH_Z) z_grad(m, st, gx, gy);
  double t, e;
  lane_shape(m, st, t, e);

  double j_acc[N];
  double h_acc[N * (N + 1) / 2];
#pragma unroll
  for (int a = 0; a < N; a++) j_acc[a] = 0.0;
#pragma unroll
  for (int a = 0; a < N * (N + 1) / 2; a++) h_acc[a] = 0.0;

#pragma unroll
  for (int s = 0; s < PPL; s++) {
    double xt = __shfl(t, c.rx(s), kWave);
    double ex = __shfl(e, c.rx(s), kWave);
    double yt = __shfl(t, 32 + c.ry(s), kWave);
    double ey = __shfl(e, 32 + c.ry(s), kWave);
    if (c.ok(s)) {
      const long long kq = c.k(s);
      double fi = c.tfi[s];
      double xk = c.x[kq];
      double et = ex * ey;
      double g[N];
      if (SHAPE == SH_2DFIXED) {
        g[0] = et;
        g[1] = 2.0 * a1 * wx * xt * et;
        g[2] = 2.0 * a1 * wy * yt * et;
        g[3] = 1.0;
      } else if (SHAPE == SH_2D) {
        g[0] = et;
        g[1] = 2.0 * a1 * wx * xt * et;
        g[2] = 2.0 * a1 * wy * yt * et;
        g[3] = -a1 * xt * xt * et - a1 * yt * yt * et;
        g[4] = 1.0;
      } else if (SHAPE == SH_3D) {
        g[0] = et;
        g[1] = 2.0 * a1 * wx * xt * et;
        g[2] = -a1 * xt * xt * et;
        g[3] = 2.0 * a1 * wy * yt * et;
        g[4] = -a1 * yt * yt * et;
        g[5] = 1.0;
      } else {
        g[0] = et;
        g[1] = 2.0 * a1 * wx * xt * et;
        g[2] = 2.0 * a1 * wy * yt * et;
        g[3] = -a1 * xt * xt * gx * et - a1 * yt * yt * gy * et;
        g[4] = 1.0;
      }
      double sc = (m.errm == EM_ALS) ? c.rqe[kq] * 2.0 / fi : c.rqe[kq];
#pragma unroll
      for (int a = 0; a < N; a++) g[a] *= sc;
      double wj, wh;
      if (m.errm == EM_MLE) {
        wj = 2.0 * (1.0 - xk / fi);
        wh = 2.0 * xk / (fi * fi);
      } else if (m.errm == EM_ALS) {
        wj = fi - asx[kq];
        wh = 1.0;
      } else if (m.errm == EM_LS) {
        wj = fi - xk;
        wh = 1.0;
      } else if (m.errm == EM_DWLS) {
        wj = (fi - xk) / xk;
        wh = 1.0 / xk;
      } else {
        double t1 = (fi - xk) / fi;
        double t2 = 1.0 / fi;
        wj = 2.0 * t1 - t1 * t1;
        wh = 2.0 * t2 * (xk * t2 - t1 + t1 * t1);
      }
      int q = 0;
#pragma unroll
      for (int a = 0; a < N; a++) {
        j_acc[a] += wj * g[a];
#pragma unroll
        for (int b = a; b < N; b++) h_acc[q++] += wh * g[a] * g[b];
      }
    }
  }
#pragma unroll
  for (int a = 0; a < N; a++) J[a] = wsum(j_acc[a]);
#pragma unroll
  for (int a = 0; a < N * (N + 1) / 2; a++) H[a] = wsum(h_acc[a]);
}

// mFitSolve (multi_fit.c:1839-1883) with LAPACK dposv "Lower" on row-major
// storage = Cholesky of the symmetric matrix given by the upper triangle.
// Non-zero on a pivot that is not > 0 (or NaN) or a NaN in the solution.
// H is the packed upper triangle; the diagonal is scaled by (1 + lambda)
// (multi_fit.c:1174-1185).
template <int N>
__device__ constexpr int upk(int r, int c) {  // index of (r, c), r <= c
  return r * N - r * (r - 1) / 2 + (c - r);
}

template <int N>
__device__ inline int chol_solve(const double* H, double lam, double* b) {
  double L[N][N];
#pragma unroll
  for (int j = 0; j < N; j++) {
    double s = (1.0 + lam) * H[upk<N>(j, j)];
#pragma unroll
    for (int k = 0; k < j; k++) s -= L[j][k] * L[j][k];
    if (!(s > 0.0)) return j + 1;
    double d = sqrt(s);
    L[j][j] = d;
#pragma unroll
    for (int i = j + 1; i < N; i++) {
      double t = H[upk<N>(j, i)];
#pragma unroll
      for (int k = 0; k < j; k++) t -= L[i][k] * L[j][k];
      L[i][j] = t / d;
    }
  }
#pragma unroll
  for (int i = 0; i < N; i++) {
    double t = b[i];
#pragma unroll
    for (int k = 0; k < i; k++) t -= L[i][k] * b[k];
    b[i] = t / L[i][i];
  }
#pragma unroll
  for (int i = N - 1; i >= 0; i--) {
    double t = b[i];
#pragma unroll
    for (int k = i + 1; k < N; k++) t -= L[k][i] * b[k];
    b[i] = t / L[i][i];
  }
#pragma unroll
  for (int i = 0; i < N; i++)
    if (isnan(b[i])) return 1;
  return 0;
}

// mFitUpdate (multi_fit.c:1933-1946).
__device__ inline void roi_follow(PeakState& s) {
  double dx = s.prm[P_XC] - s.xi;
  if ((dx < -0.5) || (dx > 1.5)) s.xi = (int)floor(s.prm[P_XC]);
  double dy = s.prm[P_YC] - s.yi;
  if ((dy < -0.5) || (dy > 1.5)) s.yi = (int)floor(s.prm[P_YC]);
}

// daoUpdate2DFixed/2D/3D/Z (dao_fit.c:1777-1880).
template <int SHAPE>
__device__ inline void apply_update(const Model& m, PeakState& s, const double* d) {
  if (SHAPE == SH_2DFIXED) {
    s.prm[P_HEIGHT] -= d[0];
    s.prm[P_XC] -= d[1];
    s.prm[P_YC] -= d[2];
    s.prm[P_BG] -= d[3];
    roi_follow(s);
  } else if (SHAPE == SH_2D) {
    s.prm[P_HEIGHT] -= d[0];
    s.prm[P_XC] -= d[1];
    s.prm[P_YC] -= d[2];
    s.prm[P_XW] -= d[3];
    s.prm[P_BG] -= d[4];
    s.prm[P_YW] = s.prm[P_XW];
    roi_follow(s);
  } else if (SHAPE == SH_3D) {
    s.prm[P_HEIGHT] -= d[0];
    s.prm[P_XC] -= d[1];
    s.prm[P_XW] -= d[2];
    s.prm[P_YC] -= d[3];
    s.prm[P_YW] -= d[4];
    s.prm[P_BG] -= d[5];
    roi_follow(s);
  } else {
    s.prm[P_HEIGHT] -= d[0];
    s.prm[P_XC] -= d[1];
    s.prm[P_YC] -= d[2];
    s.prm[P_ZC] -= d[3];
    s.prm[P_BG] -= d[4];
    if (s.prm[P_ZC] < m.zmin) s.prm[P_ZC] = m.zmin;
    if (s.prm[P_ZC] > m.zmax) s.prm[P_ZC] = m.zmax;
    roi_follow(s);
    double a, b;
    widths_from_z(m, s, a, b);
  }
}

// ROI origin inside the frame (every pixel of the ROI addressable).  The
// reference never checks this before touching pixels; a peak whose origin is
// off-image (e.g. an ERROR peak set back to RUNNING by the caller) would read
// outside its arrays there, and is turned into an ERROR peak here instead.
__device__ inline bool origin_ok(const Model& m, const Frame& fr, int xi, int yi) {
  return xi >= 0 && yi >= 0 && xi + m.roi <= fr.W && yi + m.roi <= fr.H;
}

// mFitCheck (multi_fit.c:470-502).
__device__ inline int margin_height_check(const Model& m, const Frame& fr, const PeakState& s,
                                          int* d) {
  if ((s.xi < 0) || (s.xi >= (fr.W - m.roi)) || (s.yi < 0) || (s.yi >= (fr.H - m.roi))) {
    d[C_MARGIN]++;
    return 1;
  }
  if (s.prm[P_HEIGHT] < 0.0) {
    d[C_NEG_HEIGHT]++;
    return 1;
  }
  return 0;
}

// daoCheck2DFixed/2D/3D/Z (dao_fit.c:1047-1133): 2D and 3D clamp the widths.
__device__ inline int peak_check(const Model& m, const Frame& fr, PeakState& s, int* d) {
  if (margin_height_check(m, fr, s, d)) return 1;
  if (m.shape == SH_2D || m.shape == SH_3D) {
    if (s.prm[P_XW] < m.wmin) s.prm[P_XW] = m.wmin;
    else if (s.prm[P_XW] > m.wmax) s.prm[P_XW] = m.wmax;
  }
  if (m.shape == SH_3D) {
    if (s.prm[P_YW] < m.wmin) s.prm[P_YW] = m.wmin;
    else if (s.prm[P_YW] > m.wmax) s.prm[P_YW] = m.wmax;
  }
  return 0;
}

// mFitResetPeak (multi_fit.c:1788-1798): back to the turn's start (still in
// the peak's arrays), keeping `added`, lambda times 4.
__device__ inline void reset_from(PeakState& w, const Peaks& pk, int p) {
  int added = w.added;
  double lam = w.lam;
  load_state(pk, p, w);
  w.added = added;
  w.lam = lam * 4.0;
}

// ---------------------------------------------------------------------------
// ordered executor: predecessor lists and readiness
// ---------------------------------------------------------------------------

__device__ inline bool boxes_overlap(int4 a, int4 b) {
  return (a.x < b.z) && (b.x < a.z) && (a.y < b.w) && (b.y < a.w);
}

__device__ inline bool is_big(const Sched& sc, int4 b) {
  return (b.z - b.x) > sc.cell || (b.w - b.y) > sc.cell;
}

// Calls fn(j) for every op j of the same frame that may overlap op i's box:
// ops binned in the cells that can reach the box, plus the oversized ops.
template <typename Fn>
__device__ inline void for_candidates(const Sched& sc, const Frame& fr, int4 b, int lane, Fn fn) {
  int cx0 = (b.x - sc.cell + 1) >= 0 ? (b.x - sc.cell + 1) / sc.cell : 0;
  int cy0 = (b.y - sc.cell + 1) >= 0 ? (b.y - sc.cell + 1) / sc.cell : 0;
  int cx1 = (b.z - 1) / sc.cell;
  int cy1 = (b.w - 1) / sc.cell;
  if (cx1 >= fr.ncx) cx1 = fr.ncx - 1;
  if (cy1 >= fr.ncy) cy1 = fr.ncy - 1;
  for (int cy = cy0; cy <= cy1; cy++) {
    for (int cx = cx0; cx <= cx1; cx++) {
      int c = fr.cell0 + cy * fr.ncx + cx;
      int s0 = sc.cell_start[c], s1 = sc.cell_start[c + 1];
      for (int q = s0 + lane; q < s1; q += kWave) fn(sc.cell_ops[q]);
    }
  }
}

template <typename Fn>
__device__ inline void for_big(const Sched& sc, int lane, Fn fn) {
  int nb = *sc.n_big;
  for (int q = lane; q < nb; q += kWave) fn(sc.big_ops[q]);
}

// Cell of a box (by its origin), clamped into the frame's grid: an empty box
// at the image edge may start at x = W.
__device__ inline int cell_of(const Sched& sc, const Frame& fr, int4 b) {
  int cx = min(max(b.x / sc.cell, 0), fr.ncx - 1);
  int cy = min(max(b.y / sc.cell, 0), fr.ncy - 1);
  return fr.cell0 + cy * fr.ncx + cx;
}

// cell binning: count
__global__ void k_bin_count(Sched sc, const Frame* frames, const Peaks pk) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int4 b = sc.box[i];
  if (is_big(sc, b)) {
    int q = atomicAdd(sc.n_big, 1);
    sc.big_ops[q] = i;
    return;
  }
  const Frame fr = frames[pk.frame[sc.op_peak[i]]];
  int c = cell_of(sc, fr, b);
  atomicAdd(&sc.cell_fill[c], 1);
}

// exclusive scan of cell_fill into cell_start (single block)
__global__ void k_scan_cells(Sched sc, int ncells) {
  __shared__ int part[1024];
  int t = threadIdx.x;
  int per = (ncells + blockDim.x - 1) / blockDim.x;
  int a = t * per, b = a + per < ncells ? a + per : ncells;
  int s = 0;
  for (int i = a; i < b; i++) s += sc.cell_fill[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < (int)blockDim.x; o <<= 1) {
    int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int i = a; i < b; i++) {
    sc.cell_start[i] = run;
    run += sc.cell_fill[i];
    sc.cell_fill[i] = 0;
  }
  if (t == blockDim.x - 1) sc.cell_start[ncells] = part[t];
}

__global__ void k_bin_scatter(Sched sc, const Frame* frames, const Peaks pk) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int4 b = sc.box[i];
  if (is_big(sc, b)) return;
  const Frame fr = frames[pk.frame[sc.op_peak[i]]];
  int c = cell_of(sc, fr, b);
  int q = atomicAdd(&sc.cell_fill[c], 1);
  int pos = sc.cell_start[c] + q;
  if (pos < sc.nops) sc.cell_ops[pos] = i;
}

// Predecessors of op i: earlier ops of the same frame with overlapping boxes.
// Ops are ordered by (frame, peak index) so "earlier" is j < i.
__global__ void k_preds(Sched sc, const Frame* frames, const Peaks pk) {
  int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  int lane = lane_id();
  if (i >= sc.nops) return;
  int fi = pk.frame[sc.op_peak[i]];
  const Frame fr = frames[fi];
  int4 b = sc.box[i];
  int count = 0;
  auto take = [&](int j) {
    bool hit = (j < i) && (pk.frame[sc.op_peak[j]] == fi) && boxes_overlap(b, sc.box[j]);
    // compact the hits of this wave-wide step
    unsigned long long m = __ballot(hit);
    if (hit) {
      int pos = count + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < kMaxPred) sc.pred[(long long)i * kMaxPred + pos] = j;
    }
    count += __popcll(m);
  };
  // the lambdas are called by all lanes in lock step (loop bounds uniform per
  // cell), lanes past the end pass j = i which is never a hit
  int cx0 = (b.x - sc.cell + 1) >= 0 ? (b.x - sc.cell + 1) / sc.cell : 0;
  int cy0 = (b.y - sc.cell + 1) >= 0 ? (b.y - sc.cell + 1) / sc.cell : 0;
  int cx1 = (b.z - 1) / sc.cell;
  int cy1 = (b.w - 1) / sc.cell;
  if (cx1 >= fr.ncx) cx1 = fr.ncx - 1;
  if (cy1 >= fr.ncy) cy1 = fr.ncy - 1;
  for (int cy = cy0; cy <= cy1; cy++) {
    for (int cx = cx0; cx <= cx1; cx++) {
      int c = fr.cell0 + cy * fr.ncx + cx;
      int s0 = sc.cell_start[c], s1 = sc.cell_start[c + 1];
      for (int q = s0; q < s1; q += kWave) take(q + lane < s1 ? sc.cell_ops[q + lane] : i);
    }
  }
  int nb = *sc.n_big;
  for (int q = 0; q < nb; q += kWave) take(q + lane < nb ? sc.big_ops[q + lane] : i);
  if (lane == 0) {
    sc.pred_n[i] = count;
    sc.done[i] = -1;
  }
}

// Op i may run in round r when every predecessor finished in an earlier
// round (a predecessor finishing in this same launch counts as not done).
__device__ bool op_ready(const Sched& sc, const Frame* frames, const Peaks& pk, int i, int round) {
  int lane = lane_id();
  int n = sc.pred_n[i];
  bool wait = false;
  auto pending = [&](int j) {
    int d = __hip_atomic_load(&sc.done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return d < 0 || d >= round;
  };
  if (n <= kMaxPred) {
    for (int q = lane; q < n; q += kWave)
      if (pending(sc.pred[(long long)i * kMaxPred + q])) wait = true;
  } else {
    int fi = pk.frame[sc.op_peak[i]];
    const Frame fr = frames[fi];
    int4 b = sc.box[i];
    auto chk = [&](int j) {
      if (j < i && pk.frame[sc.op_peak[j]] == fi && boxes_overlap(b, sc.box[j]) && pending(j))
        wait = true;
    };
    for_candidates(sc, fr, b, lane, chk);
    for_big(sc, lane, chk);
  }
  return !__any(wait);
}

__device__ inline void op_finish(const Sched& sc, int i, int round, int flags, const int* d) {
  if (lane_id() == 0) {
#pragma unroll
    for (int q = 0; q < C_COUNT; q++) sc.op_delta[(long long)i * C_COUNT + q] = d[q];
    sc.op_flags[i] = flags;
    __hip_atomic_store(&sc.done[i], round, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicSub(sc.remaining, 1);
  }
}

// ---------------------------------------------------------------------------
// op kernels
// ---------------------------------------------------------------------------

// One Levenberg-Marquardt turn of peak p: the body of the loop at
// multi_fit.c:1089-1345.  The turn may touch only pixels inside `box`.
// Returns 1 when the turn completed, 2 when a candidate ROI left the box: the
// frame is then exactly as before the turn (the starting contribution is put
// back), the peak state is untouched, and esc_box holds the box needed.
// d[] receives the turn's counter increments.
template <int PPL, int SHAPE>
__device__ int lm_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr, const int* roi_tab,
                       int p, int4 box, int* d, int4& esc_box) {
  constexpr int N = SHAPE == SH_2DFIXED ? 4 : (SHAPE == SH_3D ? 6 : 5);
  // The peak's arrays keep the turn's starting state until the turn ends, so
  // resets read it back from there instead of holding a second copy.
  PeakState w;
  load_state(pk, p, w);
  if (!origin_ok(m, fr, w.xi, w.yi)) {
    w.status = ST_ERROR;
    store_state(pk, p, w);
    return 1;
  }

  Roi<PPL> c;
  c.init(m, roi_tab, px, fr);
  c.load(px, fr, w.xi, w.yi);
  c.shape(m, w);

  double J[N], H[N * (N + 1) / 2], wJ[N];
  double e;
  if (eval_error<PPL>(m, c, px.asx, fr, e, d[C_NEG_FI]) == 0) w.err = e;
  const double start = w.err;
  calc_jh<PPL, SHAPE>(m, c, w, px.asx, fr, J, H);
  c.stamp(w, -1);
  w.added--;

  bool escaped = false;
  // Every retry multiplies lambda by 4 from >= 7.5e-4, so more than 60
  // cycles cannot happen; the cap only guarantees termination.
  for (int guard = 0;; guard++) {
    if (w.lam > 1.0e20 || guard > 200) {
      d[C_LOST]++;
      w.status = ST_ERROR;
      break;
    }
    w.iter++;
    d[C_ITER]++;
    w.status = ST_RUNNING;

#pragma unroll
    for (int r = 0; r < N; r++) wJ[r] = J[r];
    if (chol_solve<N>(H, w.lam, wJ) != 0) {
      d[C_DPOSV]++;
      w.status = ST_ERROR;
      w.lam = w.lam * 4.0;
      continue;
    }

    apply_update<SHAPE>(m, w, wJ);
    if (peak_check(m, fr, w, d)) {
      reset_from(w, pk, p);
      w.status = ST_ERROR;
      continue;
    }

    if (w.xi < box.x || w.yi < box.y || w.xi + m.roi > box.z || w.yi + m.roi > box.w) {
      escaped = true;
      break;
    }

    c.move_to(px, fr, w.xi, w.yi);
    c.shape(m, w);
    c.stamp(w, +1);
    w.added++;

    if (eval_error<PPL>(m, c, px.asx, fr, e, d[C_NEG_FI]) != 0) {
      c.stamp(w, -1);
      w.added--;
      reset_from(w, pk, p);
      w.status = ST_ERROR;
      continue;
    }
    w.err = e;

    if (w.err > start) {
      if (((w.err - start) / start) < m.tol) {
        w.status = ST_CONVERGED;
        break;
      }
      d[C_NON_DECR]++;
      c.stamp(w, -1);
      w.added--;
      reset_from(w, pk, p);
    } else {
      if (((start - w.err) / start) < m.tol) {
        w.status = ST_CONVERGED;
        break;
      }
      if (w.lam > 1.0e-3) w.lam = 0.75 * w.lam;
      break;
    }
  }

  if (escaped) {
    // Put the frame back: re-add the turn's starting contribution.  The
    // escape is detected before the candidate touches any pixel.
    const int cx = w.xi, cy = w.yi;
    load_state(pk, p, w);
    c.move_to(px, fr, w.xi, w.yi);
    c.shape(m, w);
    c.stamp(w, +1);
    c.store(px, fr);
    w.xi = cx;
    w.yi = cy;
    esc_box.x = min(box.x, w.xi - m.margin);
    esc_box.y = min(box.y, w.yi - m.margin);
    esc_box.z = max(box.z, w.xi + m.roi + m.margin);
    esc_box.w = max(box.w, w.yi + m.roi + m.margin);
#pragma unroll
    for (int q = 0; q < C_COUNT; q++) d[q] = 0;
    return 2;
  }

  c.store(px, fr);
  store_state(pk, p, w);
  return 1;
}

// mFitRecenterPeaks body (multi_fit.c:1656-1688) for peak p, whose centre
// left [-0.25, 1.25] of its ROI origin.
template <int PPL>
__device__ void recenter_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr,
                              const int* roi_tab, int p, int* d) {
  PeakState w;
  load_state(pk, p, w);
  if (!origin_ok(m, fr, w.xi, w.yi)) {
    w.status = ST_ERROR;
    store_state(pk, p, w);
    return;
  }
  Roi<PPL> c;
  c.init(m, roi_tab, px, fr);
  c.load(px, fr, w.xi, w.yi);
  c.shape(m, w);
  c.stamp(w, -1);
  w.added--;
  w.xi = (int)floor(w.prm[P_XC]);
  w.yi = (int)floor(w.prm[P_YC]);
  if (margin_height_check(m, fr, w, d)) {
    d[C_LOST] += 1;
    w.status = ST_ERROR;
  } else {
    c.move_to(px, fr, w.xi, w.yi);
    c.shape(m, w);
    c.stamp(w, +1);
    w.added++;
  }
  c.store(px, fr);
  store_state(pk, p, w);
}

// Add (sign +1) or remove (-1) `added` copies of state s's contribution with
// atomic adds.
template <int PPL>
__device__ void undo_stamp(const Model& m, const Pix& px, const Frame& fr, const Roi<PPL>& c,
                           const PeakState& s, int sign) {
  const int n = s.added * sign;
  if (n == 0 || !origin_ok(m, fr, s.xi, s.yi)) return;
  double t, e;
  lane_shape(m, s, t, e);
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    double ex = __shfl(e, c.rx(q), kWave);
    double ey = __shfl(e, 32 + c.ry(q), kWave);
    if (c.ok(q)) {
      long long k = fr.pix0 + (long long)s.yi * fr.W + s.xi + c.koff[q];
      double r = px.rqe[k];
      atomicAdd(&px.f[k], (double)n * (s.prm[P_HEIGHT] * (ey * ex) * r));
      atomicAdd(&px.bg[k], (double)n * (s.prm[P_BG] * r + px.scmos[k]));
      atomicAdd(&px.cnt[k], n);
      px.stale[k] = 1;
    }
  }
}

// Undo a completed LM turn of peak p: the image loses the peak's current
// contribution and regains the one it had at the attempt's start (backup),
// and the peak state returns to the backup.  Callers order or serialise
// undos of overlapping peaks; atomic adds keep concurrent undos safe.
template <int PPL>
__device__ void undo_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr, const int* roi_tab,
                          int p) {
  PeakState w, s0;
  load_state(pk, p, w);
  s0 = w;
#pragma unroll
  for (int q = 0; q < kNFit; q++) s0.prm[q] = pk.bk_prm[(long long)q * pk.cap + p];
  s0.lam = pk.bk_lam[p];
  s0.err = pk.bk_err[p];
  s0.xi = pk.bk_xi[p];
  s0.yi = pk.bk_yi[p];
  s0.status = pk.bk_status[p];
  s0.iter = pk.bk_iter[p];
  s0.added = pk.bk_added[p];
  Roi<PPL> c;
  c.init(m, roi_tab, px, fr);
  undo_stamp<PPL>(m, px, fr, c, w, -1);
  undo_stamp<PPL>(m, px, fr, c, s0, +1);
  store_state(pk, p, s0);
}

// Executor round of LM turns (multi-kernel path).
template <int PPL, int SHAPE>
__global__ __launch_bounds__(256) void k_lm(OpArgs a) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= a.sc.nops) return;
  if (a.sc.done[i] >= 0) return;
  if (!op_ready(a.sc, a.frames, a.pk, i, a.round)) return;
  const int p = a.sc.op_peak[i];
  const int fi = a.pk.frame[p];
  const Frame fr = a.frames[fi];
  int d[C_COUNT];
#pragma unroll
  for (int q = 0; q < C_COUNT; q++) d[q] = 0;
  int4 eb;
  int r = lm_turn<PPL, SHAPE>(a.m, a.px, a.pk, fr, a.roi_tab, p, a.sc.box[i], d, eb);
  if (r == 2 && lane_id() == 0) {
    a.sc.esc_box[i] = eb;
    a.sc.frame_bad[fi] = 1;
  }
  op_finish(a.sc, i, a.round, r, d);
}

// Executor round of recentering (multi-kernel path).
template <int PPL>
__global__ __launch_bounds__(256) void k_recenter(OpArgs a) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= a.sc.nops) return;
  if (a.sc.done[i] >= 0) return;
  if (!op_ready(a.sc, a.frames, a.pk, i, a.round)) return;
  const int p = a.sc.op_peak[i];
  const Frame fr = a.frames[a.pk.frame[p]];
  int d[C_COUNT];
#pragma unroll
  for (int q = 0; q < C_COUNT; q++) d[q] = 0;
  recenter_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, p, d);
  op_finish(a.sc, i, a.round, 1, d);
}

// daoNewPeaks body (dao_fit.c:1605-1751) for one new peak; the start values
// come from rows of 4 ({x,y,z,sigma}) or 7 ({x,y,z,bg,h,sx,sy}) doubles.
template <int PPL>
__global__ __launch_bounds__(256) void k_newpeak(OpArgs a, const double* rows, int fmt, int first_slot) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= a.sc.nops) return;
  if (a.sc.done[i] >= 0) return;
  if (!op_ready(a.sc, a.frames, a.pk, i, a.round)) return;
  const Model& m = a.m;
  const int p = a.sc.op_peak[i];
  const Frame fr = a.frames[a.pk.frame[p]];
  int d[C_COUNT];
#pragma unroll
  for (int q = 0; q < C_COUNT; q++) d[q] = 0;
  const int row = i;  // rows are given in op order

  PeakState w;
#pragma unroll
  for (int q = 0; q < kNFit; q++) w.prm[q] = 0.0;
  w.added = 0;
  w.iter = 0;
  w.status = ST_RUNNING;
  w.err = 0.0;
  w.lam = 1.0;
  double wxt, wyt;
  if (fmt == PF_TEXT) {
    const double* q = rows + 7LL * row;
    w.prm[P_XC] = q[0] - m.xoff;
    w.prm[P_YC] = q[1] - m.xoff;
    w.prm[P_ZC] = q[2] - m.zoff;
    w.prm[P_BG] = q[3];
    w.prm[P_HEIGHT] = q[4];
    w.prm[P_XW] = 1.0 / (2.0 * q[5] * q[5]);
    w.prm[P_YW] = 1.0 / (2.0 * q[6] * q[6]);
    if (m.shape == SH_Z) widths_from_z(m, w, wxt, wyt);
  } else {
    const double* q = rows + 4LL * row;
    w.prm[P_XC] = q[0] - m.xoff;
    w.prm[P_YC] = q[1] - m.xoff;
    w.prm[P_ZC] = q[2] - m.zoff;
    if (m.shape == SH_Z) {
      widths_from_z(m, w, wxt, wyt);
    } else {
      double width = 1.0 / (2.0 * q[3] * q[3]);
      w.prm[P_XW] = width;
      w.prm[P_YW] = width;
    }
  }
  w.xi = (int)floor(w.prm[P_XC]);
  w.yi = (int)floor(w.prm[P_YC]);
  if (fmt != PF_TEXT) {
    w.prm[P_BG] = 1.0;
    w.prm[P_HEIGHT] = 1.0;
  }

  bool good = peak_check(m, fr, w, d) == 0;
  Roi<PPL> c;
  c.init(m, a.roi_tab, a.px, fr);
  if (good) {
    c.load(a.px, fr, w.xi, w.yi);
    c.shape(m, w);
    if (fmt != PF_TEXT) {
      // mFitEstimatePeakBackground (multi_fit.c:638-662)
      double sb = 0.0;
#pragma unroll
      for (int s = 0; s < PPL; s++) {
        if (c.ok(s)) {
          long long k = c.k(s);
          sb += a.px.bgest[k] / c.rqe[k];
        }
      }
      w.prm[P_BG] = wsum(sb) / ((double)m.nroi);
      if (fmt == PF_FINDER) {
        // mFitEstimatePeakHeight (multi_fit.c:678-700)
        double sp = 0.0, sx = 0.0;
#pragma unroll
        for (int s = 0; s < PPL; s++) {
          if (c.ok(s)) {
            long long k = c.k(s);
            double t = c.psf[s] * c.rqe[k];
            sp += t * t * t;
            sx += t * t * (c.x[k] - c.f[s] - a.px.bgest[k]);
          }
        }
        w.prm[P_HEIGHT] = wsum(sx) / wsum(sp);
        if (w.prm[P_HEIGHT] < m.min_height) w.prm[P_HEIGHT] = m.min_height;
        if (w.prm[P_HEIGHT] <= 0.0) {
          if (lane_id() == 0) printf("Warning peak %d has negative estimated height!\n", row);
          good = false;
        }
      }
    }
  }
  if (good) {
    c.stamp(w, +1);
    w.added++;
    c.store(a.px, fr);
  } else {
    w.status = ST_ERROR;
  }
  store_state(a.pk, p, w);
  op_finish(a.sc, i, a.round, 1, d);
}

// Initial error of new peaks (dao_fit.c:1758-1763), in peak order.
template <int PPL>
__global__ __launch_bounds__(256) void k_peakerror(OpArgs a) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= a.sc.nops) return;
  if (a.sc.done[i] >= 0) return;
  if (!op_ready(a.sc, a.frames, a.pk, i, a.round)) return;
  const Model& m = a.m;
  const int p = a.sc.op_peak[i];
  const Frame fr = a.frames[a.pk.frame[p]];
  int d[C_COUNT];
#pragma unroll
  for (int q = 0; q < C_COUNT; q++) d[q] = 0;
  PeakState w;
  load_state(a.pk, p, w);
  if (w.status == ST_RUNNING && origin_ok(m, fr, w.xi, w.yi)) {
    Roi<PPL> c;
    c.init(m, a.roi_tab, a.px, fr);
    c.load(a.px, fr, w.xi, w.yi);
    double e;
    if (eval_error<PPL>(m, c, a.px.asx, fr, e, d[C_NEG_FI]) == 0) w.err = e;
    c.store(a.px, fr);
    store_state(a.pk, p, w);
  }
  op_finish(a.sc, i, a.round, 1, d);
}

// mFitSetPeakStatus (multi_fit.c:1809-1831): subtract a peak marked ERROR.
template <int PPL>
__global__ __launch_bounds__(256) void k_subtract(OpArgs a) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= a.sc.nops) return;
  if (a.sc.done[i] >= 0) return;
  if (!op_ready(a.sc, a.frames, a.pk, i, a.round)) return;
  const Model& m = a.m;
  const int p = a.sc.op_peak[i];
  const Frame fr = a.frames[a.pk.frame[p]];
  int d[C_COUNT];
#pragma unroll
  for (int q = 0; q < C_COUNT; q++) d[q] = 0;
  PeakState w;
  load_state(a.pk, p, w);
  if (origin_ok(m, fr, w.xi, w.yi)) {
    Roi<PPL> c;
    c.init(m, a.roi_tab, a.px, fr);
    c.load(a.px, fr, w.xi, w.yi);
    c.shape(m, w);
    c.stamp(w, -1);
    w.added--;
    c.store(a.px, fr);
    store_state(a.pk, p, w);
  }
  op_finish(a.sc, i, a.round, 1, d);
}

// Undo the completed LM ops of frames flagged bad (multi-kernel path; the
// ops of one frame may overlap, the undo uses atomic adds).
template <int PPL>
__global__ __launch_bounds__(256) void k_rollback(OpArgs a) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= a.sc.nops) return;
  const int p = a.sc.op_peak[i];
  const int fi = a.pk.frame[p];
  if (!a.sc.frame_bad[fi] || (a.sc.op_flags[i] & 1) == 0) return;
  undo_turn<PPL>(a.m, a.px, a.pk, a.frames[fi], a.roi_tab, p);
}

// ---------------------------------------------------------------------------
// Frame-per-workgroup fitter
// ---------------------------------------------------------------------------
//
// One workgroup of NW waves runs whole sweeps of one frame (doFit,
// dao_fit_c.py:332-356, or one mFitIterateLM) without returning to the host.
// The running peaks of a sweep are the ops 0..n-1 in index order; wave v
// takes ops v, v+NW, v+2NW, ... in order and publishes in fin[v] the index
// of its next unfinished op.  Op j starts when no unfinished op i < j has a
// box overlapping op j's box (ops i < j with i >= fin[i % NW]).  The wave
// holding the smallest unfinished op can always start it, so the frame makes
// progress; overlapping ops run in index order, which is the reference's
// update order (multi_fit.c:1089).  An LM turn whose ROI would leave its box
// stops the others: the ops after it that already finished are undone in
// reverse order, and the sweep resumes at the turn, whose box is now the
// whole frame (so it runs alone and cannot escape again).

constexpr int kNoEscape = 0x7fffffff;

template <int NW>
struct FpwShared {
  int esc;
  int fin[NW];
  int wsum[NW];
  int cnt[C_COUNT];
};

// Order-preserving list of the frame's peaks that satisfy `mode`
// (SEL_RUNNING or SEL_RECENTER) into op_peak[peak0 + k]; returns the count.
template <int NW>
__device__ int fpw_list(const FitArgs& a, const Frame& fr, int mode, FpwShared<NW>& sh) {
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  int count = 0;  // every thread keeps the same running total
  for (int base = 0; base < fr.nfit; base += NW * kWave) {
    int t = base + threadIdx.x;
    bool flag = false;
    if (t < fr.nfit) {
      int p = fr.peak0 + t;
      if (a.pk.status[p] == ST_RUNNING) {
        if (mode == SEL_RUNNING) {
          flag = true;
        } else {
          double dx = a.pk.prm[(long long)P_XC * a.pk.cap + p] - a.pk.xi[p];
          double dy = a.pk.prm[(long long)P_YC * a.pk.cap + p] - a.pk.yi[p];
          flag = (dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25);
        }
      }
    }
    unsigned long long mask = __ballot(flag);
    int pre = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) sh.wsum[wv] = __popcll(mask);
    __syncthreads();
    int off = count;
    int tot = 0;
    for (int u = 0; u < NW; u++) {
      if (u < wv) off += sh.wsum[u];
      tot += sh.wsum[u];
    }
    if (flag) a.op_peak[fr.peak0 + off + pre] = fr.peak0 + t;
    count += tot;
    __syncthreads();
  }
  return count;
}

// Wait until op j may start.  Returns false when an earlier op escaped (the
// op must not run now) or the wait bound was hit.
template <int NW>
__device__ bool fpw_wait(const FitArgs& a, const Frame& fr, int j, int4 b, FpwShared<NW>& sh) {
  const int lane = lane_id();
  for (long long spin = 0;; spin++) {
    bool blocked = false;
    if (lane < NW) {
      int u = lane;
      int k = __hip_atomic_load(&sh.fin[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      for (; k < j; k += NW) {
        if (boxes_overlap(b, a.box[fr.peak0 + k])) {
          blocked = true;
          break;
        }
      }
    }
    if (!__any(blocked)) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      return true;
    }
    if (__hip_atomic_load(&sh.esc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < j) return false;
    if (spin > (1ll << 22)) {
      if (lane == 0) atomicExch(a.err_flag, 1);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Run ops 0..n-1 (kind OP_LM or OP_RECENTER) of the frame.
template <int PPL, int SHAPE, int NW>
__device__ void fpw_run(const FitArgs& a, const Frame& fr, int kind, int n, FpwShared<NW>& sh) {
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const int4 whole = make_int4(0, 0, fr.W, fr.H);
  int from = 0;  // ops < from are finished
  for (;;) {
    if (threadIdx.x < NW) {
      int v = threadIdx.x;
      int first = from + ((v - from % NW) % NW + NW) % NW;  // smallest index >= from owned by v
      sh.fin[v] = first;
    }
    if (threadIdx.x == 0) sh.esc = kNoEscape;
    __syncthreads();
    int first = from + ((wv - from % NW) % NW + NW) % NW;
    for (int j = first; j < n; j += NW) {
      if (__hip_atomic_load(&sh.esc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < j) break;
      const int slot = fr.peak0 + j;
      const int4 b = a.box[slot];
      // the redone op runs first and alone: the whole frame is its box
      if (!fpw_wait<NW>(a, fr, j, b, sh)) break;
      const int p = a.op_peak[slot];
      int d[C_COUNT];
#pragma unroll
      for (int q = 0; q < C_COUNT; q++) d[q] = 0;
      int r = 1;
      if (kind == OP_LM) {
        int4 eb;
        r = lm_turn<PPL, SHAPE>(a.m, a.px, a.pk, fr, a.roi_tab, p, b, d, eb);
      } else {
        recenter_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, p, d);
      }
      if (r == 2) {
        if (lane == 0) atomicMin(&sh.esc, j);
        break;
      }
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < C_COUNT; q++) {
          a.op_delta[(long long)slot * C_COUNT + q] = d[q];
          if (d[q]) atomicAdd(&sh.cnt[q], d[q]);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0) __hip_atomic_store(&sh.fin[wv], j + NW, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    const int e = sh.esc;
    if (e == kNoEscape) break;
    // An LM turn left its box.  Every op < e has finished; undo the finished
    // ops > e (newest first), redo op e alone with the whole frame as its
    // box, and continue after it.
    if (wv == 0) {
      int hi = 0;
      for (int u = 0; u < NW; u++) hi = max(hi, sh.fin[u]);
      for (int k = hi - 1; k > e; k--) {
        if (k < n && k < sh.fin[k % NW]) {
          const int slot = fr.peak0 + k;
          undo_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, a.op_peak[slot]);
          if (lane == 0)
            for (int q = 0; q < C_COUNT; q++) sh.cnt[q] -= a.op_delta[(long long)slot * C_COUNT + q];
        }
      }
      // the op's box becomes the whole frame; it is redone first below
      if (lane == 0) a.box[fr.peak0 + e] = whole;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
    __syncthreads();
    from = e;
  }
}

template <int PPL, int SHAPE, int NW>
__global__ __launch_bounds__(NW * 64) void k_fit_frames(FitArgs a) {
  __shared__ FpwShared<NW> sh;
  const int f = a.frame_list[blockIdx.x];
  const Frame fr = a.frames[f];
  const Model& m = a.m;
  if (threadIdx.x < C_COUNT) sh.cnt[threadIdx.x] = 0;
  __syncthreads();
  int sweeps = 0;
  for (int it = 0;; it++) {
    const int n = fpw_list<NW>(a, fr, SEL_RUNNING, sh);
    if (it > 0 && (n == 0 || it > a.max_more)) break;
    if (n == 0) continue;  // the unconditional first sweep found nothing
    sweeps++;
    // reserved boxes (ROI at the sweep's start + margin) and state backups
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const int slot = fr.peak0 + k;
      const int p = a.op_peak[slot];
      int4 b;
      b.x = max(0, a.pk.xi[p] - m.margin);
      b.y = max(0, a.pk.yi[p] - m.margin);
      b.z = min(fr.W, a.pk.xi[p] + m.roi + m.margin);
      b.w = min(fr.H, a.pk.yi[p] + m.roi + m.margin);
      a.box[slot] = b;
      for (int q = 0; q < kNFit; q++)
        a.pk.bk_prm[(long long)q * a.pk.cap + p] = a.pk.prm[(long long)q * a.pk.cap + p];
      a.pk.bk_lam[p] = a.pk.lam[p];
      a.pk.bk_err[p] = a.pk.err[p];
      a.pk.bk_xi[p] = a.pk.xi[p];
      a.pk.bk_yi[p] = a.pk.yi[p];
      a.pk.bk_status[p] = a.pk.status[p];
      a.pk.bk_iter[p] = a.pk.iter[p];
      a.pk.bk_added[p] = a.pk.added[p];
    }
    __syncthreads();
    fpw_run<PPL, SHAPE, NW>(a, fr, OP_LM, n, sh);
    // mFitRecenterPeaks (multi_fit.c:1644-1690)
    const int nr = fpw_list<NW>(a, fr, SEL_RECENTER, sh);
    if (nr > 0) {
      for (int k = threadIdx.x; k < nr; k += blockDim.x) {
        const int slot = fr.peak0 + k;
        const int p = a.op_peak[slot];
        int nx = (int)floor(a.pk.prm[(long long)P_XC * a.pk.cap + p]);
        int ny = (int)floor(a.pk.prm[(long long)P_YC * a.pk.cap + p]);
        int4 b;
        b.x = max(0, min(fr.W, min(a.pk.xi[p], nx)));
        b.y = max(0, min(fr.H, min(a.pk.yi[p], ny)));
        b.z = max(b.x, min(fr.W, max(a.pk.xi[p], nx) + m.roi));
        b.w = max(b.y, min(fr.H, max(a.pk.yi[p], ny) + m.roi));
        a.box[slot] = b;
      }
      __syncthreads();
      fpw_run<PPL, SHAPE, NW>(a, fr, OP_RECENTER, nr, sh);
    }
  }
  __syncthreads();
  if (threadIdx.x < C_COUNT && sh.cnt[threadIdx.x] != 0)
    atomicAdd(&a.frame_counters[f * C_COUNT + threadIdx.x], sh.cnt[threadIdx.x]);
  if (threadIdx.x == 0) a.frame_sweeps[f] = sweeps;
}

// ---------------------------------------------------------------------------
// simple per-pixel / per-peak kernels
// ---------------------------------------------------------------------------

// mFitNewImage (multi_fit.c:1382-1408) for pixels [0, n) of one frame.
__global__ void k_new_image(Pix px, long long pix0, long long n, const double* img) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long k = pix0 + i;
  px.x[k] = img[i] * px.rqe[k] + px.scmos[k];
  px.cnt[k] = 0;
  px.bg[k] = 0.0;
  px.f[k] = 0.0;
  px.stale[k] = 1;
}

// mFitNewBackground (multi_fit.c:1360-1370).
__global__ void k_new_background(Pix px, long long pix0, long long n, const double* bg) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long k = pix0 + i;
  px.bgest[k] = bg[i] * px.rqe[k];
  px.stale[k] = 1;
}

// mFitAnscombeTransformImage (multi_fit.c:90-97).
__global__ void k_anscombe(Pix px, long long pix0, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long k = pix0 + i;
  px.asx[k] = anscombe(px.x[k]);
}

// mFitGetFitImage (f / rqe) and mFitGetResidual (x - f).
__global__ void k_fit_image(Pix px, long long pix0, long long n, double* out, int residual) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long k = pix0 + i;
  out[i] = residual ? px.x[k] - px.f[k] : px.f[k] / px.rqe[k];
}

// Build the LM op list of the running peaks and their reserved boxes: the
// ROI at the sweep's start plus `margin` pixels, widened for peaks that
// escaped in an earlier attempt of this sweep.
__global__ void k_box_from_peaks(Sched sc, const Peaks pk, const Frame* frames, Model m,
                                 const int4* override_box) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int p = sc.op_peak[i];
  const Frame fr = frames[pk.frame[p]];
  int4 b;
  b.x = max(0, pk.xi[p] - m.margin);
  b.y = max(0, pk.yi[p] - m.margin);
  b.z = min(fr.W, pk.xi[p] + m.roi + m.margin);
  b.w = min(fr.H, pk.yi[p] + m.roi + m.margin);
  if (override_box != nullptr) {
    int4 o = override_box[p];
    if (o.z > o.x) {
      b.x = max(0, min(b.x, o.x));
      b.y = max(0, min(b.y, o.y));
      b.z = min(fr.W, max(b.z, o.z));
      b.w = min(fr.H, max(b.w, o.w));
    }
  }
  sc.box[i] = b;
  // back up the peak state for a possible rollback
  for (int q = 0; q < kNFit; q++) pk.bk_prm[(long long)q * pk.cap + p] = pk.prm[(long long)q * pk.cap + p];
  pk.bk_lam[p] = pk.lam[p];
  pk.bk_err[p] = pk.err[p];
  pk.bk_xi[p] = pk.xi[p];
  pk.bk_yi[p] = pk.yi[p];
  pk.bk_status[p] = pk.status[p];
  pk.bk_iter[p] = pk.iter[p];
  pk.bk_added[p] = pk.added[p];
  sc.op_flags[i] = 0;
}

// Box = the ROI at the peak's current origin (new peaks, subtraction).
__global__ void k_box_roi(Sched sc, const Peaks pk, const Frame* frames, Model m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int p = sc.op_peak[i];
  const Frame fr = frames[pk.frame[p]];
  int4 b;
  b.x = max(0, min(fr.W, pk.xi[p]));
  b.y = max(0, min(fr.H, pk.yi[p]));
  b.z = max(b.x, min(fr.W, pk.xi[p] + m.roi));
  b.w = max(b.y, min(fr.H, pk.yi[p] + m.roi));
  sc.box[i] = b;
  sc.op_flags[i] = 0;
}

// Box for a new peak from its start row: floor of the start position.
__global__ void k_box_newpeak(Sched sc, const Peaks pk, const Frame* frames, Model m, const double* rows,
                              int fmt, int first_slot) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int p = sc.op_peak[i];
  const Frame fr = frames[pk.frame[p]];
  int row = i;
  const double* q = rows + (fmt == PF_TEXT ? 7LL : 4LL) * row;
  int xi = (int)floor(q[0] - m.xoff);
  int yi = (int)floor(q[1] - m.xoff);
  int4 b;
  b.x = max(0, min(fr.W, xi));
  b.y = max(0, min(fr.H, yi));
  b.z = max(b.x, min(fr.W, xi + m.roi));
  b.w = max(b.y, min(fr.H, yi + m.roi));
  sc.box[i] = b;
  sc.op_flags[i] = 0;
}

// Recenter candidates: box covers the old and the new ROI origin.
__global__ void k_box_recenter(Sched sc, const Peaks pk, const Frame* frames, Model m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int p = sc.op_peak[i];
  const Frame fr = frames[pk.frame[p]];
  int nx = (int)floor(pk.prm[(long long)P_XC * pk.cap + p]);
  int ny = (int)floor(pk.prm[(long long)P_YC * pk.cap + p]);
  int x0 = min(pk.xi[p], nx), y0 = min(pk.yi[p], ny);
  int x1 = max(pk.xi[p], nx) + m.roi, y1 = max(pk.yi[p], ny) + m.roi;
  int4 b;
  b.x = max(0, min(fr.W, x0));
  b.y = max(0, min(fr.H, y0));
  b.z = max(b.x, min(fr.W, x1));
  b.w = max(b.y, min(fr.H, y1));
  sc.box[i] = b;
  sc.op_flags[i] = 0;
}

// Per-frame counter totals of the ops that completed in good frames.
__global__ void k_commit_counters(Sched sc, const Peaks pk, int* frame_counters) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int fi = pk.frame[sc.op_peak[i]];
  if (sc.frame_bad[fi] || (sc.op_flags[i] & 1) == 0) return;
  for (int q = 0; q < C_COUNT; q++) {
    int v = sc.op_delta[(long long)i * C_COUNT + q];
    if (v) atomicAdd(&frame_counters[fi * C_COUNT + q], v);
  }
}

// Escaped ops leave the box they need for the next attempt.
__global__ void k_collect_escapes(Sched sc, const Peaks pk, int4* override_box) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  if (sc.op_flags[i] & 2) override_box[sc.op_peak[i]] = sc.esc_box[i];
}

// Per-peak derived properties: bg_sum, fg_sum, sum (multi_fit.c:1518-1632).
template <int PPL>
__global__ __launch_bounds__(256) void k_peak_sums(Model m, Pix px, Peaks pk, const Frame* frames,
                                                   const int* roi_tab, int slot0, int n, int what,
                                                   double* out) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= n) return;
  const int p = slot0 + i;
  const Frame fr = frames[pk.frame[p]];
  PeakState w;
  load_state(pk, p, w);
  Roi<PPL> c;
  c.init(m, roi_tab, px, fr);
  double t, e;
  lane_shape(m, w, t, e);
  double ps = 0.0, a = 0.0;
#pragma unroll
  for (int s = 0; s < PPL; s++) {
    double ex = __shfl(e, c.rx(s), kWave);
    double ey = __shfl(e, 32 + c.ry(s), kWave);
    if (c.ok(s)) {
      double psf = ey * ex;
      ps += psf;
      int gx = w.xi + c.rx(s), gy = w.yi + c.ry(s);
      if (what == 0 && gx >= 0 && gy >= 0 && gx < fr.W && gy < fr.H) {
        long long k = fr.pix0 + (long long)gy * fr.W + gx;
        double b = (px.x[k] - px.scmos[k]) / px.rqe[k] - w.prm[P_HEIGHT] * psf;
        a += b * psf * psf;
      } else if (what == 1) {
        a += psf * psf;
      }
    }
  }
  ps = wsum(ps);
  a = wsum(a);
  if (lane_id() == 0) {
    double v;
    if (what == 0) {
      v = ((a > 0.0) && (ps > 0.0)) ? a / (ps * ps) : 1.0;
    } else if (what == 1) {
      a = w.prm[P_HEIGHT] * a;
      v = ps > 0.0 ? a / ps : 1.0;
    } else {
      v = ps * w.prm[P_HEIGHT];
    }
    out[i] = v;
  }
}

// "jacobian" property: the 'b' vector at each peak's current state.
template <int PPL, int SHAPE>
__global__ __launch_bounds__(256) void k_peak_jacobian(Model m, Pix px, Peaks pk, const Frame* frames,
                                                       const int* roi_tab, int slot0, int n, double* out) {
  constexpr int N = SHAPE == SH_2DFIXED ? 4 : (SHAPE == SH_3D ? 6 : 5);
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= n) return;
  const int p = slot0 + i;
  const Frame fr = frames[pk.frame[p]];
  PeakState w;
  load_state(pk, p, w);
  if (!origin_ok(m, fr, w.xi, w.yi)) {
    if (lane_id() == 0)
      for (int q = 0; q < N; q++) out[(long long)i * N + q] = 0.0;
    return;
  }
  Roi<PPL> c;
  c.init(m, roi_tab, px, fr);
  c.load(px, fr, w.xi, w.yi);
  double J[N], H[N * (N + 1) / 2];
  calc_jh<PPL, SHAPE>(m, c, w, px.asx, fr, J, H);
  if (lane_id() == 0)
    for (int q = 0; q < N; q++) out[(long long)i * N + q] = J[q];
}

// Gather peak slots: dst slot t takes src slot map[t] (stable compaction).
__global__ void k_gather_peaks(Peaks dst, Peaks src, const int* map, int slot0, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int s = slot0 + map[t], d = slot0 + t;
  for (int q = 0; q < kNFit; q++) dst.prm[(long long)q * dst.cap + d] = src.prm[(long long)q * src.cap + s];
  dst.lam[d] = src.lam[s];
  dst.err[d] = src.err[s];
  dst.xi[d] = src.xi[s];
  dst.yi[d] = src.yi[s];
  dst.status[d] = src.status[s];
  dst.iter[d] = src.iter[s];
  dst.added[d] = src.added[s];
  dst.frame[d] = src.frame[s];
}

__global__ void k_set_status(Peaks pk, int slot0, int n, const int* status) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  pk.status[slot0 + t] = status[t];
}


__global__ void k_flag(Peaks pk, const Frame* frames, long long nslots, int mode, const int* frame_bad,
                       int* flag) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  int fi = pk.frame[t];
  int sel = 0;
  if (fi >= 0) {
    const Frame fr = frames[fi];
    bool live = (t - fr.peak0) < fr.nfit;
    if (live && pk.status[t] == ST_RUNNING) {
      if (mode == SEL_RUNNING) {
        sel = 1;
      } else if (mode == SEL_RUNNING_BAD) {
        sel = frame_bad[fi] ? 1 : 0;
      } else {
        // mFitRecenterPeaks test (multi_fit.c:1658-1661)
        double dx = pk.prm[(long long)P_XC * pk.cap + t] - pk.xi[t];
        double dy = pk.prm[(long long)P_YC * pk.cap + t] - pk.yi[t];
        sel = ((dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25)) ? 1 : 0;
      }
    }
  }
  flag[t] = sel;
}

// Order-preserving compaction of flagged slots into out[], count in *count
// (one block; each thread owns a contiguous chunk).
__global__ void k_compact(const int* flag, long long n, int* out, int* count) {
  __shared__ int part[1024];
  int t = threadIdx.x;
  long long per = (n + blockDim.x - 1) / blockDim.x;
  long long a = t * per, b = a + per < n ? a + per : n;
  int s = 0;
  for (long long i = a; i < b; i++) s += flag[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < (int)blockDim.x; o <<= 1) {
    int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (long long i = a; i < b; i++)
    if (flag[i]) out[run++] = (int)i;
  if (t == blockDim.x - 1) *count = part[t];
}

__global__ void k_fill_i32(int* p, long long n, int v) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ void k_iota(int* p, int n, int base) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = base + t;
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------

static int ppl_for(int nroi) {
  int p = (nroi + kWave - 1) / kWave;
  if (p <= 2) return 2;
  if (p <= 4) return 4;
  if (p <= 6) return 6;
  if (p <= 8) return 8;
  if (p <= 12) return 12;
  return 16;
}

#define DAO_PPL_SWITCH(nroi, KCALL)           \
  switch (ppl_for(nroi)) {                    \
    case 2: { constexpr int PPL = 2; KCALL; } break;   \
    case 4: { constexpr int PPL = 4; KCALL; } break;   \
    case 6: { constexpr int PPL = 6; KCALL; } break;   \
    case 8: { constexpr int PPL = 8; KCALL; } break;   \
    case 12: { constexpr int PPL = 12; KCALL; } break; \
    default: { constexpr int PPL = 16; KCALL; } break; \
  }

#define DAO_SHAPE_SWITCH(shape, KCALL)                        \
  switch (shape) {                                            \
    case SH_2DFIXED: { constexpr int SHP = SH_2DFIXED; KCALL; } break; \
    case SH_2D: { constexpr int SHP = SH_2D; KCALL; } break;           \
    case SH_3D: { constexpr int SHP = SH_3D; KCALL; } break;           \
    default: { constexpr int SHP = SH_Z; KCALL; } break;               \
  }

static inline dim3 waves_grid(int n) { return dim3((n + kWavesPerBlock - 1) / kWavesPerBlock); }
static inline dim3 flat_grid(long long n) { return dim3((unsigned)((n + 255) / 256)); }

void launch_op_round(int kind, const OpArgs& a, hipStream_t st, const double* rows, int fmt, int first_slot) {
  if (a.sc.nops == 0) return;
  dim3 g = waves_grid(a.sc.nops), b(256);
  switch (kind) {
    case OP_LM:
      DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, (k_lm<PPL, SHP><<<g, b, 0, st>>>(a))));
      break;
    case OP_RECENTER:
      DAO_PPL_SWITCH(a.m.nroi, (k_recenter<PPL><<<g, b, 0, st>>>(a)));
      break;
    case OP_NEWPEAK:
      DAO_PPL_SWITCH(a.m.nroi, (k_newpeak<PPL><<<g, b, 0, st>>>(a, rows, fmt, first_slot)));
      break;
    case OP_SUBTRACT:
      DAO_PPL_SWITCH(a.m.nroi, (k_subtract<PPL><<<g, b, 0, st>>>(a)));
      break;
    default:
      DAO_PPL_SWITCH(a.m.nroi, (k_peakerror<PPL><<<g, b, 0, st>>>(a)));
      break;
  }
}

void launch_rollback(const OpArgs& a, hipStream_t st) {
  if (a.sc.nops == 0) return;
  DAO_PPL_SWITCH(a.m.nroi, (k_rollback<PPL><<<waves_grid(a.sc.nops), 256, 0, st>>>(a)));
}

void launch_binning(const Sched& sc, const Frame* frames, const Peaks& pk, int ncells, hipStream_t st) {
  if (sc.nops == 0) return;
  // cell_fill counts ops per cell here, is zeroed by the scan, and counts
  // again as the scatter's cursor: clear it first
  k_fill_i32<<<flat_grid(ncells), 256, 0, st>>>(sc.cell_fill, ncells, 0);
  k_bin_count<<<flat_grid(sc.nops), 256, 0, st>>>(sc, frames, pk);
  k_scan_cells<<<1, 1024, 0, st>>>(sc, ncells);
  k_bin_scatter<<<flat_grid(sc.nops), 256, 0, st>>>(sc, frames, pk);
  k_preds<<<waves_grid(sc.nops), 256, 0, st>>>(sc, frames, pk);
}

void launch_box(int kind, const Sched& sc, const Peaks& pk, const Frame* frames, const Model& m,
                const int4* override_box, const double* rows, int fmt, int first_slot, hipStream_t st) {
  if (sc.nops == 0) return;
  dim3 g = flat_grid(sc.nops);
  switch (kind) {
    case OP_LM: k_box_from_peaks<<<g, 256, 0, st>>>(sc, pk, frames, m, override_box); break;
    case OP_RECENTER: k_box_recenter<<<g, 256, 0, st>>>(sc, pk, frames, m); break;
    case OP_NEWPEAK: k_box_newpeak<<<g, 256, 0, st>>>(sc, pk, frames, m, rows, fmt, first_slot); break;
    default: k_box_roi<<<g, 256, 0, st>>>(sc, pk, frames, m); break;
  }
}

void launch_commit(const Sched& sc, const Peaks& pk, int* frame_counters, hipStream_t st) {
  if (sc.nops == 0) return;
  k_commit_counters<<<flat_grid(sc.nops), 256, 0, st>>>(sc, pk, frame_counters);
}

void launch_collect_escapes(const Sched& sc, const Peaks& pk, int4* override_box, hipStream_t st) {
  if (sc.nops == 0) return;
  k_collect_escapes<<<flat_grid(sc.nops), 256, 0, st>>>(sc, pk, override_box);
}

void launch_new_image(const Pix& px, long long pix0, long long n, const double* img, hipStream_t st) {
  k_new_image<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, img);
}
void launch_new_background(const Pix& px, long long pix0, long long n, const double* bg, hipStream_t st) {
  k_new_background<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, bg);
}
void launch_anscombe(const Pix& px, long long pix0, long long n, hipStream_t st) {
  k_anscombe<<<flat_grid(n), 256, 0, st>>>(px, pix0, n);
}
void launch_fit_image(const Pix& px, long long pix0, long long n, double* out, int residual, hipStream_t st) {
  k_fit_image<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, out, residual);
}
void launch_peak_sums(const Model& m, const Pix& px, const Peaks& pk, const Frame* frames, const int* roi_tab,
                      int slot0, int n, int what, double* out, hipStream_t st) {
  if (n == 0) return;
  DAO_PPL_SWITCH(m.nroi, (k_peak_sums<PPL><<<waves_grid(n), 256, 0, st>>>(m, px, pk, frames, roi_tab, slot0,
                                                                         n, what, out)));
}
void launch_peak_jacobian(const Model& m, const Pix& px, const Peaks& pk, const Frame* frames,
                          const int* roi_tab, int slot0, int n, double* out, hipStream_t st) {
  if (n == 0) return;
  DAO_PPL_SWITCH(m.nroi, DAO_SHAPE_SWITCH(m.shape, (k_peak_jacobian<PPL, SHP><<<waves_grid(n), 256, 0, st>>>(
                                                       m, px, pk, frames, roi_tab, slot0, n, out))));
}
void launch_gather_peaks(const Peaks& dst, const Peaks& src, const int* map, int slot0, int n, hipStream_t st) {
  if (n == 0) return;
  k_gather_peaks<<<flat_grid(n), 256, 0, st>>>(dst, src, map, slot0, n);
}
void launch_set_status(const Peaks& pk, int slot0, int n, const int* status, hipStream_t st) {
  if (n == 0) return;
  k_set_status<<<flat_grid(n), 256, 0, st>>>(pk, slot0, n, status);
}

void launch_select(const Peaks& pk, const Frame* frames, long long nslots, int mode, const int* frame_bad,
                   int* flag, int* out, int* count, hipStream_t st) {
  k_flag<<<flat_grid(nslots), 256, 0, st>>>(pk, frames, nslots, mode, frame_bad, flag);
  k_compact<<<1, 1024, 0, st>>>(flag, nslots, out, count);
}
void launch_fill_i32(int* p, long long n, int v, hipStream_t st) {
  if (n > 0) k_fill_i32<<<flat_grid(n), 256, 0, st>>>(p, n, v);
}
void launch_iota(int* p, int n, int base, hipStream_t st) {
  if (n > 0) k_iota<<<flat_grid(n), 256, 0, st>>>(p, n, base);
}

// Pack the live peaks of every frame into rows of 9 doubles
// {x, y, z, height, background, xsigma, ysigma, error, iterations} + status,
// frame f's peaks starting at row offset[f].
__global__ void k_pack_peaks(Peaks pk, const Frame* frames, const int* offset, Model m, long long nslots,
                             double* out, int* status) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  int fi = pk.frame[t];
  if (fi < 0) return;
  const Frame fr = frames[fi];
  long long j = t - fr.peak0;
  if (j < 0 || j >= fr.nfit) return;
  long long r = offset[fi] + j;
  double* o = out + 9 * r;
  o[0] = pk.prm[(long long)P_XC * pk.cap + t] + m.xoff;
  o[1] = pk.prm[(long long)P_YC * pk.cap + t] + m.xoff;
  o[2] = pk.prm[(long long)P_ZC * pk.cap + t] + m.zoff;
  o[3] = pk.prm[(long long)P_HEIGHT * pk.cap + t];
  o[4] = pk.prm[(long long)P_BG * pk.cap + t];
  o[5] = sqrt(1.0 / (2.0 * pk.prm[(long long)P_XW * pk.cap + t]));
  o[6] = sqrt(1.0 / (2.0 * pk.prm[(long long)P_YW * pk.cap + t]));
  o[7] = pk.err[t];
  o[8] = (double)pk.iter[t];
  status[r] = pk.status[t];
}

void launch_pack_peaks(const Peaks& pk, const Frame* frames, const int* offset, const Model& m, long long nslots,
                       double* out, int* status, hipStream_t st) {
  k_pack_peaks<<<flat_grid(nslots), 256, 0, st>>>(pk, frames, offset, m, nslots, out, status);
}

constexpr int kFpwWaves = 8;

void launch_fit_frames(const FitArgs& a, int nframes, hipStream_t st) {
  if (nframes == 0) return;
  DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, (k_fit_frames<PPL, SHP, kFpwWaves>
                                                         <<<nframes, kFpwWaves * 64, 0, st>>>(a))));
}

}  // namespace dao
