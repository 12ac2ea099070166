/*
 * dao_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * This is the oracle: a plain-C, FP64, single-threaded restatement of the
 * 3D-DAOSTORM multi-emitter Levenberg-Marquardt fitter of the reference
 * (storm_analysis/sa_library/multi_fit.c and storm_analysis/sa_library/dao_fit.c).
 * It exists to check the HIP implementation and to be timed as the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * It exports the reference's C-ABI symbol names (daoInitialize, mFitIterateLM,
 * ...) and a fitData handle whose leading fields have the reference's layout
 * (storm_analysis/sa_library/multi_fit.h:80-143), so the same ctypes surface
 * drives either library.
 *
 * How each part is pinned (see tests/test_oracle_*.py):
 *   ROI table            closed form of test_multifit.py:418-431 (test_mfit_11)
 *   peak shape           closed-form Gaussian at pixel centres
 *   error functions      closed forms of the five error models
 *   Jacobian             finite-difference gradient of the error function
 *   Hessian              finite-difference Hessian at a zero-residual image
 *   Cholesky solve       numpy.linalg.solve; singular/indefinite -> failure
 *   width(z)             closed form w = wo*sqrt(1+X^2+A X^3+B X^4)
 *   storage / compaction test_multifit.py test_mfit_1,2,3,6,7 (exact counts)
 *   finder start values  test_mfit_4 (height 10%, background 1e-6)
 *   text/hdf5 round trip test_mfit_8,10 (1e-6)
 *   significance, RQE    test_mfit_9,12
 *   LM state machine     test_mfit_5 (two-peak fit), noise-free convergence to
 *                        the truth for every model x error model, monotone
 *                        accepted errors of an isolated peak, counter
 *                        bookkeeping invariants, and a second, independently
 *                        written pure-Python restatement on tiny frames
 *                        (tests/lm_pure.py, tests/test_oracle_lm.py).
 * The compiled reference is not used: multi_fit.c needs an external LAPACK
 * (dposv_, multi_fit.c:17-18,1851) that is not among the reference's sources,
 * so it is not built here.
 *
 * Quirks of the reference that are kept on purpose:
 *   - the per-pixel t_fi / err_i cache with stale flags (multi_fit.c:111-195):
 *     when an error evaluation stops early on fi <= 0 the pixels after the
 *     failing one keep their cached fi, and the Jacobian reads those;
 *   - the starting error's return value is ignored (multi_fit.c:1122);
 *   - an error increase smaller than tol is CONVERGED with the new params
 *     (multi_fit.c:1259-1283);
 *   - lambda only decreases while above 1e-3 (multi_fit.c:1318-1320);
 *   - 3D and Z use the MLE error (multi_fit.c:1008; dao_fit.c:1524-1571);
 *   - Z stores wo^2 (dao_fit.c:1566-1567) and dwx/dz omits the 1/d factor
 *     (dao_fit.c:904-916);
 *   - the 2D check clamps XWIDTH only; YWIDTH follows at the next update
 *     (dao_fit.c:1061-1088, 1812);
 *   - ERROR peaks are dropped only when peak storage regrows
 *     (multi_fit.c:1429-1477).
 */

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NFIT 7
#define P_HEIGHT 0
#define P_XC 1
#define P_XW 2
#define P_YC 3
#define P_YW 4
#define P_BG 5
#define P_ZC 6

#define ST_RUNNING 0
#define ST_CONVERGED 1
#define ST_ERROR 2

/* multi_fit.h:42-49 */
#define LM_START 1.0
#define LM_DOWN 0.75
#define LM_MAX 1.0e+20
#define LM_MIN 1.0e-3
#define LM_UP 4.0
#define PEAK_BLOCK 500

enum { SH_2DFIXED = 0, SH_2D = 1, SH_3D = 2, SH_Z = 3 };
enum { EM_MLE = 0, EM_ALS = 1, EM_LS = 2, EM_DWLS = 3, EM_FWLS = 4 };

typedef struct OPeak {
  int added;
  int index;
  int iterations;
  int status;
  int xi;
  int yi;
  double error;
  double lambda;
  double params[NFIT];
  double wx_term;
  double wy_term;
  double *psf; /* roi_n_index entries */
  double *xt;  /* roi_size entries each */
  double *ext;
  double *yt;
  double *eyt;
} OPeak;

typedef struct OModel {
  int shape;
  int errm;
  int roi;
  int zfit;
  double min_z, max_z;
  double width_min, width_max;
  double wxz[5];
  double wyz[5];
} OModel;

/* Leading fields follow storm_analysis/sa_library/multi_fit.h:80-143. */
typedef struct fitData {
  int n_dposv;
  int n_iterations;
  int n_lost;
  int n_margin;
  int n_neg_fi;
  int n_neg_height;
  int n_non_converged;
  int n_non_decr;

  int fit_size_x;
  int fit_size_y;
  int image_size_x;
  int image_size_y;
  int jac_size;
  int max_nfit;
  int nfit;
  int roi_n_index;

  double minimum_height;
  double xoff;
  double yoff;
  double zoff;
  double tolerance;

  int *bg_counts;
  int *roi_x_index;
  int *roi_y_index;
  int *stale;

  double *as_xi;
  double *bg_data;
  double *bg_estimate;
  double *err_i;
  double *f_data;
  double *rqe;
  double *scmos_term;
  double *t_fi;
  double *x_data;

  OPeak *working_peak;
  OPeak *fit;
  OModel *fit_model;

  void *fn_slots[9];

  /* oracle-only instrumentation: largest ROI displacement seen in one turn,
   * and a histogram of per-turn displacements (last bin: >= 31 px) */
  int max_turn_shift;
  int turn_shift_hist[32];
  int turn_shift;
} fitData;

/* ------------------------------------------------------------------ */
/* peak storage                                                         */
/* ------------------------------------------------------------------ */

static void peak_arrays(fitData *fd, OPeak *p)
{
  int r = fd->fit_size_x;
  p->psf = (double *)calloc((size_t)r * r, sizeof(double));
  p->xt = (double *)calloc(r, sizeof(double));
  p->ext = (double *)calloc(r, sizeof(double));
  p->yt = (double *)calloc(r, sizeof(double));
  p->eyt = (double *)calloc(r, sizeof(double));
}

static void peak_free(OPeak *p)
{
  free(p->psf);
  free(p->xt);
  free(p->ext);
  free(p->yt);
  free(p->eyt);
}

/* daoCopyPeak (dao_fit.c:1182-1210) + mFitCopyPeak (multi_fit.c:560-575). */
static void peak_copy(const fitData *fd, const OPeak *src, OPeak *dst)
{
  int r = fd->fit_size_x;
  dst->added = src->added;
  dst->index = src->index;
  dst->iterations = src->iterations;
  dst->status = src->status;
  dst->xi = src->xi;
  dst->yi = src->yi;
  dst->error = src->error;
  dst->lambda = src->lambda;
  memcpy(dst->params, src->params, sizeof(double) * NFIT);
  dst->wx_term = src->wx_term;
  dst->wy_term = src->wy_term;
  memcpy(dst->psf, src->psf, sizeof(double) * fd->roi_n_index);
  memcpy(dst->xt, src->xt, sizeof(double) * r);
  memcpy(dst->ext, src->ext, sizeof(double) * r);
  memcpy(dst->yt, src->yt, sizeof(double) * r);
  memcpy(dst->eyt, src->eyt, sizeof(double) * r);
}

/* ------------------------------------------------------------------ */
/* image bookkeeping                                                    */
/* ------------------------------------------------------------------ */

/* mFitAddPeak (multi_fit.c:28-57) with sign=+1, mFitSubtractPeak
 * (multi_fit.c:1894-1922) with sign=-1. */
static void peak_stamp(fitData *fd, int sign)
{
  OPeak *p = fd->working_peak;
  int W = fd->image_size_x;
  int base = p->yi * W + p->xi;
  double h = p->params[P_HEIGHT];
  double b = p->params[P_BG];
  p->added += sign;
  for (int j = 0; j < fd->roi_n_index; j++) {
    int k = base + fd->roi_y_index[j] * W + fd->roi_x_index[j];
    double q = fd->rqe[k];
    if (sign > 0) {
      fd->f_data[k] += h * p->psf[j] * q;
      fd->bg_counts[k] += 1;
      fd->bg_data[k] += b * q + fd->scmos_term[k];
    } else {
      fd->f_data[k] -= h * p->psf[j] * q;
      fd->bg_counts[k] -= 1;
      fd->bg_data[k] -= (b * q + fd->scmos_term[k]);
    }
    fd->stale[k] = 1;
  }
}

/* mFitAnscombe (multi_fit.c:67-78). */
static double anscombe(double v)
{
  if (v < (-0.375 + 1.0e-6)) return 0.0;
  return 2.0 * sqrt(v + 0.375);
}

/* The five error functions (multi_fit.c:111-458), one loop.  Returns 1 when
 * MLE/FWLS meet fi <= 0 (the peak's error is then left untouched). */
static int peak_error(fitData *fd)
{
  OPeak *p = fd->working_peak;
  int W = fd->image_size_x;
  int em = fd->fit_model->errm;
  if (p->status != ST_RUNNING) return 0;
  int base = p->yi * W + p->xi;
  double sum = 0.0;
  for (int j = 0; j < fd->roi_n_index; j++) {
    int k = base + fd->roi_y_index[j] * W + fd->roi_x_index[j];
    if (fd->stale[k] != 0) {
      double fi = fd->f_data[k] + fd->bg_data[k] / ((double)fd->bg_counts[k]);
      double xk = fd->x_data[k];
      double e;
      switch (em) {
      case EM_MLE:
        if (fi <= 0.0) { fd->n_neg_fi++; return 1; }
        fd->t_fi[k] = fi;
        e = 2.0 * ((fi - xk) - xk * log(fi / xk));
        break;
      case EM_ALS: {
        double a = anscombe(fi);
        if (a == 0.0) fd->n_neg_fi++;
        fd->t_fi[k] = a;
        double d = a - fd->as_xi[k];
        e = d * d;
        break;
      }
      case EM_LS: {
        fd->t_fi[k] = fi;
        double d = fi - xk;
        e = d * d;
        break;
      }
      case EM_DWLS: {
        fd->t_fi[k] = fi;
        double d = fi - xk;
        e = d * d / xk;
        break;
      }
      default: { /* EM_FWLS */
        if (fi <= 0.0) { fd->n_neg_fi++; return 1; }
        fd->t_fi[k] = fi;
        double d = fi - xk;
        e = d * d / fi;
        break;
      }
      }
      fd->err_i[k] = e;
      fd->stale[k] = 0;
    }
    sum += fd->err_i[k];
  }
  p->error = sum;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Gaussian model                                                       */
/* ------------------------------------------------------------------ */

/* daoCalcPeakShape (dao_fit.c:968-998).  The y offset uses xoff, as the
 * reference does (dao_fit.c:979); xoff == yoff always. */
static void peak_shape(fitData *fd)
{
  OPeak *p = fd->working_peak;
  int r = fd->fit_size_x;
  double ox = -(p->params[P_XC] - (double)p->xi) - fd->xoff;
  double oy = -(p->params[P_YC] - (double)p->yi) - fd->xoff;
  for (int i = 0; i < r; i++) {
    double t = (double)i + ox;
    p->xt[i] = t;
    p->ext[i] = exp(-t * t * p->params[P_XW]);
  }
  for (int i = 0; i < r; i++) {
    double t = (double)i + oy;
    p->yt[i] = t;
    p->eyt[i] = exp(-t * t * p->params[P_YW]);
  }
  for (int j = 0; j < fd->roi_n_index; j++) {
    p->psf[j] = p->eyt[fd->roi_y_index[j]] * p->ext[fd->roi_x_index[j]];
  }
}

/* daoCalcWidthsFromZ (dao_fit.c:1009-1035); the stored wo is already squared. */
static void widths_from_z(const OModel *m, OPeak *p)
{
  double u = (p->params[P_ZC] - m->wxz[1]) / m->wxz[2];
  double u2 = u * u;
  double s = 1.0 + u2 + m->wxz[3] * u2 * u + m->wxz[4] * u2 * u2;
  p->wx_term = s * s;
  p->params[P_XW] = 2.0 / (m->wxz[0] * s);

  u = (p->params[P_ZC] - m->wyz[1]) / m->wyz[2];
  u2 = u * u;
  s = 1.0 + u2 + m->wyz[3] * u2 * u + m->wyz[4] * u2 * u2;
  p->wy_term = s * s;
  p->params[P_YW] = 2.0 / (m->wyz[0] * s);
}

/* d(width)/dz chain factors of daoCalcJHZ (dao_fit.c:904-916). */
static void z_gradients(const OModel *m, const OPeak *p, double *gx, double *gy)
{
  double u = (p->params[P_ZC] - m->wxz[1]) / m->wxz[2];
  double u2 = u * u;
  double zt = 2.0 * u + 3.0 * m->wxz[3] * u2 + 4.0 * m->wxz[4] * u2 * u;
  *gx = -2.0 * zt / (m->wxz[0] * p->wx_term);
  u = (p->params[P_ZC] - m->wyz[1]) / m->wyz[2];
  u2 = u * u;
  zt = 2.0 * u + 3.0 * m->wyz[3] * u2 + 4.0 * m->wyz[4] * u2 * u;
  *gy = -2.0 * zt / (m->wyz[0] * p->wy_term);
}

/* daoCalcJH* (dao_fit.c:48-958): derivative of the model at each ROI pixel,
 * then the error model's weights.  J and H are jac_size and jac_size^2,
 * H filled in its upper triangle (row-major). */
static void peak_jh(fitData *fd, double *J, double *H)
{
  OPeak *p = fd->working_peak;
  OModel *m = fd->fit_model;
  int n = fd->jac_size;
  int W = fd->image_size_x;
  int base = p->yi * W + p->xi;
  double a1 = p->params[P_HEIGHT];
  double wx = p->params[P_XW];
  double wy = (m->shape == SH_3D || m->shape == SH_Z) ? p->params[P_YW] : wx;
  double gx = 0.0, gy = 0.0;
  if (m->shape == SH_Z) z_gradients(m, p, &gx, &gy);

  for (int l = 0; l < n; l++) J[l] = 0.0;
  for (int l = 0; l < n * n; l++) H[l] = 0.0;

  for (int j = 0; j < fd->roi_n_index; j++) {
    int k = base + fd->roi_y_index[j] * W + fd->roi_x_index[j];
    double fi = fd->t_fi[k];
    double xk = fd->x_data[k];
    double xt = p->xt[fd->roi_x_index[j]];
    double ex = p->ext[fd->roi_x_index[j]];
    double yt = p->yt[fd->roi_y_index[j]];
    double ey = p->eyt[fd->roi_y_index[j]];
    double e = ex * ey;
    double g[6];

    switch (m->shape) {
    case SH_2DFIXED:
      g[0] = e;
      g[1] = 2.0 * a1 * wx * xt * e;
      g[2] = 2.0 * a1 * wy * yt * e;
      g[3] = 1.0;
      break;
    case SH_2D:
      g[0] = e;
      g[1] = 2.0 * a1 * wx * xt * e;
      g[2] = 2.0 * a1 * wy * yt * e;
      g[3] = -a1 * xt * xt * e - a1 * yt * yt * e;
      g[4] = 1.0;
      break;
    case SH_3D:
      g[0] = e;
      g[1] = 2.0 * a1 * wx * xt * e;
      g[2] = -a1 * xt * xt * e;
      g[3] = 2.0 * a1 * wy * yt * e;
      g[4] = -a1 * yt * yt * e;
      g[5] = 1.0;
      break;
    default: /* SH_Z */
      g[0] = e;
      g[1] = 2.0 * a1 * wx * xt * e;
      g[2] = 2.0 * a1 * wy * yt * e;
      g[3] = -a1 * xt * xt * gx * e - a1 * yt * yt * gy * e;
      g[4] = 1.0;
      break;
    }

    /* scale by d(transformed fit)/d(fit): rqe, or rqe*2/as_fi for ALS */
    double s = (m->errm == EM_ALS) ? fd->rqe[k] * 2.0 / fi : fd->rqe[k];
    for (int l = 0; l < n; l++) g[l] *= s;

    double wj, wh;
    switch (m->errm) {
    case EM_MLE:
      wj = 2.0 * (1.0 - xk / fi);
      wh = 2.0 * xk / (fi * fi);
      break;
    case EM_ALS:
      wj = fi - fd->as_xi[k];
      wh = 1.0;
      break;
    case EM_LS:
      wj = fi - xk;
      wh = 1.0;
      break;
    case EM_DWLS:
      wj = (fi - xk) / xk;
      wh = 1.0 / xk;
      break;
    default: { /* EM_FWLS */
      double t1 = (fi - xk) / fi;
      double t2 = 1.0 / fi;
      wj = 2.0 * t1 - t1 * t1;
      wh = 2.0 * t2 * (xk * t2 - t1 + t1 * t1);
      break;
    }
    }
    for (int l = 0; l < n; l++) {
      J[l] += wj * g[l];
      for (int c = l; c < n; c++) H[l * n + c] += wh * g[l] * g[c];
    }
  }
}

/* mFitSolve (multi_fit.c:1839-1883): dposv "Lower" on row-major storage reads
 * the filled upper triangle.  Returns non-zero on a non-positive or NaN pivot
 * (LAPACK dpotrf convention) or a NaN in the solution. */
static int chol_solve(const double *A, double *b, int n)
{
  double L[NFIT][NFIT];
  for (int j = 0; j < n; j++) {
    double s = A[j * n + j];
    for (int k = 0; k < j; k++) s -= L[j][k] * L[j][k];
    if (!(s > 0.0)) return j + 1;
    double d = sqrt(s);
    L[j][j] = d;
    for (int i = j + 1; i < n; i++) {
      double t = A[j * n + i]; /* element (i,j) of the symmetric matrix */
      for (int k = 0; k < j; k++) t -= L[i][k] * L[j][k];
      L[i][j] = t / d;
    }
  }
  for (int i = 0; i < n; i++) {
    double t = b[i];
    for (int k = 0; k < i; k++) t -= L[i][k] * b[k];
    b[i] = t / L[i][i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double t = b[i];
    for (int k = i + 1; k < n; k++) t -= L[k][i] * b[k];
    b[i] = t / L[i][i];
  }
  for (int i = 0; i < n; i++)
    if (isnan(b[i])) return 1;
  return 0;
}

/* mFitUpdate (multi_fit.c:1933-1946): move the ROI with hysteresis. */
static void roi_follow(OPeak *p)
{
  double dx = p->params[P_XC] - p->xi;
  if ((dx < -0.5) || (dx > 1.5)) p->xi = (int)floor(p->params[P_XC]);
  double dy = p->params[P_YC] - p->yi;
  if ((dy < -0.5) || (dy > 1.5)) p->yi = (int)floor(p->params[P_YC]);
}

/* daoUpdate2DFixed/2D/3D/Z (dao_fit.c:1777-1880). */
static void peak_update(fitData *fd, const double *d)
{
  OPeak *p = fd->working_peak;
  OModel *m = fd->fit_model;
  switch (m->shape) {
  case SH_2DFIXED:
    p->params[P_HEIGHT] -= d[0];
    p->params[P_XC] -= d[1];
    p->params[P_YC] -= d[2];
    p->params[P_BG] -= d[3];
    roi_follow(p);
    break;
  case SH_2D:
    p->params[P_HEIGHT] -= d[0];
    p->params[P_XC] -= d[1];
    p->params[P_YC] -= d[2];
    p->params[P_XW] -= d[3];
    p->params[P_BG] -= d[4];
    p->params[P_YW] = p->params[P_XW];
    roi_follow(p);
    break;
  case SH_3D:
    p->params[P_HEIGHT] -= d[0];
    p->params[P_XC] -= d[1];
    p->params[P_XW] -= d[2];
    p->params[P_YC] -= d[3];
    p->params[P_YW] -= d[4];
    p->params[P_BG] -= d[5];
    roi_follow(p);
    break;
  default: /* SH_Z */
    p->params[P_HEIGHT] -= d[0];
    p->params[P_XC] -= d[1];
    p->params[P_YC] -= d[2];
    p->params[P_ZC] -= d[3];
    p->params[P_BG] -= d[4];
    if (p->params[P_ZC] < m->min_z) p->params[P_ZC] = m->min_z;
    if (p->params[P_ZC] > m->max_z) p->params[P_ZC] = m->max_z;
    roi_follow(p);
    widths_from_z(m, p);
    break;
  }
}

/* mFitCheck (multi_fit.c:470-502). */
static int margin_height_check(fitData *fd)
{
  OPeak *p = fd->working_peak;
  if ((p->xi < 0) || (p->xi >= (fd->image_size_x - fd->fit_size_x)) || (p->yi < 0) ||
      (p->yi >= (fd->image_size_y - fd->fit_size_y))) {
    fd->n_margin++;
    return 1;
  }
  if (p->params[P_HEIGHT] < 0.0) {
    fd->n_neg_height++;
    return 1;
  }
  return 0;
}

/* daoCheck2DFixed/2D/3D/Z (dao_fit.c:1047-1133). */
static int peak_check(fitData *fd)
{
  OPeak *p = fd->working_peak;
  OModel *m = fd->fit_model;
  if (margin_height_check(fd)) return 1;
  if (m->shape == SH_2D || m->shape == SH_3D) {
    if (p->params[P_XW] < m->width_min) p->params[P_XW] = m->width_min;
    else if (p->params[P_XW] > m->width_max) p->params[P_XW] = m->width_max;
  }
  if (m->shape == SH_3D) {
    if (p->params[P_YW] < m->width_min) p->params[P_YW] = m->width_min;
    else if (p->params[P_YW] > m->width_max) p->params[P_YW] = m->width_max;
  }
  return 0;
}

/* mFitResetPeak (multi_fit.c:1788-1798). */
static void peak_reset(fitData *fd, int i)
{
  OPeak *w = fd->working_peak;
  int added = w->added;
  double lam = w->lambda;
  peak_copy(fd, &fd->fit[i], w);
  w->added = added;
  w->lambda = lam * LM_UP;
}

static void note_shift(fitData *fd, int xi0, int yi0)
{
  int dx = abs(fd->working_peak->xi - xi0);
  int dy = abs(fd->working_peak->yi - yi0);
  int d = dx > dy ? dx : dy;
  if (d > fd->max_turn_shift) fd->max_turn_shift = d;
  if (d > fd->turn_shift) fd->turn_shift = d;
}

/* mFitRecenterPeaks (multi_fit.c:1644-1690). */
static void recenter_peaks(fitData *fd)
{
  for (int i = 0; i < fd->nfit; i++) {
    OPeak *p = &fd->fit[i];
    if (p->status != ST_RUNNING) continue;
    double dx = p->params[P_XC] - p->xi;
    double dy = p->params[P_YC] - p->yi;
    if ((dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25)) {
      OPeak *w = fd->working_peak;
      peak_copy(fd, p, w);
      peak_stamp(fd, -1);
      w->xi = (int)floor(p->params[P_XC]);
      w->yi = (int)floor(p->params[P_YC]);
      if (margin_height_check(fd)) {
        fd->n_lost += 1;
        w->status = ST_ERROR;
      } else {
        peak_shape(fd);
        peak_stamp(fd, +1);
      }
      peak_copy(fd, w, p);
    }
  }
}

/* ------------------------------------------------------------------ */
/* exported API                                                         */
/* ------------------------------------------------------------------ */

/* mFitIterateLM (multi_fit.c:1072-1349): one Gauss-Seidel sweep. */
void mFitIterateLM(fitData *fd)
{
  int n = fd->jac_size;
  double J[NFIT], H[NFIT * NFIT], wJ[NFIT], wH[NFIT * NFIT];
  OPeak *w = fd->working_peak;

  for (int i = 0; i < fd->nfit; i++) {
    if (fd->fit[i].status != ST_RUNNING) continue;
    peak_copy(fd, &fd->fit[i], w);
    int xi0 = w->xi, yi0 = w->yi;
    fd->turn_shift = 0;

    peak_error(fd); /* return value ignored (multi_fit.c:1116-1123) */
    double start = w->error;
    peak_jh(fd, J, H);
    peak_stamp(fd, -1);

    for (;;) {
      if (w->lambda > LM_MAX) {
        fd->n_lost++;
        w->status = ST_ERROR;
        break;
      }
      w->iterations++;
      fd->n_iterations++;
      w->status = ST_RUNNING;

      for (int a = 0; a < n; a++) {
        wJ[a] = J[a];
        for (int c = 0; c < n; c++)
          wH[a * n + c] = (a == c) ? (1.0 + w->lambda) * H[a * n + c] : H[a * n + c];
      }
      if (chol_solve(wH, wJ, n) != 0) {
        fd->n_dposv++;
        w->status = ST_ERROR;
        w->lambda = w->lambda * LM_UP;
        continue;
      }

      peak_update(fd, wJ);
      if (peak_check(fd)) {
        peak_reset(fd, i);
        w->status = ST_ERROR;
        continue;
      }
      note_shift(fd, xi0, yi0);

      peak_shape(fd);
      peak_stamp(fd, +1);

      if (peak_error(fd)) {
        peak_stamp(fd, -1);
        peak_reset(fd, i);
        w->status = ST_ERROR;
        continue;
      }

      if (w->error > start) {
        if (((w->error - start) / start) < fd->tolerance) {
          w->status = ST_CONVERGED;
          break;
        }
        fd->n_non_decr++;
        peak_stamp(fd, -1);
        peak_reset(fd, i);
      } else {
        if (((start - w->error) / start) < fd->tolerance) {
          w->status = ST_CONVERGED;
          break;
        }
        if (w->lambda > LM_MIN) w->lambda = LM_DOWN * w->lambda;
        break;
      }
    }
    peak_copy(fd, w, &fd->fit[i]);
    fd->turn_shift_hist[fd->turn_shift < 31 ? fd->turn_shift : 31]++;
  }
  recenter_peaks(fd);
}

/* mFitInitialize (multi_fit.c:926-1013) + mFitInitializeROIIndexing
 * (multi_fit.c:1031-1060) + daoInitialize (dao_fit.c:1251-1309). */
fitData *daoInitialize(double *rqe, double *scmos_cal, double tol, int im_size_x, int im_size_y,
                       int roi_size)
{
  size_t npix = (size_t)im_size_x * im_size_y;
  fitData *fd = (fitData *)calloc(1, sizeof(fitData));
  fd->image_size_x = im_size_x;
  fd->image_size_y = im_size_y;
  fd->tolerance = tol;
  fd->minimum_height = -1.0;

  fd->rqe = (double *)malloc(sizeof(double) * npix);
  fd->scmos_term = (double *)malloc(sizeof(double) * npix);
  memcpy(fd->rqe, rqe, sizeof(double) * npix);
  memcpy(fd->scmos_term, scmos_cal, sizeof(double) * npix);
  fd->bg_counts = (int *)calloc(npix, sizeof(int));
  fd->stale = (int *)calloc(npix, sizeof(int));
  fd->as_xi = (double *)calloc(npix, sizeof(double));
  fd->bg_data = (double *)calloc(npix, sizeof(double));
  fd->bg_estimate = (double *)calloc(npix, sizeof(double));
  fd->err_i = (double *)calloc(npix, sizeof(double));
  fd->f_data = (double *)calloc(npix, sizeof(double));
  fd->t_fi = (double *)calloc(npix, sizeof(double));
  fd->x_data = (double *)calloc(npix, sizeof(double));

  /* circular ROI: pixel centres within r/2 of the ROI centre, row-major */
  fd->fit_size_x = roi_size;
  fd->fit_size_y = roi_size;
  fd->roi_x_index = (int *)malloc(sizeof(int) * roi_size * roi_size);
  fd->roi_y_index = (int *)malloc(sizeof(int) * roi_size * roi_size);
  double c = 0.5 * (double)roi_size;
  int cnt = 0;
  for (int i = 0; i < roi_size; i++) {
    double dy = (double)i - c + 0.5;
    for (int j = 0; j < roi_size; j++) {
      double dx = (double)j - c + 0.5;
      if ((dx * dx + dy * dy) <= c * c) {
        fd->roi_x_index[cnt] = j;
        fd->roi_y_index[cnt] = i;
        cnt++;
      }
    }
  }
  fd->roi_n_index = cnt;

  fd->xoff = 0.5 * (double)roi_size - 1.0;
  fd->yoff = 0.5 * (double)roi_size - 1.0;
  fd->zoff = 0.0;

  fd->fit_model = (OModel *)calloc(1, sizeof(OModel));
  fd->fit_model->roi = roi_size;
  fd->working_peak = (OPeak *)calloc(1, sizeof(OPeak));
  peak_arrays(fd, fd->working_peak);
  return fd;
}

static void set_model(fitData *fd, int shape, int errm, int jac)
{
  fd->fit_model->shape = shape;
  fd->fit_model->errm = errm;
  fd->jac_size = jac;
}

/* daoInitialize2DFixed* (dao_fit.c:1319-1399). */
void daoInitialize2DFixed(fitData *fd) { set_model(fd, SH_2DFIXED, EM_MLE, 4); }
void daoInitialize2DFixedALS(fitData *fd) { set_model(fd, SH_2DFIXED, EM_ALS, 4); }
void daoInitialize2DFixedLS(fitData *fd) { set_model(fd, SH_2DFIXED, EM_LS, 4); }
void daoInitialize2DFixedDWLS(fitData *fd) { set_model(fd, SH_2DFIXED, EM_DWLS, 4); }
void daoInitialize2DFixedFWLS(fitData *fd) { set_model(fd, SH_2DFIXED, EM_FWLS, 4); }

/* daoInitialize2D* (dao_fit.c:1411-1514). */
static void init_2d(fitData *fd, int errm, double wmin, double wmax)
{
  set_model(fd, SH_2D, errm, 5);
  fd->fit_model->width_min = wmin;
  fd->fit_model->width_max = wmax;
}
void daoInitialize2D(fitData *fd, double a, double b) { init_2d(fd, EM_MLE, a, b); }
void daoInitialize2DALS(fitData *fd, double a, double b) { init_2d(fd, EM_ALS, a, b); }
void daoInitialize2DLS(fitData *fd, double a, double b) { init_2d(fd, EM_LS, a, b); }
void daoInitialize2DDWLS(fitData *fd, double a, double b) { init_2d(fd, EM_DWLS, a, b); }
void daoInitialize2DFWLS(fitData *fd, double a, double b) { init_2d(fd, EM_FWLS, a, b); }

/* daoInitialize3D (dao_fit.c:1524-1534): keeps the MLE error function. */
void daoInitialize3D(fitData *fd, double wmin, double wmax)
{
  set_model(fd, SH_3D, EM_MLE, 6);
  fd->fit_model->width_min = wmin;
  fd->fit_model->width_max = wmax;
}

/* daoInitializeZ (dao_fit.c:1548-1571). */
void daoInitializeZ(fitData *fd, double *wx, double *wy, double z_min, double z_max)
{
  set_model(fd, SH_Z, EM_MLE, 5);
  OModel *m = fd->fit_model;
  m->zfit = 1;
  for (int i = 0; i < 5; i++) {
    m->wxz[i] = wx[i];
    m->wyz[i] = wy[i];
  }
  m->wxz[0] = m->wxz[0] * m->wxz[0];
  m->wyz[0] = m->wyz[0] * m->wyz[0];
  m->min_z = z_min;
  m->max_z = z_max;
}

/* mFitNewPeaks (multi_fit.c:1419-1499): storage growth drops ERROR peaks. */
static void grow_peaks(fitData *fd, int n_new)
{
  if ((fd->nfit + n_new) > fd->max_nfit) {
    int n_alloc = PEAK_BLOCK * ((fd->nfit + n_new) / PEAK_BLOCK + 1);
    OPeak *np = (OPeak *)calloc(n_alloc, sizeof(OPeak));
    for (int j = 0; j < n_alloc; j++) peak_arrays(fd, &np[j]);
    int kept = 0;
    for (int j = 0; j < fd->nfit; j++) {
      if (fd->fit[j].status != ST_ERROR) {
        peak_copy(fd, &fd->fit[j], &np[kept]);
        kept++;
      }
    }
    if (fd->fit != NULL) {
      for (int j = 0; j < fd->max_nfit; j++) peak_free(&fd->fit[j]);
      free(fd->fit);
    }
    fd->fit = np;
    fd->max_nfit = n_alloc;
    fd->nfit = kept;
  }
  for (int i = fd->nfit; i < fd->nfit + n_new; i++) {
    OPeak *p = &fd->fit[i];
    p->added = 0;
    p->index = i;
    p->iterations = 0;
    p->status = ST_RUNNING;
    p->error = 0.0;
    p->lambda = LM_START;
  }
}

/* mFitEstimatePeakBackground (multi_fit.c:638-662). */
static void estimate_background(fitData *fd)
{
  OPeak *p = fd->working_peak;
  int W = fd->image_size_x;
  int base = p->yi * W + p->xi;
  double s = 0.0;
  for (int j = 0; j < fd->roi_n_index; j++) {
    int k = base + fd->roi_y_index[j] * W + fd->roi_x_index[j];
    s += fd->bg_estimate[k] / fd->rqe[k];
  }
  p->params[P_BG] = s / ((double)fd->roi_n_index);
}

/* mFitEstimatePeakHeight (multi_fit.c:678-700). */
static void estimate_height(fitData *fd)
{
  OPeak *p = fd->working_peak;
  int W = fd->image_size_x;
  int base = p->yi * W + p->xi;
  double sp = 0.0, sx = 0.0;
  for (int j = 0; j < fd->roi_n_index; j++) {
    int k = base + fd->roi_y_index[j] * W + fd->roi_x_index[j];
    double t = p->psf[j] * fd->rqe[k];
    sp += t * t * t;
    sx += t * t * (fd->x_data[k] - fd->f_data[k] - fd->bg_estimate[k]);
  }
  p->params[P_HEIGHT] = sx / sp;
}

/* daoNewPeaks (dao_fit.c:1582-1766). */
void daoNewPeaks(fitData *fd, double *pp, char *p_type, int n_peaks)
{
  OModel *m = fd->fit_model;
  OPeak *w = fd->working_peak;
  int finder = !strcmp(p_type, "finder");
  int four = finder || !strcmp(p_type, "testing");
  int seven = !strcmp(p_type, "text") || !strcmp(p_type, "hdf5");

  grow_peaks(fd, n_peaks);
  int start = fd->nfit;
  int stop = fd->nfit + n_peaks;

  if (four || seven) {
    for (int i = start; i < stop; i++) {
      OPeak *p = &fd->fit[i];
      const double *q = four ? &pp[4 * (i - start)] : &pp[7 * (i - start)];
      p->params[P_XC] = q[0] - fd->xoff;
      p->params[P_YC] = q[1] - fd->yoff;
      p->params[P_ZC] = q[2] - fd->zoff;
      if (four) {
        if (m->zfit) {
          widths_from_z(m, p);
        } else {
          double width = 1.0 / (2.0 * q[3] * q[3]);
          p->params[P_XW] = width;
          p->params[P_YW] = width;
        }
      } else {
        p->params[P_BG] = q[3];
        p->params[P_HEIGHT] = q[4];
        p->params[P_XW] = 1.0 / (2.0 * q[5] * q[5]);
        p->params[P_YW] = 1.0 / (2.0 * q[6] * q[6]);
        if (m->zfit) widths_from_z(m, p);
      }
      p->xi = (int)floor(p->params[P_XC]);
      p->yi = (int)floor(p->params[P_YC]);
      if (four) {
        p->params[P_BG] = 1.0;
        p->params[P_HEIGHT] = 1.0;
      }

      peak_copy(fd, p, w);
      if (peak_check(fd)) {
        w->status = ST_ERROR;
        peak_copy(fd, w, p);
        continue;
      }
      peak_shape(fd);
      if (four) {
        estimate_background(fd);
        if (finder) {
          estimate_height(fd);
          if (w->params[P_HEIGHT] < fd->minimum_height) w->params[P_HEIGHT] = fd->minimum_height;
          if (w->params[P_HEIGHT] <= 0.0) {
            printf("Warning peak %d has negative estimated height!\n", (i - start));
            w->status = ST_ERROR;
            peak_copy(fd, w, p);
            continue;
          }
        }
      }
      peak_stamp(fd, +1);
      peak_copy(fd, w, p);
    }
  } else {
    printf("Unrecognized peak type '%s'!\n", p_type);
  }

  for (int i = start; i < stop; i++) {
    peak_copy(fd, &fd->fit[i], w);
    peak_error(fd);
    peak_copy(fd, w, &fd->fit[i]);
  }
  fd->nfit = stop;
}

void daoCleanup(fitData *fd)
{
  if (fd->fit != NULL) {
    for (int j = 0; j < fd->max_nfit; j++) peak_free(&fd->fit[j]);
    free(fd->fit);
  }
  peak_free(fd->working_peak);
  free(fd->working_peak);
  free(fd->fit_model);
  free(fd->bg_counts);
  free(fd->roi_x_index);
  free(fd->roi_y_index);
  free(fd->stale);
  free(fd->as_xi);
  free(fd->bg_data);
  free(fd->bg_estimate);
  free(fd->err_i);
  free(fd->f_data);
  free(fd->rqe);
  free(fd->scmos_term);
  free(fd->t_fi);
  free(fd->x_data);
  free(fd);
}

/* mFitNewImage (multi_fit.c:1382-1408). */
void mFitNewImage(fitData *fd, double *img)
{
  size_t npix = (size_t)fd->image_size_x * fd->image_size_y;
  for (size_t i = 0; i < npix; i++) {
    fd->x_data[i] = img[i] * fd->rqe[i] + fd->scmos_term[i];
    fd->bg_counts[i] = 0;
    fd->bg_data[i] = 0.0;
    fd->err_i[i] = 0.0;
    fd->f_data[i] = 0.0;
    fd->stale[i] = 1;
  }
  fd->nfit = 0;
}

/* mFitNewBackground (multi_fit.c:1360-1370). */
void mFitNewBackground(fitData *fd, double *bg)
{
  size_t npix = (size_t)fd->image_size_x * fd->image_size_y;
  for (size_t i = 0; i < npix; i++) {
    fd->bg_estimate[i] = bg[i] * fd->rqe[i];
    fd->stale[i] = 1;
  }
}

/* mFitAnscombeTransformImage (multi_fit.c:90-97). */
void mFitAnscombeTransformImage(fitData *fd)
{
  size_t npix = (size_t)fd->image_size_x * fd->image_size_y;
  for (size_t i = 0; i < npix; i++) fd->as_xi[i] = anscombe(fd->x_data[i]);
}

int mFitGetUnconverged(fitData *fd)
{
  int c = 0;
  for (int i = 0; i < fd->nfit; i++)
    if (fd->fit[i].status == ST_RUNNING) c++;
  return c;
}

int mFitGetNError(fitData *fd)
{
  int c = 0;
  for (int i = 0; i < fd->nfit; i++)
    if (fd->fit[i].status == ST_ERROR) c++;
  return c;
}

void mFitGetFitImage(fitData *fd, double *out)
{
  size_t npix = (size_t)fd->image_size_x * fd->image_size_y;
  for (size_t i = 0; i < npix; i++) out[i] = fd->f_data[i] / fd->rqe[i];
}

void mFitGetResidual(fitData *fd, double *out)
{
  size_t npix = (size_t)fd->image_size_x * fd->image_size_y;
  for (size_t i = 0; i < npix; i++) out[i] = fd->x_data[i] - fd->f_data[i];
}

/* mFitPeakBgSum (multi_fit.c:1518-1565). */
static double peak_bg_sum(fitData *fd, const OPeak *p)
{
  int W = fd->image_size_x;
  int base = p->yi * W + p->xi;
  double bs = 0.0, ps = 0.0;
  for (int j = 0; j < fd->roi_n_index; j++) {
    int k = base + fd->roi_y_index[j] * W + fd->roi_x_index[j];
    double psf = p->psf[j];
    ps += psf;
    double b = (fd->x_data[k] - fd->scmos_term[k]) / fd->rqe[k] - p->params[P_HEIGHT] * psf;
    bs += b * psf * psf;
  }
  if ((bs > 0.0) && (ps > 0.0)) return bs / (ps * ps);
  return 1.0;
}

/* mFitPeakFgSum (multi_fit.c:1578-1610). */
static double peak_fg_sum(fitData *fd, const OPeak *p)
{
  double fs = 0.0, ps = 0.0;
  for (int j = 0; j < fd->roi_n_index; j++) {
    ps += p->psf[j];
    fs += p->psf[j] * p->psf[j];
  }
  fs = p->params[P_HEIGHT] * fs;
  if (ps > 0.0) return fs / ps;
  return 1.0;
}

/* mFitPeakSum (multi_fit.c:1620-1632). */
static double peak_sum(fitData *fd, const OPeak *p)
{
  double s = 0.0;
  for (int j = 0; j < fd->roi_n_index; j++) s += p->psf[j];
  return s * p->params[P_HEIGHT];
}

/* mFitGetPeakPropertyDouble (multi_fit.c:754-839).  "jacobian" writes the
 * nfit x jac_size 'b' vectors at the current state (the mounted reference's
 * copy loop at multi_fit.c:801-803 is empty; this writes what it computes). */
void mFitGetPeakPropertyDouble(fitData *fd, double *v, char *what)
{
  int n = fd->nfit;
  if (!strcmp(what, "background")) {
    for (int i = 0; i < n; i++) v[i] = fd->fit[i].params[P_BG];
  } else if (!strcmp(what, "bg_sum")) {
    for (int i = 0; i < n; i++) v[i] = peak_bg_sum(fd, &fd->fit[i]);
  } else if (!strcmp(what, "error")) {
    for (int i = 0; i < n; i++) v[i] = fd->fit[i].error;
  } else if (!strcmp(what, "fg_sum") || !strcmp(what, "fg_sum_sc")) {
    for (int i = 0; i < n; i++) v[i] = peak_fg_sum(fd, &fd->fit[i]);
  } else if (!strcmp(what, "height")) {
    for (int i = 0; i < n; i++) v[i] = fd->fit[i].params[P_HEIGHT];
  } else if (!strcmp(what, "jacobian")) {
    double J[NFIT], H[NFIT * NFIT];
    for (int i = 0; i < n; i++) {
      peak_copy(fd, &fd->fit[i], fd->working_peak);
      peak_jh(fd, J, H);
      for (int k = 0; k < fd->jac_size; k++) v[i * fd->jac_size + k] = J[k];
    }
  } else if (!strcmp(what, "sum")) {
    for (int i = 0; i < n; i++) v[i] = peak_sum(fd, &fd->fit[i]);
  } else if (!strcmp(what, "x")) {
    for (int i = 0; i < n; i++) v[i] = fd->fit[i].params[P_XC] + fd->xoff;
  } else if (!strcmp(what, "xsigma")) {
    for (int i = 0; i < n; i++) v[i] = sqrt(1.0 / (2.0 * fd->fit[i].params[P_XW]));
  } else if (!strcmp(what, "y")) {
    for (int i = 0; i < n; i++) v[i] = fd->fit[i].params[P_YC] + fd->yoff;
  } else if (!strcmp(what, "ysigma")) {
    for (int i = 0; i < n; i++) v[i] = sqrt(1.0 / (2.0 * fd->fit[i].params[P_YW]));
  } else if (!strcmp(what, "z")) {
    for (int i = 0; i < n; i++) v[i] = fd->fit[i].params[P_ZC] + fd->zoff;
  } else {
    printf("Unrecognized parameter '%s'!\n", what);
  }
}

/* mFitGetPeakPropertyInt (multi_fit.c:851-868). */
void mFitGetPeakPropertyInt(fitData *fd, int32_t *v, char *what)
{
  if (!strcmp(what, "iterations")) {
    for (int i = 0; i < fd->nfit; i++) v[i] = fd->fit[i].iterations;
  } else if (!strcmp(what, "status")) {
    for (int i = 0; i < fd->nfit; i++) v[i] = fd->fit[i].status;
  } else {
    printf("Unrecognized parameter '%s'!\n", what);
  }
}

/* mFitRemoveErrorPeaks (multi_fit.c:1704-1735): stable compaction. */
void mFitRemoveErrorPeaks(fitData *fd)
{
  int k = 0;
  for (int j = 0; j < fd->nfit; j++) {
    if (fd->fit[j].status != ST_ERROR) {
      if (j != k) peak_copy(fd, &fd->fit[j], &fd->fit[k]);
      k++;
    }
  }
  fd->nfit = k;
}

/* mFitRemoveRunningPeaks (multi_fit.c:1750-1779). */
void mFitRemoveRunningPeaks(fitData *fd)
{
  int k = 0;
  for (int j = 0; j < fd->nfit; j++) {
    if (fd->fit[j].status != ST_RUNNING) {
      if (j != k) peak_copy(fd, &fd->fit[j], &fd->fit[k]);
      k++;
    } else {
      fd->n_non_converged++;
    }
  }
  fd->nfit = k;
}

/* mFitSetPeakStatus (multi_fit.c:1809-1831). */
void mFitSetPeakStatus(fitData *fd, int32_t *status)
{
  for (int i = 0; i < fd->nfit; i++) {
    if (status[i] == ST_ERROR && fd->fit[i].added > 0) {
      peak_copy(fd, &fd->fit[i], fd->working_peak);
      peak_stamp(fd, -1);
      peak_copy(fd, fd->working_peak, &fd->fit[i]);
    }
    fd->fit[i].status = status[i];
  }
}

/* ------------------------------------------------------------------ */
/* oracle-only helpers for the tests                                    */
/* ------------------------------------------------------------------ */

/* Error and J/H of peak i at its current state without changing it
 * (error written to *err, J to J[jac], H to H[jac*jac]).  Returns the error
 * function's status. */
int oraPeakErrorJH(fitData *fd, int i, double *err, double *J, double *H)
{
  OPeak *w = fd->working_peak;
  peak_copy(fd, &fd->fit[i], w);
  w->status = ST_RUNNING;
  int rc = peak_error(fd);
  *err = w->error;
  peak_jh(fd, J, H);
  return rc;
}

/* Solve (A) x = b with the fitter's Cholesky; A is n x n row-major upper. */
int oraSolve(double *A, double *b, int n) { return chol_solve(A, b, n); }

/* Largest ROI shift (pixels, Chebyshev) seen inside one LM turn so far. */
int oraMaxTurnShift(fitData *fd) { return fd->max_turn_shift; }

/* Histogram of per-turn ROI displacements, 32 bins. */
void oraTurnShiftHist(fitData *fd, int32_t *out)
{
  for (int i = 0; i < 32; i++) out[i] = fd->turn_shift_hist[i];
}
