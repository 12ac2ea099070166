/*
 * mp_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle of the multiplane DAO fitter: a plain-C, FP64, single-threaded
 * restatement of storm_analysis/multi_plane/mp_fit.c (mpIterateLM :180-487,
 * mpCheckError :44-67, mpResetWorkingPeaks :496-516, mpUpdate :604-690,
 * mpWeightIndex / mpWeightInterpolate :697-741) and
 * storm_analysis/multi_plane/mp_fit_dao.c (mpDaoInitialize2DChannel :32-71,
 * mpDaoNewPeaks :82-153, mpDaoUpdate :173-209), built on the single-channel
 * oracle (dao_oracle.c, compiled into the same library).  Only tests/ load
 * it; the product library never links or calls it.
 *
 * It exports the reference's entry points (mpInitialize, mpDaoInitialize2DChannel,
 * mpSetTransforms, mpSetWeights, mpSetWeightsIndexing, mpDaoNewPeaks,
 * mpIterateLM, mpCleanup) and an mpFit handle whose leading fields have the
 * reference's layout (mp_fit.h:15-61, read by mp_fit_c.py:15-52), so the
 * same Python surface drives it and the HIP library.
 *
 * Quirks of the reference kept on purpose:
 *   - the tolerance the Python passes is the single-channel one times the
 *     number of channels (mp_fit_c.py:266);
 *   - heights are clamped at 0.01 instead of failing (mp_fit_dao.c:184-186),
 *     and the clamped heights weight the channels' x / y updates;
 *   - only XWIDTH is updated; YWIDTH keeps its start value (mp_fit_dao.c:193-204),
 *     while the 2D Jacobian uses XWIDTH for both axes;
 *   - the width clamp runs through daoCheck2D before the positions move, so
 *     its margin test sees the old ROI origin (mp_fit_dao.c:203);
 *   - a failed solve stops the channel loop: later channels' cycle counts are
 *     not incremented in that cycle (mp_fit.c:243-269);
 *   - lambda tests read channel 0's lambda (mp_fit.c:234, 365);
 *   - every channel recentres on its own, then groups with any ERROR channel
 *     are subtracted and set to ERROR everywhere (mp_fit.c:459-475).
 */

#include "dao_oracle.c"

typedef struct mpFit {
  int im_size_x;
  int im_size_y;
  int independent_heights;
  int n_channels;
  int n_weights;

  double tolerance;
  double w_z_offset;
  double w_z_scale;
  double zmin;
  double zmax;

  double *xt_0toN;
  double *yt_0toN;
  double *xt_Nto0;
  double *yt_Nto0;

  double *w_bg;
  double *w_h;
  double *w_x;
  double *w_y;
  double *w_z;
  double *heights;

  double **jacobian;
  double **w_jacobian;
  double **hessian;
  double **w_hessian;

  fitData **fit_data;

  void *fn_slots[5];
} mpFit;

/* mpInitialize (mp_fit.c:147-178). */
mpFit *mpInitialize(double tolerance, int n_channels, int independent_heights, int im_size_x, int im_size_y)
{
  mpFit *mp = (mpFit *)calloc(1, sizeof(mpFit));
  mp->im_size_x = im_size_x;
  mp->im_size_y = im_size_y;
  mp->independent_heights = independent_heights;
  mp->n_channels = n_channels;
  mp->tolerance = tolerance;
  mp->xt_0toN = (double *)calloc(3 * n_channels, sizeof(double));
  mp->yt_0toN = (double *)calloc(3 * n_channels, sizeof(double));
  mp->xt_Nto0 = (double *)calloc(3 * n_channels, sizeof(double));
  mp->yt_Nto0 = (double *)calloc(3 * n_channels, sizeof(double));
  mp->fit_data = (fitData **)calloc(n_channels, sizeof(fitData *));
  mp->jacobian = (double **)calloc(n_channels, sizeof(double *));
  mp->w_jacobian = (double **)calloc(n_channels, sizeof(double *));
  mp->hessian = (double **)calloc(n_channels, sizeof(double *));
  mp->w_hessian = (double **)calloc(n_channels, sizeof(double *));
  return mp;
}

/* mpDaoInitialize2DChannel (mp_fit_dao.c:32-71). */
void mpDaoInitialize2DChannel(mpFit *mp, double *rqe, double *variance, double width_min, double width_max,
                              int roi_size, int channel)
{
  fitData *fd = daoInitialize(rqe, variance, mp->tolerance, mp->im_size_x, mp->im_size_y, roi_size);
  daoInitialize2D(fd, width_min, width_max);
  fd->minimum_height = 1.0;
  mp->fit_data[channel] = fd;
  mp->jacobian[channel] = (double *)calloc(fd->jac_size, sizeof(double));
  mp->w_jacobian[channel] = (double *)calloc(fd->jac_size, sizeof(double));
  mp->hessian[channel] = (double *)calloc(fd->jac_size * fd->jac_size, sizeof(double));
  mp->w_hessian[channel] = (double *)calloc(fd->jac_size * fd->jac_size, sizeof(double));
}

/* mpSetTransforms (mp_fit.c:527-540): by channel, then coefficient. */
void mpSetTransforms(mpFit *mp, double *xt_0toN, double *yt_0toN, double *xt_Nto0, double *yt_Nto0)
{
  size_t b = sizeof(double) * 3 * mp->n_channels;
  memcpy(mp->xt_0toN, xt_0toN, b);
  memcpy(mp->yt_0toN, yt_0toN, b);
  memcpy(mp->xt_Nto0, xt_Nto0, b);
  memcpy(mp->yt_Nto0, yt_Nto0, b);
}

/* mpSetWeights (mp_fit.c:553-582): indexed by z, then channel. */
void mpSetWeights(mpFit *mp, double *w_bg, double *w_h, double *w_x, double *w_y, double *w_z, int z_size)
{
  size_t n = (size_t)mp->n_channels * z_size;
  mp->n_weights = z_size;
  double **dst[5] = {&mp->w_bg, &mp->w_h, &mp->w_x, &mp->w_y, &mp->w_z};
  double *src[5] = {w_bg, w_h, w_x, w_y, w_z};
  for (int q = 0; q < 5; q++) {
    free(*dst[q]);
    *dst[q] = (double *)malloc(sizeof(double) * n);
    memcpy(*dst[q], src[q], sizeof(double) * n);
  }
  free(mp->heights);
  mp->heights = (double *)malloc(sizeof(double) * mp->n_channels);
  for (int c = 0; c < mp->n_channels; c++) mp->heights[c] = 1.0;
}

/* mpSetWeightsIndexing (mp_fit.c:590-594). */
void mpSetWeightsIndexing(mpFit *mp, double z_offset, double z_scale)
{
  mp->w_z_offset = z_offset;
  mp->w_z_scale = z_scale;
}

/* Remove group i from every channel where it is not yet ERROR, then set it
 * to ERROR everywhere, if any channel has it in ERROR (mpCheckError). */
static void check_group(mpFit *mp, int i)
{
  int bad = 0;
  for (int c = 0; c < mp->n_channels; c++)
    if (mp->fit_data[c]->fit[i].status == ST_ERROR) bad = 1;
  if (!bad) return;
  for (int c = 0; c < mp->n_channels; c++) {
    fitData *fd = mp->fit_data[c];
    if (fd->fit[i].status != ST_ERROR) {
      peak_copy(fd, &fd->fit[i], fd->working_peak);
      peak_stamp(fd, -1);
      peak_copy(fd, fd->working_peak, &fd->fit[i]);
    }
    fd->fit[i].status = ST_ERROR;
  }
}

/* mpDaoNewPeaks (mp_fit_dao.c:82-153): channel 0 takes the rows as given,
 * channel c > 0 their x, y mapped by the 0 -> c transforms. */
void mpDaoNewPeaks(mpFit *mp, double *pp, char *p_type, int n_peaks)
{
  int w = (!strcmp(p_type, "finder") || !strcmp(p_type, "testing")) ? 4 : 7;
  int start = mp->fit_data[0]->nfit;
  double *mapped = (double *)malloc(sizeof(double) * (size_t)(n_peaks > 0 ? n_peaks : 1) * w);
  for (int c = 0; c < mp->n_channels; c++) {
    if (c == 0) {
      daoNewPeaks(mp->fit_data[0], pp, p_type, n_peaks);
      continue;
    }
    const double *xt = &mp->xt_0toN[3 * c], *yt = &mp->yt_0toN[3 * c];
    for (int j = 0; j < n_peaks; j++) {
      const double *q = &pp[w * j];
      double *o = &mapped[w * j];
      for (int k = 0; k < w; k++) o[k] = q[k];
      o[0] = xt[0] + q[0] * xt[1] + q[1] * xt[2];
      o[1] = yt[0] + q[0] * yt[1] + q[1] * yt[2];
    }
    daoNewPeaks(mp->fit_data[c], mapped, p_type, n_peaks);
  }
  free(mapped);
  /* the new groups are the last n_peaks.  The reference checks
   * [old nfit, old nfit + n): after a regrowth that dropped k ERROR peaks
   * that misses the first k new groups (and reads k past the live ones);
   * the new groups are checked here */
  (void)start;
  int stop = mp->fit_data[0]->nfit;
  for (int i = (stop - n_peaks > 0 ? stop - n_peaks : 0); i < stop; i++) check_group(mp, i);
}

/* mpWeightIndex (mp_fit.c:697-724). */
static void weight_index(const mpFit *mp, double *dz, int *zi)
{
  const double s = mp->w_z_scale * (mp->fit_data[0]->working_peak->params[P_ZC] - mp->w_z_offset);
  *zi = (int)s;
  *dz = s - (double)*zi;
  if (*zi < 0) *zi = 0;
  if (*zi > mp->n_weights - 2) *zi = mp->n_weights - 2;
}

/* mpWeightInterpolate (mp_fit.c:731-741). */
static double weight_at(const double *wts, double dz, int zi, int nc, int c)
{
  double w1 = wts[nc * zi + c];
  double w2 = wts[nc * (zi + 1) + c];
  return w1 + (w2 - w1) * dz;
}

/* mpDaoUpdate + mpUpdate (mp_fit_dao.c:173-209, mp_fit.c:604-690): the
 * deltas of every channel are in w_jacobian. */
static void mp_update(mpFit *mp)
{
  int nc = mp->n_channels;
  for (int c = 0; c < nc; c++) {
    OPeak *p = mp->fit_data[c]->working_peak;
    p->params[P_HEIGHT] -= mp->w_jacobian[c][0];
    if (p->params[P_HEIGHT] < 0.01) p->params[P_HEIGHT] = 0.01;
    mp->heights[c] = p->params[P_HEIGHT];
  }
  for (int c = 0; c < nc; c++) {
    mp->fit_data[c]->working_peak->params[P_XW] -= mp->w_jacobian[c][3];
    (void)peak_check(mp->fit_data[c]); /* daoCheck2D: margin test, then the width clamp */
  }

  double dz;
  int zi;
  weight_index(mp, &dz, &zi);
  OPeak *p0 = mp->fit_data[0]->working_peak;
  double xoff = mp->fit_data[0]->xoff, yoff = mp->fit_data[0]->yoff;
  for (int axis = 0; axis < 2; axis++) {
    const double *tr = axis == 0 ? mp->xt_Nto0 : mp->yt_Nto0;
    const double *wts = axis == 0 ? mp->w_x : mp->w_y;
    double sum = 0.0, norm = 0.0;
    for (int c = 0; c < nc; c++) {
      double w = weight_at(wts, dz, zi, nc, c);
      double d = tr[3 * c + 1] * mp->w_jacobian[c][1];
      d += tr[3 * c + 2] * mp->w_jacobian[c][2];
      sum += d * w * mp->heights[c];
      norm += w * mp->heights[c];
    }
    p0->params[axis == 0 ? P_XC : P_YC] -= sum / norm;
  }
  for (int c = 1; c < nc; c++) {
    OPeak *p = mp->fit_data[c]->working_peak;
    double t = mp->xt_0toN[3 * c];
    t += mp->xt_0toN[3 * c + 1] * (p0->params[P_XC] + xoff);
    t += mp->xt_0toN[3 * c + 2] * (p0->params[P_YC] + yoff);
    p->params[P_XC] = t - xoff;
    t = mp->yt_0toN[3 * c];
    t += mp->yt_0toN[3 * c + 1] * (p0->params[P_XC] + xoff);
    t += mp->yt_0toN[3 * c + 2] * (p0->params[P_YC] + yoff);
    p->params[P_YC] = t - yoff;
  }
  for (int c = 0; c < nc; c++) roi_follow(mp->fit_data[c]->working_peak);
  for (int c = 0; c < nc; c++) mp->fit_data[c]->working_peak->params[P_BG] -= mp->w_jacobian[c][4];
}

/* mpResetWorkingPeaks (mp_fit.c:496-516). */
static void reset_all(mpFit *mp, int i)
{
  for (int c = 0; c < mp->n_channels; c++) {
    fitData *fd = mp->fit_data[c];
    peak_reset(fd, i);
    fd->working_peak->status = ST_ERROR;
  }
}

/* mpIterateLM (mp_fit.c:180-487). */
void mpIterateLM(mpFit *mp)
{
  int nc = mp->n_channels;
  int nfit = mp->fit_data[0]->nfit;
  for (int i = 0; i < nfit; i++) {
    if (mp->fit_data[0]->fit[i].status != ST_RUNNING) continue;
    double start = 0.0;
    for (int c = 0; c < nc; c++) {
      fitData *fd = mp->fit_data[c];
      peak_copy(fd, &fd->fit[i], fd->working_peak);
      peak_error(fd);
      start += fd->working_peak->error;
      peak_jh(fd, mp->jacobian[c], mp->hessian[c]);
      peak_stamp(fd, -1);
    }
    for (;;) {
      if (mp->fit_data[0]->working_peak->lambda > LM_MAX) {
        for (int c = 0; c < nc; c++) {
          mp->fit_data[c]->n_lost++;
          mp->fit_data[c]->working_peak->status = ST_ERROR;
        }
        break;
      }
      for (int c = 0; c < nc; c++) mp->fit_data[c]->working_peak->status = ST_RUNNING;

      int bad = 0;
      for (int c = 0; c < nc && !bad; c++) {
        fitData *fd = mp->fit_data[c];
        int n = fd->jac_size;
        fd->working_peak->iterations++;
        fd->n_iterations++;
        for (int a = 0; a < n; a++) {
          mp->w_jacobian[c][a] = mp->jacobian[c][a];
          for (int b = 0; b < n; b++)
            mp->w_hessian[c][a * n + b] =
                (a == b) ? (1.0 + fd->working_peak->lambda) * mp->hessian[c][a * n + b] : mp->hessian[c][a * n + b];
        }
        if (chol_solve(mp->w_hessian[c], mp->w_jacobian[c], n) != 0) {
          fd->n_dposv++;
          bad = 1;
        }
      }
      if (bad) {
        for (int c = 0; c < nc; c++) {
          OPeak *w = mp->fit_data[c]->working_peak;
          w->status = ST_ERROR;
          w->lambda = w->lambda * LM_UP;
        }
        continue;
      }

      mp_update(mp);
      for (int c = 0; c < nc; c++)
        if (peak_check(mp->fit_data[c])) bad = 1;
      if (bad) {
        reset_all(mp, i);
        continue;
      }

      for (int c = 0; c < nc; c++) {
        peak_shape(mp->fit_data[c]);
        peak_stamp(mp->fit_data[c], +1);
      }
      double cur = 0.0;
      for (int c = 0; c < nc; c++) {
        if (peak_error(mp->fit_data[c]))
          bad = 1;
        else
          cur += mp->fit_data[c]->working_peak->error;
      }
      if (bad) {
        for (int c = 0; c < nc; c++) peak_stamp(mp->fit_data[c], -1);
        reset_all(mp, i);
        continue;
      }

      if (cur > start) {
        if (((cur - start) / start) < mp->tolerance) {
          for (int c = 0; c < nc; c++) mp->fit_data[c]->working_peak->status = ST_CONVERGED;
          break;
        }
        for (int c = 0; c < nc; c++) {
          mp->fit_data[c]->n_non_decr++;
          peak_stamp(mp->fit_data[c], -1);
        }
        reset_all(mp, i);
        continue;
      }
      if (((start - cur) / start) < mp->tolerance) {
        for (int c = 0; c < nc; c++) mp->fit_data[c]->working_peak->status = ST_CONVERGED;
      } else if (mp->fit_data[0]->working_peak->lambda > LM_MIN) {
        for (int c = 0; c < nc; c++) {
          OPeak *w = mp->fit_data[c]->working_peak;
          w->lambda = w->lambda * LM_DOWN;
        }
      }
      break;
    }
    int status = mp->fit_data[0]->working_peak->status;
    for (int c = 0; c < nc; c++) {
      fitData *fd = mp->fit_data[c];
      fd->working_peak->status = status;
      peak_copy(fd, fd->working_peak, &fd->fit[i]);
    }
  }
  for (int c = 0; c < nc; c++) recenter_peaks(mp->fit_data[c]);
  for (int i = 0; i < nfit; i++) check_group(mp, i);
}

/* mpCleanup (mp_fit.c:73-108). */
void mpCleanup(mpFit *mp)
{
  for (int c = 0; c < mp->n_channels; c++) {
    daoCleanup(mp->fit_data[c]);
    free(mp->jacobian[c]);
    free(mp->w_jacobian[c]);
    free(mp->hessian[c]);
    free(mp->w_hessian[c]);
  }
  free(mp->xt_0toN);
  free(mp->yt_0toN);
  free(mp->xt_Nto0);
  free(mp->yt_Nto0);
  free(mp->w_bg);
  free(mp->w_h);
  free(mp->w_x);
  free(mp->w_y);
  free(mp->w_z);
  free(mp->heights);
  free(mp->jacobian);
  free(mp->w_jacobian);
  free(mp->hessian);
  free(mp->w_hessian);
  free(mp->fit_data);
  free(mp);
}
