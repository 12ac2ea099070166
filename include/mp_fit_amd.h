/*
 * mp_fit_amd.h -- C-ABI of the multiplane DAO fitter (MI355X, gfx950), in
 * libdao_fit.so and loadable as libmp_fit_dao.so (a library that holds
 * nothing of its own and resolves every symbol from libdao_fit.so), as the
 * reference's Python loads it (storm_analysis/multi_plane/mp_fit_dao_c.py:17-36,
 * mp_fit_c.py:55-127).
 *
 * Every function replaces the reference function of the same name in
 * storm_analysis/multi_plane/mp_fit.c and mp_fit_dao.c, with the same
 * arguments.  The per-channel fitData handles (mpFit.fit_data[c]) work with
 * the mFit* entry points of dao_fit_amd.h (properties, fit images, status,
 * compaction), as in the reference.  As there, there are no return codes; a
 * failed call never aborts the process and daoGetLastError() reports it.
 * The fit runs on the GPU: one wavefront per multiplane frame runs its
 * peak groups in the reference's index order.
 */
#ifndef MP_FIT_AMD_H
#define MP_FIT_AMD_H

#include "dao_fit_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Handle.  The leading fields have the layout of the reference's mpFit
 * (storm_analysis/multi_plane/mp_fit.h:15-61), which the Python binding
 * reads (mp_fit_c.py:15-52): fit_data[c] is channel c's fitData handle. */
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

  void *jacobian;
  void *w_jacobian;
  void *hessian;
  void *w_hessian;

  fitData **fit_data;

  void *fn_slots[5];
} mpFit;

/* mpInitialize (mp_fit.c:147-178).  n_channels is at most 8.  NULL (and
 * daoGetLastError set) without a GPU or with bad sizes. */
mpFit *mpInitialize(double tolerance, int n_channels, int independent_heights, int im_size_x, int im_size_y);
/* mpDaoInitialize2DChannel (mp_fit_dao.c:32-71): channel c fits 2D Gaussians
 * (MLE) with the given calibration (copied) and width range; every channel
 * uses the ROI size of the first one initialised. */
void mpDaoInitialize2DChannel(mpFit *mp, double *rqe, double *variance, double width_min, double width_max,
                              int roi_size, int channel);
/* mpSetTransforms (mp_fit.c:527-540): 3 coefficients per channel. */
void mpSetTransforms(mpFit *mp, double *xt_0toN, double *yt_0toN, double *xt_Nto0, double *yt_Nto0);
/* mpSetWeights (mp_fit.c:553-582): z_size x n_channels each. */
void mpSetWeights(mpFit *mp, double *w_bg, double *w_h, double *w_x, double *w_y, double *w_z, int z_size);
/* mpSetWeightsIndexing (mp_fit.c:590-594). */
void mpSetWeightsIndexing(mpFit *mp, double z_offset, double z_scale);
/* mpDaoNewPeaks (mp_fit_dao.c:82-153): n_peaks rows for channel 0, mapped
 * to the other channels; groups in ERROR in some channel are removed from
 * all. */
void mpDaoNewPeaks(mpFit *mp, double *peak_params, char *p_type, int n_peaks);
/* mpIterateLM (mp_fit.c:180-487): one sweep, then recentering and the ERROR
 * synchronisation. */
void mpIterateLM(mpFit *mp);
/* MultiFitter.doFit for the multiplane fitter (dao_fit_c.py:332-356 driving
 * mpIterateLM): one sweep, then more while channel 0 has RUNNING peaks, at
 * most max_iterations more.  No counterpart in the reference's C library.
 * Returns the sweeps run (-1 on failure). */
int mpDoFit(mpFit *mp, int max_iterations);
/* mpCleanup (mp_fit.c:73-108). */
void mpCleanup(mpFit *mp);

/*
 * Batched multiplane fitting.  No counterpart in the reference, which fits
 * one frame-set (a frame of every channel) at a time: here n_sets frame-sets
 * of one size share one device engine and one launch fits them all, one
 * wavefront per frame-set, with mpIterateLM's arithmetic (the same results
 * as the single-set handle).  Channel c of set s is engine frame s * C + c;
 * arrays that hold every frame are in that order.  Calls return 0 (mpBatchFit:
 * the largest sweep count of a set) or a negative code with
 * daoGetLastError() set.
 */
typedef struct mpBatch mpBatch;
/* 2D Gaussian MLE fitters, ROI roi_size, tolerance as mpInitialize's (the
 * single-channel one times the channel count), peak_cap groups per set. */
mpBatch *mpBatchCreate(int n_sets, int n_channels, int im_size_x, int im_size_y, int roi_size, double tolerance,
                       int peak_cap);
void mpBatchDestroy(mpBatch *b);
/* mpDaoInitialize2DChannel for channel `channel` of every set. */
int mpBatchInitializeChannel(mpBatch *b, int channel, const double *rqe, const double *variance, double width_min,
                             double width_max);
/* mpSetTransforms. */
int mpBatchSetTransforms(mpBatch *b, const double *xt_0toN, const double *yt_0toN, const double *xt_Nto0,
                         const double *yt_Nto0);
/* The x / y weights of mpSetWeights (z_size x n_channels each) and mpSetWeightsIndexing. */
int mpBatchSetWeights(mpBatch *b, const double *w_x, const double *w_y, int z_size, double z_offset, double z_scale);
/* mFitNewImage / mFitNewBackground of every frame (host, or device with on_device). */
int mpBatchNewImages(mpBatch *b, const double *images, int on_device);
int mpBatchNewBackgrounds(mpBatch *b, const double *backgrounds, int on_device);
/* mpDaoNewPeaks of every set: counts[s] channel-0 rows for set s, sets in order. */
int mpBatchNewPeaks(mpBatch *b, const double *rows, const int *counts, const char *p_type, int on_device);
/* MultiFitter.doFit driving mpIterateLM, for every set. */
int mpBatchFit(mpBatch *b, int max_iterations);
/* Groups held by set `set` (-1 if out of range). */
int mpBatchNFit(mpBatch *b, int set);
/* Every frame's peaks in frame order: rows of 9 doubles (x, y, z, height,
 * background, xsigma, ysigma, error, iterations) and status. */
int mpBatchGetAllPeaks(mpBatch *b, double *out, int32_t *status, int on_device);
/* mFitRemoveErrorPeaks / mFitRemoveRunningPeaks of every frame. */
int mpBatchRemoveErrorPeaks(mpBatch *b);
int mpBatchRemoveRunningPeaks(mpBatch *b);
/* Counters of one channel of one set since the last call (fitData order, as
 * daoBatchCounters). */
void mpBatchCounters(mpBatch *b, int set, int channel, int32_t *out8);

#ifdef __cplusplus
}
#endif

#endif /* MP_FIT_AMD_H */
