/*
 * dao_batch_amd.h -- batch entry points of libdao_fit.so.
 *
 * These have no counterpart in the reference: its fitter handles one frame
 * per fitData (storm_analysis/sa_library/dao_fit_c.py:257-549) and callers
 * loop over frames.  A daoBatch holds many frames (of different sizes) on the
 * GPU at once; daoBatchFit runs MultiFitter.doFit (dao_fit_c.py:332-356) for
 * all of them together, each frame with exactly the per-frame semantics of
 * mFitIterateLM (multi_fit.c:1072-1349): frames share no state and the result
 * of a frame does not depend on the other frames in the batch.
 *
 * Frame data are concatenated in frame order, each frame row-major
 * (height x width).  Peak rows use the daoNewPeaks layouts (n x 4 for
 * "finder"/"testing", n x 7 for "text"/"hdf5").  Pointers flagged
 * `on_device` are HIP device pointers (for example torch CUDA tensors).
 * Calls are synchronous, one call per batch at a time.
 *
 * Errors.  No entry point aborts the process.  Functions returning int
 * return 0 (or the count they describe) on success and a negative code on
 * failure: -1 argument out of range, -2 HIP runtime failure, -3 more peaks
 * than a frame's capacity (checked before any work: nothing changes), -4
 * executor made no progress, -5 some frames failed (daoBatchFailedFrames;
 * the other frames' results stand), -6 other.  daoGetLastError
 * (dao_fit_amd.h) returns the calling thread's last code and message.
 */
#ifndef DAO_BATCH_AMD_H
#define DAO_BATCH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct daoBatch daoBatch;

/* Model selection: shape 0 = 2DFixed, 1 = 2D, 2 = 3D, 3 = Z (dao_fit.c:1319-1571);
 * error model 0 = MLE, 1 = ALS, 2 = LS, 3 = DWLS, 4 = FWLS (multi_fit.c:111-458;
 * 3D and Z support MLE only, as in the reference). */
typedef struct daoBatchModel {
  int shape;
  int error_model;
  int roi_size;
  int margin;       /* reserved-box margin of the scheduler, pixels (default 2) */
  double tolerance; /* 1e-6 in the reference (dao_fit_c.py:286) */
  double width_min; /* 2D / 3D width clamp, 1/(2 sigma^2) */
  double width_max;
  double wx[5];     /* Z: [wo (px), c (um), d (um), A, B] */
  double wy[5];
  double z_min;
  double z_max;
  int precision;    /* 0: FP64 throughout (the reference's arithmetic);
                       1: mixed -- FP32 per-pixel model, deviance terms,
                       Jacobian and Hessian, FP64 fit-image accumulators,
                       deviance sums, Cholesky solve and accept/reject
                       (MLE only; other error models run in FP64) */
} daoBatchModel;

/* Serving.  The frame sizes given at creation are the slots' capacities; a
 * slot can then hold any frame of at most that many pixels (and scheduling
 * cells), one at a time. */

/* Put a frame of width x height into slot `frame`: image, background, optional
 * calibration (rqe and scmos_cal together, else rqe = 1 and scmos = 0) and
 * n_rows start peaks in the daoNewPeaks layout.  Replaces what the slot held.
 * Returns 0 (and changes nothing) if the frame or its peaks exceed the slot. */
int daoBatchLoadFrame(daoBatch *b, int frame, int width, int height, const double *image,
                      const double *background, const double *rqe, const double *scmos_cal, const double *rows,
                      int n_rows, const char *p_type, int on_device);
/* Empty slot `frame`; daoBatchFit skips empty slots. */
int daoBatchClearFrame(daoBatch *b, int frame);
/* daoBatchLoadFrame for n slots at once (uniform calibration: rqe = 1,
 * scmos = 0), with one synchronisation: frames[k] receives the
 * widths[k] x heights[k] image and background and n_rows[k] start peaks.
 * Returns 0 (and changes nothing) if any frame does not fit its slot. */
int daoBatchLoadFrames(daoBatch *b, int n, const int *frames, const int *widths, const int *heights,
                       const double *const *images, const double *const *backgrounds, const double *const *rows,
                       const int *n_rows, const char *p_type, int on_device);
/* Empty n slots at once. */
int daoBatchClearFrames(daoBatch *b, int n, const int *frames);

/* Create a batch of n_frames frames; peak_caps[f] bounds the peaks of frame f.
 * Returns NULL without a GPU or with unsupported sizes. */
daoBatch *daoBatchCreate(const daoBatchModel *model, int n_frames, const int *widths, const int *heights,
                         const int *peak_caps);
void daoBatchDestroy(daoBatch *b);

/* Per-frame camera calibration (host arrays; NULL = rqe 1 / scmos 0). */
int daoBatchSetCalibration(daoBatch *b, int frame, const double *rqe, const double *scmos_cal);

/* mFitNewImage / mFitNewBackground (multi_fit.c:1360-1408) for every frame. */
int daoBatchNewImages(daoBatch *b, const double *images, int on_device);
int daoBatchNewBackgrounds(daoBatch *b, const double *backgrounds, int on_device);
/* daoBatchNewImages and daoBatchNewBackgrounds in one pass over the pixels
 * (the same values; no counterpart in the reference, which sets image and
 * background in two calls). */
int daoBatchNewFrames(daoBatch *b, const double *images, const double *backgrounds, int on_device);
/* mFitAnscombeTransformImage for every frame (ALS models). */
int daoBatchAnscombe(daoBatch *b);

/* daoNewPeaks (dao_fit.c:1582-1766) for every frame: counts[f] (n_frames
 * entries) rows for frame f, rows concatenated in frame order.  -3, with
 * nothing changed, if a frame's held peaks plus its new ones exceed its
 * capacity. */
int daoBatchNewPeaks(daoBatch *b, const double *rows, const int *counts, const char *p_type, int on_device);

/* doFit for every frame; returns the number of sweeps run, or -5 when a
 * frame's scheduler wait exceeded its bound: that frame is emptied (its
 * state mid-fit is not the sequential one) and flagged, the other frames
 * are fitted as usual and can be read. */
int daoBatchFit(daoBatch *b, int max_iterations);
/* Frames that failed in the last daoBatchFit: flags[f] = 1 (n_frames
 * entries); returns their number. */
int daoBatchFailedFrames(daoBatch *b, int32_t *flags);
/* Polls after which a scheduler wait of the fitter gives up and fails its
 * frame (0: the default, far beyond any wait of a correct schedule). */
void daoBatchSetWaitLimit(daoBatch *b, long long polls);
/* Teams (daoBatchSetTuning's team > 0) hand a frame's tail to one workgroup
 * once an LM sweep would run at most `ops` peaks (a long tail of sweeps with
 * few RUNNING peaks costs about twice as much through a team); 0: never,
 * < 0: the default (32).  The fits do not depend on it. */
void daoBatchSetTeamHandoff(daoBatch *b, int ops);

/* Per-frame read-back.  out: nfit x 9 doubles per peak
 * {x, y, z, height, background, xsigma, ysigma, error, iterations}. */
int daoBatchNFrames(daoBatch *b);
int daoBatchNFit(daoBatch *b, int frame);
int daoBatchGetPeaks(daoBatch *b, int frame, double *out, int32_t *status);
/* All frames at once, concatenated in frame order (sum of nfit rows): into
 * host arrays, or into device arrays (no host copy). */
int daoBatchGetAllPeaks(daoBatch *b, double *out, int32_t *status);
int daoBatchGetAllPeaksDevice(daoBatch *b, double *out, int32_t *status);
/* Counters of the frame since the last call, fitData order:
 * n_dposv, n_iterations, n_lost, n_margin, n_neg_fi, n_neg_height, n_non_converged, n_non_decr. */
void daoBatchCounters(daoBatch *b, int frame, int32_t *out8);
/* Fit image (f / rqe) of one frame (mFitGetFitImage, multi_fit.c:708-720). */
int daoBatchGetFitImage(daoBatch *b, int frame, double *out);
/* Fit images (residual = 0: f / rqe) or residuals (1: x - f,
 * mFitGetResidual multi_fit.c:880-888) of every frame, concatenated in frame
 * order at the frames' current sizes, into host or device memory. */
int daoBatchGetFitImages(daoBatch *b, double *out, int residual, int on_device);
/* mFitRemoveErrorPeaks (multi_fit.c:1704-1740) for every frame, on the device. */
int daoBatchRemoveErrorPeaks(daoBatch *b);
/* mFitRemoveRunningPeaks (multi_fit.c:1750-1779) for every frame, on the
 * device; each removed peak counts in its frame's n_non_converged. */
int daoBatchRemoveRunningPeaks(daoBatch *b);
/* mFitSetPeakStatus (multi_fit.c:1809-1831) for every frame: status holds
 * one entry per live peak, frames in order (the rows of
 * daoBatchGetAllPeaks); a peak that becomes ERROR while in the fit image is
 * subtracted.  status is a host array, or a device array with on_device. */
int daoBatchSetPeakStatus(daoBatch *b, const int32_t *status, int on_device);
/* The "significance" property (fg_sum / sqrt(bg_sum), dao_fit_c.py:398-439,
 * multi_fit.c:1518-1606) of every live peak, frames in order, into a
 * device array.  No counterpart in the reference's C library. */
int daoBatchGetSignificanceDevice(daoBatch *b, double *out);
/* daoBatchFit limited to the frames whose active[f] (host, one per frame)
 * is non-zero; the others are left as they are.  No counterpart in the
 * reference: a batch runs the reference's per-frame find / fit loop
 * (fitting.py:866-902), in which only some frames fit at each step. */
int daoBatchFitActive(daoBatch *b, int max_iterations, const int32_t *active);

/* Instrumentation: enable HIP-event timing of the LM kernel; stats out[6] =
 * {LM kernel ms, LM launches, sweeps, executor rounds, escape re-runs, max attempts}. */
void daoBatchSetTiming(daoBatch *b, int on);
/* Tuning of the fitter: waves per workgroup (0: automatic -- the kernel's
 * maximum, 8 for ROIs up to 512 pixels, else 4, and 4 when a launch fits more
 * frames than the device has CUs; fewer waves leave room for more frames per
 * CU), the reserved-box margin in pixels (>= 0; -1 keeps it) and the team
 * size: a frame with n peaks is fitted by min(64, ceil(n / team))
 * workgroups on as many CUs (0: one workgroup per frame; -1 keeps it), which
 * shortens the frame's latency.  Results are the same for every setting: a
 * turn that leaves its reserved box is rolled back exactly, from the pixel
 * journal of every turn undone. */
void daoBatchSetTuning(daoBatch *b, int waves, int margin, int team);
void daoBatchStats(daoBatch *b, double *out6);

#ifdef __cplusplus
}
#endif

#endif /* DAO_BATCH_AMD_H */
