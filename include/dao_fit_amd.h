/*
 * dao_fit_amd.h -- C-ABI of libdao_fit.so, the MI355X (gfx950) build of the
 * 3D-DAOSTORM multi-emitter Levenberg-Marquardt fitter.
 *
 * Every function below replaces the reference function of the same name in
 * storm_analysis/sa_library/{multi_fit.c,dao_fit.c}, with the argument types
 * the reference's ctypes binding declares (storm_analysis/sa_library/
 * dao_fit_c.py:147-238).  Inputs are copied by the callee; outputs go to
 * caller-allocated arrays; as in the reference there are no return codes
 * (per-peak failures are status == ERROR plus the counters in the handle).
 * A failure of the call itself (a HIP runtime error, the scheduler giving
 * up) never aborts the process: the call returns and daoGetLastError()
 * reports it.  Calls are synchronous: each returns when its GPU work has
 * finished.  Handles share no state and may be used from different
 * threads, one call per handle at a time.
 *
 * The batch / serving entry points (daoBatch*) have no counterpart in the
 * reference; they are declared in dao_batch_amd.h.
 */
#ifndef DAO_FIT_AMD_H
#define DAO_FIT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Handle.  The leading fields have the layout of the reference's fitData
 * (storm_analysis/sa_library/multi_fit.h:80-143), which Python reads and
 * writes through ctypes (dao_fit_c.py:22-77): the counters, the sizes, nfit,
 * max_nfit, jac_size and the ROI index tables are kept current on the host
 * after every call.  The image/peak array pointers are NULL here (that state
 * lives on the GPU); roi_x_index / roi_y_index point to host tables.
 */
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

  void *working_peak;
  void *fit;
  void *fit_model;

  void *fn_slots[9];
} fitData;

/* daoInitialize (dao_fit.c:1251-1309 + mFitInitialize multi_fit.c:926-1013 +
 * mFitInitializeROIIndexing multi_fit.c:1031-1060).  Returns NULL when no GPU
 * is available or the sizes are unsupported (roi_size must be 1..32). */
fitData *daoInitialize(double *rqe, double *scmos_cal, double tol, int im_size_x, int im_size_y, int roi_size);

/* daoInitialize2DFixed{,ALS,LS,DWLS,FWLS} (dao_fit.c:1319-1399). */
void daoInitialize2DFixed(fitData *fd);
void daoInitialize2DFixedALS(fitData *fd);
void daoInitialize2DFixedLS(fitData *fd);
void daoInitialize2DFixedDWLS(fitData *fd);
void daoInitialize2DFixedFWLS(fitData *fd);

/* daoInitialize2D{,ALS,LS,DWLS,FWLS} (dao_fit.c:1411-1514). */
void daoInitialize2D(fitData *fd, double width_min, double width_max);
void daoInitialize2DALS(fitData *fd, double width_min, double width_max);
void daoInitialize2DLS(fitData *fd, double width_min, double width_max);
void daoInitialize2DDWLS(fitData *fd, double width_min, double width_max);
void daoInitialize2DFWLS(fitData *fd, double width_min, double width_max);

/* daoInitialize3D (dao_fit.c:1524-1534). */
void daoInitialize3D(fitData *fd, double width_min, double width_max);

/* daoInitializeZ (dao_fit.c:1548-1571): wx/wy = [wo (px), c (um), d (um), A, B]. */
void daoInitializeZ(fitData *fd, double *wx_vs_z, double *wy_vs_z, double z_min, double z_max);

/* daoNewPeaks (dao_fit.c:1582-1766): n x 4 {x,y,z,sigma} for "finder" and
 * "testing", n x 7 {x,y,z,background,height,xsigma,ysigma} for "text" and
 * "hdf5".  Storage grows as mFitNewPeaks does (multi_fit.c:1419-1499). */
void daoNewPeaks(fitData *fd, double *peak_params, char *p_type, int n_peaks);

/* daoCleanup (dao_fit.c:1143-1170). */
void daoCleanup(fitData *fd);

/* mFitNewImage (multi_fit.c:1382-1408). */
void mFitNewImage(fitData *fd, double *image);
/* mFitNewBackground (multi_fit.c:1360-1370). */
void mFitNewBackground(fitData *fd, double *background);
/* mFitAnscombeTransformImage (multi_fit.c:90-97). */
void mFitAnscombeTransformImage(fitData *fd);
/* mFitIterateLM (multi_fit.c:1072-1349): one sweep over the RUNNING peaks in
 * index order, then mFitRecenterPeaks (multi_fit.c:1644-1690). */
void mFitIterateLM(fitData *fd);
/* MultiFitter.doFit (dao_fit_c.py:332-356) in one call: no counterpart in
 * the reference's C library (its Python loops over mFitIterateLM and
 * mFitGetUnconverged).  Same result as that loop; returns the sweeps run. */
int mFitDoFit(fitData *fd, int max_iterations);
/* mFitGetUnconverged (multi_fit.c:898-910). */
int mFitGetUnconverged(fitData *fd);
/* mFitGetNError (multi_fit.c:730-742). */
int mFitGetNError(fitData *fd);
/* mFitGetFitImage (multi_fit.c:708-720): f / rqe, image_size_y x image_size_x. */
void mFitGetFitImage(fitData *fd, double *out);
/* mFitGetResidual (multi_fit.c:880-888): x - f. */
void mFitGetResidual(fitData *fd, double *out);
/* mFitGetPeakPropertyDouble (multi_fit.c:754-839): background, bg_sum, error,
 * fg_sum, fg_sum_sc, height, jacobian (nfit x jac_size), sum, x, xsigma, y,
 * ysigma, z. */
void mFitGetPeakPropertyDouble(fitData *fd, double *values, char *what);
/* mFitGetPeakPropertyInt (multi_fit.c:851-868): iterations, status. */
void mFitGetPeakPropertyInt(fitData *fd, int32_t *values, char *what);
/* mFitRemoveErrorPeaks (multi_fit.c:1704-1735). */
void mFitRemoveErrorPeaks(fitData *fd);
/* mFitRemoveRunningPeaks (multi_fit.c:1750-1779). */
void mFitRemoveRunningPeaks(fitData *fd);
/* mFitSetPeakStatus (multi_fit.c:1809-1831). */
void mFitSetPeakStatus(fitData *fd, int32_t *status);
/* No counterpart in the reference: fit the frame with a team of
 * min(64, ceil(n_peaks / peaks_per_workgroup)) workgroups on as many CUs
 * instead of one (0, the default: one workgroup).  Same results; shorter
 * doFit / mFitIterateLM latency for frames with many peaks. */
void daoSetTeam(fitData *fd, int peaks_per_workgroup);
/* No counterpart in the reference: the code of the calling thread's last
 * failed call of this library (0: none; the codes of dao_batch_amd.h) and,
 * if msg is not NULL, its message (at most len bytes with the terminating
 * zero).  Clears the error. */
int daoGetLastError(char *msg, int len);
/* No counterpart in the reference: polls after which a scheduler wait of
 * the fitter gives up (0: the default).  A wait that gives up fails the
 * call (mFitIterateLM / mFitDoFit) and drops the frame's peaks. */
void daoSetWaitLimit(fitData *fd, long long polls);

#ifdef __cplusplus
}
#endif

#endif /* DAO_FIT_AMD_H */
