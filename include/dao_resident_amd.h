/*
 * dao_resident_amd.h -- the resident fitter of libdao_fit.so: frames admitted
 * into a kernel that is already running.
 *
 * No counterpart in the reference, which fits one frame per MultiFitter call
 * (storm_analysis/sa_library/dao_fit_c.py:257-549, driven frame by frame by
 * fitting.PeakFitter, fitting.py:701-824).  A daoResident serves requests on
 * the slots of a daoBatch (dao_batch_amd.h): a request names a free slot and
 * hands over one frame (image, background, start peaks); workgroups of a
 * resident kernel take requests in submission order as they free up, fit
 * each frame with the per-frame semantics of mFitNewImage, mFitNewBackground,
 * daoNewPeaks and MultiFitter.doFit (multi_fit.c:1360-1408,
 * dao_fit.c:1582-1766, dao_fit_c.py:332-356), and write the peaks into pinned
 * host memory.  Submitting a request is a few host memory writes: no kernel
 * launch and no HIP call per request.  The kernel's workgroups leave once
 * no frame is being fitted and no request has been taken for `idle_seconds`;
 * the next submission (or wait) starts it again.  Frames share no state: a frame's result does not depend on which
 * other frames are on the device with it, and equals daoBatchFit's for the
 * same frame.
 *
 * Limits: the MLE error model; uniform calibration (rqe = 1, scmos = 0); a
 * frame at most the slot's pixel capacity and peak capacity.  While a
 * daoResident exists, its daoBatch must not be used otherwise.
 * daoResidentSubmit, daoResidentWait and daoResidentResult may be called from
 * different threads.
 *
 * Errors as dao_batch_amd.h: negative codes, message via daoGetLastError.
 */
#ifndef DAO_RESIDENT_AMD_H
#define DAO_RESIDENT_AMD_H

#include <stdint.h>

#include "dao_batch_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct daoResident daoResident;

/* A resident fitter over the slots of b: `workgroups` workgroups (0: one per
 * compute unit but 8, which stay free for other kernels -- a resident
 * workgroup holds all of its CU's registers), doFit with at most
 * max_iterations sweeps after the first (2000 in the reference), workgroups
 * leave once idle for idle_seconds.  team_peaks = 0: each request is fitted
 * by one workgroup (the most frames per second); > 0: a request of n start
 * peaks by a team of min(64, ceil(n / team_peaks)) workgroups on as many CUs,
 * fewer when workgroups are already given to requests not yet collected (at
 * least one): lower latency per frame while the device has room; results
 * identical for every team size.  NULL (daoGetLastError) for
 * an error model other than MLE. */
daoResident *daoResidentCreate(daoBatch *b, int workgroups, int max_iterations, double idle_seconds, int team_peaks);
/* Stops the kernel once the outstanding requests are done, frees everything
 * (not b). */
void daoResidentDestroy(daoResident *r);

/* Queue one frame of width x height into free slot `slot`: image and
 * background (row-major, height x width) and n_rows start peaks in the
 * daoNewPeaks layout of p_type.  Host arrays are copied before the call
 * returns; with on_device they are device pointers that must stay valid until
 * the request is done.  *ticket receives the request's number.  -1 if the
 * slot holds a request not yet collected or the frame exceeds the slot. */
int daoResidentSubmit(daoResident *r, int slot, int width, int height, const double *image, const double *background,
                      const double *rows, int n_rows, const char *p_type, int on_device, int64_t *ticket);
/* Wait up to timeout_seconds for requests to finish; writes the slots of at
 * most max_slots finished, not yet reported requests to slots_out and returns
 * their number (0 on timeout). */
int daoResidentWait(daoResident *r, double timeout_seconds, int max_slots, int32_t *slots_out);
/* Results of the finished request in `slot`, which frees the slot: out (n x 9
 * doubles {x, y, z, height, background, xsigma, ysigma, error, iterations}),
 * status (n), counters8 (the frame's fitData counters, daoBatchCounters'
 * order) and info4 {sweeps, device seconds from taking the request to its
 * completion, of them loading the frame and its start peaks, 0}; any output
 * may be NULL.  Returns the number of peaks n,
 * -5 if the frame failed (a scheduler wait gave up, or the request was
 * refused: a pixel not finite and positive, a non-finite background or start
 * peak; no peaks), -1 if the slot holds no finished request. */
int daoResidentResult(daoResident *r, int slot, double *out, int32_t *status, int32_t *counters8, double *info4);
/* Requests submitted so far, and how many times the kernel was launched. */
void daoResidentStats(daoResident *r, int64_t *out2);

#ifdef __cplusplus
}
#endif

#endif /* DAO_RESIDENT_AMD_H */
