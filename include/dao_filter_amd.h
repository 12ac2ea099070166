/*
 * dao_filter_amd.h -- post-fit peak filters of libdao_fit.so.
 *
 * After each doFit the reference's fitting loop (storm_analysis/sa_library/fitting.py:718-782)
 * downloads the peaks' positions and status and runs three neighbour searches
 * on the host, each on a kd-tree (storm_analysis/sa_library/ia_utilities.c:221-430):
 * runningIfHasNeighbors before new peaks are added, then markDimmerPeaks and
 * markLowSignificancePeaks.  Here they run on the GPU with a per-frame grid
 * hash, so a batch of frames' peak lists never leaves device memory.
 *
 * Status values are the fitter's: RUNNING 0, CONVERGED 1, ERROR 2
 * (storm_analysis/sa_library/multi_fit.h:33-35).
 *
 * Results are the reference's: a peak is a neighbour when its squared distance,
 * (x_k - x_i)^2 + (y_k - y_i)^2 evaluated without fused multiply-add, is <= r^2
 * (the kd-tree's test, storm_analysis/dbscan/kdtree.c:269-273), and the final
 * status of every peak does not depend on the order the reference visits them
 * (a peak only ever moves CONVERGED -> RUNNING or to ERROR, and only a peak
 * itself is set to ERROR).  Positions must be finite and radii positive: for
 * those inputs the kd-tree returns every peak within the radius, except that its
 * pruning test (kdtree.c:283, |dx| < range) can skip a peak whose squared
 * distance is exactly r^2 (which one depends on the tree's shape); the grid
 * hash here always includes it.
 *
 * Return 0 (or the removed count) on success, -1 for bad sizes or null
 * pointers, -2 for a device error.  Calls are synchronous.
 */
#ifndef DAO_FILTER_AMD_H
#define DAO_FILTER_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Drop-ins for libia_utilities (host arrays, one peak list) ---- */

/* markDimmerPeaks (ia_utilities.c:221-284): every non-ERROR peak with a
 * brighter peak (h greater) within r_removal becomes ERROR; CONVERGED peaks
 * within r_neighbors of a removed peak become RUNNING.  Returns the number
 * removed, or -1 / -2 on error. */
int markDimmerPeaks(double *x, double *y, double *h, int32_t *status, double r_removal, double r_neighbors, int np);

/* markLowSignificancePeaks (ia_utilities.c:287-339): every non-ERROR peak with
 * sig <= min_sig becomes ERROR; CONVERGED peaks within r_neighbors of a
 * removed peak become RUNNING.  Returns the number removed, or -1 / -2. */
int markLowSignificancePeaks(double *x, double *y, double *sig, int32_t *status, double min_sig, double r_neighbors,
                             int np);

/* runningIfHasNeighbors (ia_utilities.c:394-427): every current peak that is
 * neither RUNNING nor ERROR and has a new peak within radius becomes RUNNING.
 * The reference returns nothing; errors are therefore silent here too. */
void runningIfHasNeighbors(double *c_x, double *c_y, double *n_x, double *n_y, int32_t *status, double radius,
                           int nc, int nn);

/* ---- Batched forms (device arrays; no counterpart in the reference) ----
 * A batch holds the peak lists of nframes frames back to back; frame f owns
 * peaks offsets[f] .. offsets[f+1]-1 (offsets: nframes + 1 ints in device
 * memory, non-decreasing).  Each frame is filtered on its own, as one call of
 * the function above per frame would.  removed (device, nframes ints) receives
 * each frame's removed count; status is updated in place. */
int daoBatchMarkDimmerPeaks(const double *x, const double *y, const double *h, int32_t *status, const int *offsets,
                            int nframes, double r_removal, double r_neighbors, int *removed);

int daoBatchMarkLowSignificancePeaks(const double *x, const double *y, const double *sig, int32_t *status,
                                     const int *offsets, int nframes, double min_sig, double r_neighbors,
                                     int *removed);

/* Current peaks c_offsets, new peaks n_offsets (both nframes + 1 device ints). */
int daoBatchRunningIfHasNeighbors(const double *c_x, const double *c_y, const int *c_offsets, const double *n_x,
                                  const double *n_y, const int *n_offsets, int32_t *status, int nframes,
                                  double radius);

#ifdef __cplusplus
}
#endif

#endif /* DAO_FILTER_AMD_H */
