/*
 * dao_finder_amd.h -- local-maxima peak finder of libdao_fit.so.
 *
 * The reference finds new peaks with findLocalMaxima
 * (storm_analysis/sa_library/ia_utilities.c:105-220), driven by
 * MaximaFinder.findMaxima (storm_analysis/sa_library/ia_utilities_c.py:132-252):
 * every pixel inside the margin that is above threshold, not yet taken, and a
 * maximum over a cylinder (radius in x/y, z_range planes) becomes a peak, in
 * (plane, row, column) order, and its taken counter is incremented.  Ties are
 * broken as the reference does: a neighbour at or before the pixel in both x
 * and y must be strictly greater to veto it, any other neighbour greater or
 * equal.  Here the search runs on the GPU, one workgroup per image row, and the
 * peaks are written in the reference's order by a stable compaction, so the
 * peak list, the heights and the taken counters are the reference's exactly.
 *
 * Return 0 on success, -1 for bad sizes, radius or pointers, -2 for a device
 * error.  Calls are synchronous.
 */
#ifndef DAO_FINDER_AMD_H
#define DAO_FINDER_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's finder state (ia_utilities.c:19-38, ia_utilities_c.py:30-44),
 * field for field: images[z] and taken[z] are host arrays of ysize x xsize. */
typedef struct flmData {
  int margin;
  int n_peaks;
  int z_range;

  int xsize;
  int ysize;
  int zsize;

  double radius;
  double threshold;

  double *z_values;

  int32_t **taken;
  double **images;
} flmData;

/* calcMaxPeaks (ia_utilities.c:52-75): pixels inside the margin above
 * threshold and not taken (an upper bound on the peaks findLocalMaxima returns). */
int calcMaxPeaks(flmData *flm_data);

/* findLocalMaxima (ia_utilities.c:105-160): at most flm_data->n_peaks peaks into
 * z, y, x, h (host arrays); flm_data->n_peaks is set to the number found and
 * the taken counters of the found pixels are incremented.  On an error
 * n_peaks is set to 0 and nothing else changes. */
void findLocalMaxima(flmData *flm_data, double *z, double *y, double *x, double *h);

/* The same search on device arrays: images and taken are zsize x ysize x
 * xsize (contiguous, device memory), z_values zsize host doubles; up to
 * max_peaks peaks go to the device arrays z, y, x, h and their count to
 * *n_found (host).  No counterpart in the reference. */
int daoFindLocalMaxima(const double *images, int32_t *taken, int zsize, int ysize, int xsize, int margin,
                       int z_range, double radius, double threshold, const double *z_values, int max_peaks,
                       double *z, double *y, double *x, double *h, int *n_found);

/* daoFindLocalMaxima that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same search and results as daoFindLocalMaxima, with:
 * z_values (zsize doubles) in device memory; counts (2 ints, device memory)
 * receiving {n_found, maxima before the max_peaks cut}; the work queued on
 * `stream` (NULL: the null stream) and not waited for.  The scratch is a
 * fixed device array of 17 MiB that each GPU holds until the process ends:
 * calls with more than 2^24 pixels (zsize * ysize * xsize), 65536 rows
 * (zsize * ysize) or 4096 planes are refused with -1 and a message.  Every
 * call on one GPU uses that GPU's scratch: calls that overlap on one GPU
 * (other streams, other threads) race.  0 on success, -1 bad arguments, -2
 * device error. */
int daoFindLocalMaximaNoAlloc(const double *images, int32_t *taken, int zsize, int ysize, int xsize, int margin,
                              int z_range, double radius, double threshold, const double *z_values, int max_peaks,
                              double *z, double *y, double *x, double *h, int *counts, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DAO_FINDER_AMD_H */
