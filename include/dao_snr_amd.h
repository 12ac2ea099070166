/*
 * dao_snr_amd.h -- the peak finder's per-pixel arithmetic, in libdao_fit.so.
 *
 * PeakFinderGaussian.peakFinder (storm_analysis/sa_library/fitting.py:444-501)
 * turns the image, the background estimate and the current fit image into the
 * signal-to-noise image that findLocalMaxima searches, around two matched-filter
 * convolutions:
 *   bg_var     = background + fit (+ camera_variance)        -> convolved (variance filter)
 *   foreground = image - background - fit                    -> convolved (foreground filter)
 *   bg_var[bg_var <= 0.1] = 0.1;  snr = foreground / sqrt(bg_var) * peak_mask
 * daoFinderPrepare and daoFinderSNR are the two halves, over n frames in
 * device memory (n x npix float64), with the reference's operations in the
 * reference's order (all IEEE-exact: +, -, *, /, sqrt), so the results are
 * bit-identical to numpy's on the same inputs.  Return 0, -1 for bad
 * arguments, -2 for a device error.  Synchronous.  No counterpart in the
 * reference's C (it is numpy there).
 */
#ifndef DAO_SNR_AMD_H
#define DAO_SNR_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* bg_var = background + fit (+ camera_variance, npix values shared by all
 * frames, may be NULL); foreground = image - background - fit. */
int daoFinderPrepare(const double *image, const double *background, const double *fit,
                     const double *camera_variance, int n, long long npix, double *bg_var, double *foreground);

/* snr = foreground / sqrt(max-clamped bg_var) * peak_mask (npix values shared
 * by all frames, may be NULL for none); values of bg_var <= 0.1 are set to 0.1. */
int daoFinderSNR(const double *foreground, const double *bg_var, const double *peak_mask, int n, long long npix,
                 double *snr);

#ifdef __cplusplus
}
#endif

#endif /* DAO_SNR_AMD_H */
