/*
 * dao_mfilter_amd.h -- libmatched_filter.so: FFT image convolution on the GPU.
 *
 * A drop-in for the reference's matched filter library
 * (storm_analysis/sa_library/matched_filter.c:60-230, bound by
 * storm_analysis/sa_library/matched_filter_c.py:18-34), which the peak finders
 * use for the background estimate and the foreground / variance smoothing
 * (storm_analysis/sa_library/fitting.py:250-272,444-501).  Same symbol names,
 * signatures and results (to FFT rounding): the caller passes the PSF already
 * recentred (numpy.roll by half its size on each axis, recenter_psf.py), the
 * filter keeps FFT(psf) / (x_size * y_size), and convolve() returns the
 * inverse FFT of FFT(image) * that product, unnormalised, i.e. the circular
 * convolution of the image with the PSF.  Arrays are x_size x y_size float64,
 * y the fast axis.  The transforms run in hipFFT, the spectrum product in a
 * HIP kernel, with the filter's arrays in device memory.
 */
#ifndef DAO_MFILTER_AMD_H
#define DAO_MFILTER_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct daoMatchedFilter filter;

/* initialize (matched_filter.c:165-230): psf is the recentred PSF (host);
 * max_diff > 0 configures convolveMemo.  estimate (FFTW's planning flag) has no
 * effect.  NULL when the device or hipFFT fails. */
filter *initialize(double *psf, double max_diff, int x_size, int y_size, int estimate);

/* convolve (matched_filter.c:93-105): result = image (*) psf, host arrays. */
void convolve(filter *flt, double *image, double *result);

/* convolveMemo (matched_filter.c:118-152): convolve only when some pixel
 * differs from the previous image by more than max_diff, else return the
 * previous result (the PSF itself before the first convolution, as the
 * reference's work buffer holds it). */
void convolveMemo(filter *flt, double *image, double *result);

/* cleanup (matched_filter.c:60-81). */
void cleanup(filter *flt);

/* n images (device, n x x_size x y_size, n <= 65535) convolved into result
 * (device) in one batched transform; 0 on success, -1 for bad arguments, -2 for a device
 * error.  The images are left untouched (the transform runs on a copy in
 * result; result may also be images itself).  No counterpart in the reference. */
int daoMatchedFilterConvolveBatch(filter *flt, const double *images, int n, double *result);

#ifdef __cplusplus
}
#endif

#endif /* DAO_MFILTER_AMD_H */
