"""
Matched-filter oracle -- test infrastructure only.

Used by tests/ (and bench.py's CPU leg) to check the device FFT convolution
(libmatched_filter.so: initialize / convolve / convolveMemo and the batched
form).  A numpy restatement of storm_analysis/sa_library/matched_filter.c:60-230
and recenter_psf.py:9-17:

- the PSF is recentred by numpy.roll of half its size on each axis;
- initialize keeps psf_fft = rfft2(psf) / (x_size * y_size);
- convolve returns FFTW's unnormalised c2r of rfft2(image) * psf_fft, i.e.
  numpy's irfft2 (which divides by x_size * y_size) times x_size * y_size;
- convolveMemo convolves only when some pixel differs from the previous
  image by more than max_diff, otherwise returns the previous result; the
  previous image starts at zero and the previous result at the PSF (the
  reference's work buffer holds the PSF after initialize).
FFTW itself (an external library the reference links) is replaced by numpy's
FFT; both compute the same discrete Fourier transform, so results agree to
FFT rounding.  brute_circular() is the plain circular convolution sum that
pins the restatement.
"""
import numpy


def recenter_psf(psf):
    """recenterPSF (recenter_psf.py:9-17)."""
    out = numpy.roll(psf, int(psf.shape[0] / 2), axis=0)
    return numpy.roll(out, int(psf.shape[1] / 2), axis=1)


class MatchedFilterOracle(object):
    def __init__(self, psf, memoize=False, max_diff=0.1):
        rc = numpy.ascontiguousarray(recenter_psf(psf), dtype=numpy.float64)
        self.shape = rc.shape
        self.psf_fft = numpy.fft.rfft2(rc) * (1.0 / float(rc.shape[0] * rc.shape[1]))
        self.memoize = memoize
        self.max_diff = max_diff
        self.old = numpy.zeros(self.shape)
        self.last = rc.copy()

    def _convolve(self, image):
        spec = numpy.fft.rfft2(image) * self.psf_fft
        return numpy.fft.irfft2(spec, s=self.shape) * float(self.shape[0] * self.shape[1])

    def convolve(self, image):
        image = numpy.asarray(image, dtype=numpy.float64)
        if self.memoize:
            if numpy.any(numpy.abs(image - self.old) > self.max_diff):
                self.old = image.copy()
                self.last = self._convolve(image)
            return self.last.copy()
        self.last = self._convolve(image)
        return self.last.copy()


def brute_circular(image, psf):
    """sum_k image[(i - k) mod X, (j - l) mod Y] * recentred_psf[k, l]."""
    rc = recenter_psf(psf)
    out = numpy.zeros(image.shape)
    for k in range(image.shape[0]):
        for l in range(image.shape[1]):
            if rc[k, l] != 0.0:
                out += rc[k, l] * numpy.roll(numpy.roll(image, k, axis=0), l, axis=1)
    return out


def gaussian_psf(shape, sigma, center=None):
    """A unit-sum Gaussian drawn at the array centre (the finders' PSFs)."""
    x, y = numpy.mgrid[0:shape[0], 0:shape[1]]
    cx, cy = (shape[0] / 2, shape[1] / 2) if center is None else center
    g = numpy.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * sigma * sigma))
    return g / g.sum()


def finder_prepare(image, background, fit, camera_variance=None):
    """bg_var and foreground before the matched filters (fitting.py:452-478)."""
    bg_var = background + fit
    if camera_variance is not None:
        bg_var += camera_variance
    foreground = image - background - fit
    return bg_var, foreground


def finder_snr(foreground, bg_var, peak_mask=None):
    """The signal-to-noise image after the matched filters (fitting.py:466-494)."""
    bg_var = bg_var.copy()
    mask = (bg_var <= 0.1)
    if numpy.sum(mask) > 0:
        bg_var[mask] = 0.1
    snr = foreground / numpy.sqrt(bg_var)
    if peak_mask is not None:
        snr = snr * peak_mask
    return snr


def peak_finder_snr(image, background, fit, fg_filter=None, var_filter=None, peak_mask=None, camera_variance=None):
    """PeakFinderGaussian.peakFinder up to the maxima search (fitting.py:444-498)."""
    bg_var, foreground = finder_prepare(image, background, fit, camera_variance)
    if var_filter is not None:
        bg_var = var_filter.convolve(bg_var)
    if fg_filter is not None:
        foreground = fg_filter.convolve(foreground)
    return finder_snr(foreground, bg_var, peak_mask)
