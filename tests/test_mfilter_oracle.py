"""
Pins of the matched-filter oracle (oracle/mfilter_oracle.py), CPU only:

- the FFT restatement against the plain circular-convolution sum on small
  arrays of odd and even sizes with arbitrary (non-symmetric) PSFs, which
  fixes the recentring, the normalisation and the orientation;
- a centred delta PSF is the identity;
- the reference's own expectations (storm_analysis/test/test_matched_filter.py:
  test_matched_filter1-4) restated with the same sizes and settings.
(The reference's C needs FFTW, a library outside its sources, so it is not
built here.)
"""
import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mfilter_oracle as MO  # noqa: E402


@pytest.mark.parametrize("shape", [(8, 11), (12, 9), (7, 7), (10, 16)])
def test_fft_convolution_is_the_circular_sum(shape):
    rng = numpy.random.default_rng(shape[0] * 100 + shape[1])
    psf = rng.uniform(-1.0, 1.0, size=shape)
    image = rng.uniform(0.0, 5.0, size=shape)
    got = MO.MatchedFilterOracle(psf).convolve(image)
    want = MO.brute_circular(image, psf)
    assert numpy.max(numpy.abs(got - want)) < 1e-12 * numpy.max(numpy.abs(want))
    # the plausible mistakes: no recentring, or the correlation instead of the convolution
    wrong = MO.brute_circular(image, numpy.roll(numpy.roll(psf, -(shape[0] // 2), 0), -(shape[1] // 2), 1))
    assert numpy.max(numpy.abs(got - wrong)) > 1e-3
    flipped = MO.brute_circular(image, psf[::-1, ::-1])
    assert numpy.max(numpy.abs(got - flipped)) > 1e-3


def test_centred_delta_is_identity():
    """A delta where the recentring roll (by int(size / 2)) lands it on the
    origin -- index size - size // 2, which for odd sizes is one past the middle."""
    psf = numpy.zeros((9, 14))
    psf[9 - 9 // 2, (14 - 14 // 2) % 14] = 1.0
    image = numpy.random.default_rng(2).uniform(size=(9, 14))
    numpy.testing.assert_allclose(MO.MatchedFilterOracle(psf).convolve(image), image, rtol=0, atol=1e-14)


def check_reference_tests(make, sum_tol=1e-6):
    """test_matched_filter1-4 (test_matched_filter.py:14-160) with MatchedFilter-like objects from make()."""
    # 1: a unit-sum filter preserves the image sum
    x_size, y_size = 80, 90
    psf = MO.gaussian_psf((x_size, y_size), 1.0)
    flt = make(psf)
    for i in range(1, 5):
        image = numpy.zeros((x_size, y_size))
        image[x_size // 2, y_size // 2] = float(i)
        assert abs(numpy.sum(image) - numpy.sum(flt.convolve(image))) < sum_tol
    # 2: the height of a PSF-shaped object is recovered
    psf = numpy.exp(-(numpy.mgrid[0:x_size, 0:y_size][0] - x_size / 2) ** 2 / 8.0 -
                    (numpy.mgrid[0:x_size, 0:y_size][1] - y_size / 2) ** 2 / 8.0)
    psf_norm = psf / numpy.sum(psf)
    flt = make(psf_norm)
    rescale = 1.0 / numpy.sum(psf * psf_norm)
    conv = flt.convolve(10.0 * psf)
    assert abs(numpy.amax(conv) * rescale - 10.0) / 10.0 < 1e-2
    # 3: memoization
    x_size, y_size = 40, 50
    flt = make(MO.gaussian_psf((x_size, y_size), 2.0), memoize=True, max_diff=0.1)

    def img(v):
        a = numpy.zeros((x_size, y_size))
        a[x_size // 2, y_size // 2] = v
        return a
    res1 = flt.convolve(img(1.0))
    res2 = flt.convolve(img(1.05))
    assert abs(numpy.max(res2 - res1)) < 1e-12
    res2 = flt.convolve(img(1.1))
    assert abs(numpy.max(res2 - res1)) > 1e-12
    res1 = flt.convolve(img(1.0))
    assert abs(numpy.max(res2 - res1)) > 1e-12
    res2 = flt.convolve(img(1.0))
    assert abs(numpy.max(res2 - res1)) < 1e-12
    # 4: the FFT planning flag has no effect
    psf = MO.gaussian_psf((80, 90), 1.0)
    f1, f2 = make(psf, fftw_estimate=False), make(psf, fftw_estimate=True)
    for i in range(1, 5):
        image = numpy.zeros((80, 90))
        image[40, 45] = float(i)
        assert numpy.allclose(f1.convolve(image), f2.convolve(image), rtol=0, atol=1e-12)


def oracle_factory(psf, fftw_estimate=False, memoize=False, max_diff=0.1):
    return MO.MatchedFilterOracle(psf, memoize=memoize, max_diff=max_diff)


def test_reference_tests_on_the_oracle():
    check_reference_tests(oracle_factory)


def test_memo_returns_the_psf_before_any_convolution():
    """The reference's work buffer holds the recentred PSF after initialize:
    a first memoized call whose image is within max_diff of zero returns it."""
    psf = MO.gaussian_psf((16, 20), 2.0)
    flt = MO.MatchedFilterOracle(psf, memoize=True, max_diff=0.1)
    numpy.testing.assert_array_equal(flt.convolve(numpy.full((16, 20), 0.05)), MO.recenter_psf(psf))


# Hand values for the finder's per-pixel arithmetic (fitting.py:452-494),
# worked out by hand from
#   image      [[10, 5], [3, 1], [4, 6]]
#   background [[2, 2], [1, 0.05], [-1, 0.5]]
#   fit        [[1, 0], [0, 0], [0.5, -0.4]]
#   camera var [[0.5, 0], [0, 0.02], [0.1, 0.05]]
# bg_var = background + fit + camera var:
#   [[3.5, 2], [1, 0.07], [-0.4, 0.15]]
# foreground = image - background - fit:
#   [[7, 3], [2, 0.95], [4.5, 5.9]]
# bg_var clamped where <= 0.1 (0.07 and -0.4 -> 0.1; 0.15 stays), then
# snr = foreground / sqrt(bg_var) * mask with mask [[1, 1], [1, 1], [1, 0]]:
#   7/sqrt(3.5) = 3.7416573867739413   3/sqrt(2) = 2.1213203435596424
#   2/sqrt(1) = 2                      0.95/sqrt(0.1) = 3.0041637771599603
#   4.5/sqrt(0.1) = 14.230249470757707 5.9/sqrt(0.15) * 0 = 0
FINDER_HAND = {
    "image": [[10.0, 5.0], [3.0, 1.0], [4.0, 6.0]],
    "background": [[2.0, 2.0], [1.0, 0.05], [-1.0, 0.5]],
    "fit": [[1.0, 0.0], [0.0, 0.0], [0.5, -0.4]],
    "camera": [[0.5, 0.0], [0.0, 0.02], [0.1, 0.05]],
    "mask": [[1.0, 1.0], [1.0, 1.0], [1.0, 0.0]],
    "bg_var": [[3.5, 2.0], [1.0, 0.07], [-0.4, 0.15]],
    "foreground": [[7.0, 3.0], [2.0, 0.95], [4.5, 5.9]],
    "snr": [[3.7416573867739413, 2.1213203435596424], [2.0, 3.0041637771599603], [14.230249470757707, 0.0]],
    # without the mask the last pixel is 5.9 / sqrt(0.15)
    "snr_unmasked_last": 15.233734495082507,
}


def finder_hand(name):
    return numpy.array(FINDER_HAND[name], dtype=numpy.float64)


def test_finder_arithmetic_hand_values():
    """
    The clamp replaces every variance <= 0.1 (including negative ones) by
    0.1 and leaves 0.15 alone.  Plausible slips that fail: no clamp (NaN at
    -0.4), clamping the standard deviation at 0.1 instead of the variance,
    the camera variance left out, the fit image added to the foreground.
    """
    bg_var, fg = MO.finder_prepare(finder_hand("image"), finder_hand("background"), finder_hand("fit"),
                                   finder_hand("camera"))
    numpy.testing.assert_allclose(bg_var, finder_hand("bg_var"), rtol=1e-15, atol=1e-15)
    numpy.testing.assert_allclose(fg, finder_hand("foreground"), rtol=1e-15, atol=1e-15)
    snr = MO.finder_snr(fg, bg_var, finder_hand("mask"))
    numpy.testing.assert_allclose(snr, finder_hand("snr"), rtol=1e-14)
    assert abs(MO.finder_snr(fg, bg_var)[2, 1] - FINDER_HAND["snr_unmasked_last"]) < 1e-13
