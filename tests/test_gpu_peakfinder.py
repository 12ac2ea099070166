"""
The peak finder's per-pixel arithmetic on the GPU (daoFinderPrepare /
daoFinderSNR through storm_analysis_amd.sa_library.dao_peakfinder) against
the oracle's restatement of PeakFinderGaussian.peakFinder
(oracle/mfilter_oracle.py: finder_prepare, finder_snr, peak_finder_snr):
bit-identical on the same inputs; through the matched filters within the
convolution's tolerance (max |gpu - oracle| <= 1e-10 x max |oracle|); and the
device-resident chain ending in the device finder gives the maxima the
oracle finder gives on the same search image.
"""
import os
import sys

import numpy
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import finder_oracle as FD  # noqa: E402
import mfilter_oracle as MO  # noqa: E402

from storm_analysis_amd.sa_library import dao_finder, dao_mfilter, dao_peakfinder as PF  # noqa: E402


def inputs(seed, n, shape):
    rng = numpy.random.default_rng(seed)
    image = rng.uniform(0.0, 200.0, size=(n,) + shape)
    background = rng.uniform(-1.0, 2.0, size=(n,) + shape)  # bg_var <= 0.1 in places
    fit = rng.uniform(-1.0, 1.0, size=(n,) + shape)
    background[:, 0, 0], fit[:, 0, 0] = 0.05, 0.05  # exactly 0.1
    return image, background, fit


def cuda(a):
    import torch
    return torch.from_numpy(numpy.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("with_maps", [False, True])
def test_prepare_and_snr_bit_identical(with_maps):
    shape = (40, 56)
    image, background, fit = inputs(3, 4, shape)
    rng = numpy.random.default_rng(4)
    cam = rng.uniform(0.0, 3.0, size=shape) if with_maps else None
    mask = (rng.uniform(size=shape) > 0.3).astype(numpy.float64) if with_maps else None
    bg_var, fg = PF.finderPrepare(cuda(image), cuda(background), cuda(fit), cam)
    snr = PF.finderSNR(fg, bg_var, mask).cpu().numpy()
    for f in range(4):
        w_var, w_fg = MO.finder_prepare(image[f], background[f], fit[f], cam)
        numpy.testing.assert_array_equal(bg_var[f].cpu().numpy(), w_var)
        numpy.testing.assert_array_equal(fg[f].cpu().numpy(), w_fg)
        numpy.testing.assert_array_equal(snr[f], MO.finder_snr(w_fg, w_var, mask))


@pytest.mark.gpu
def test_prepare_and_snr_hand_values():
    """The hand-worked values of test_mfilter_oracle.FINDER_HAND through the HIP kernels."""
    from test_mfilter_oracle import FINDER_HAND, finder_hand
    bg_var, fg = PF.finderPrepare(cuda(finder_hand("image")[None]), cuda(finder_hand("background")[None]),
                                  cuda(finder_hand("fit")[None]), finder_hand("camera"))
    numpy.testing.assert_allclose(bg_var[0].cpu().numpy(), finder_hand("bg_var"), rtol=1e-15, atol=1e-15)
    numpy.testing.assert_allclose(fg[0].cpu().numpy(), finder_hand("foreground"), rtol=1e-15, atol=1e-15)
    snr = PF.finderSNR(fg, bg_var, finder_hand("mask")).cpu().numpy()[0]
    numpy.testing.assert_allclose(snr, finder_hand("snr"), rtol=1e-14)
    plain = PF.finderSNR(fg, bg_var).cpu().numpy()[0]
    assert abs(plain[2, 1] - FINDER_HAND["snr_unmasked_last"]) < 1e-13


@pytest.mark.gpu
def test_search_image_through_the_matched_filters():
    shape = (96, 80)
    image, background, fit = inputs(5, 3, shape)
    background = numpy.abs(background) + 5.0
    psf = MO.gaussian_psf(shape, 1.5)
    fg_f, var_f = dao_mfilter.MatchedFilter(psf), dao_mfilter.MatchedFilter(psf * psf)
    mask = numpy.ones(shape)
    mask[:, :4] = 0.0
    snr = PF.peakFinderSNR(cuda(image), cuda(background), cuda(fit), fg_f, var_f, mask).cpu().numpy()
    o_fg, o_var = MO.MatchedFilterOracle(psf), MO.MatchedFilterOracle(psf * psf)
    for f in range(3):
        want = MO.peak_finder_snr(image[f], background[f], fit[f], o_fg, o_var, mask)
        assert numpy.max(numpy.abs(snr[f] - want)) <= 1e-10 * numpy.max(numpy.abs(want))
    fg_f.cleanup()
    var_f.cleanup()


@pytest.mark.gpu
def test_device_chain_to_maxima():
    """ingested-like frames -> search image -> maxima, all on the GPU; the
    oracle finder on the same search image returns the same peaks."""
    shape = (128, 128)
    rng = numpy.random.default_rng(6)
    yy, xx = numpy.mgrid[0:shape[0], 0:shape[1]]
    frames = []
    for _ in range(3):
        img = 10.0 + rng.normal(0.0, 1.0, size=shape)
        for y, x, h in zip(rng.uniform(5, 123, 40), rng.uniform(5, 123, 40), rng.uniform(50, 200, 40)):
            img += h * numpy.exp(-((yy - y) ** 2 + (xx - x) ** 2) / 4.5)
        frames.append(img)
    image = numpy.stack(frames)
    background = numpy.full(image.shape, 10.0)
    fit = numpy.zeros(image.shape)
    psf = MO.gaussian_psf(shape, 1.5)
    fg_f, var_f = dao_mfilter.MatchedFilter(psf), dao_mfilter.MatchedFilter(psf * psf)
    snr = PF.peakFinderSNR(cuda(image), cuda(background), cuda(fit), fg_f, var_f, None)
    for f in range(3):
        mxf = dao_finder.MaximaFinder(margin=5, radius=3.0, threshold=8.0, z_values=[0.0])
        x, y, z = mxf.findMaxima(snr[f:f + 1])
        host = snr[f].cpu().numpy()
        tk = [numpy.zeros(shape, dtype=numpy.int32)]
        wz, wy, wx, wh = FD.find_local_maxima([host], tk, 6, 1, 3.0, 8.0, [0.0], host.size)
        numpy.testing.assert_array_equal(x, wx)
        numpy.testing.assert_array_equal(y, wy)
        assert x.size >= 20
    fg_f.cleanup()
    var_f.cleanup()
