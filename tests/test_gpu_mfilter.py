"""
GPU FFT convolution (libmatched_filter.so through
storm_analysis_amd.sa_library.dao_mfilter.MatchedFilter) against the oracle
(oracle/mfilter_oracle.py).  Tolerance: the two compute the same discrete
Fourier transforms in float64 with different summation orders, so results
agree to FFT rounding; the tests require max |gpu - oracle| <= 1e-10 x
max |oracle| (an error in the recentring, normalisation, orientation or the
spectrum product is of order 1).
"""
import os
import sys

import numpy
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mfilter_oracle as MO  # noqa: E402
from test_mfilter_oracle import check_reference_tests  # noqa: E402

from storm_analysis_amd.sa_library import dao_mfilter  # noqa: E402

REL_TOL = 1e-10


def close(got, want):
    err = numpy.max(numpy.abs(got - want))
    assert err <= REL_TOL * numpy.max(numpy.abs(want)), err


@pytest.mark.gpu
def test_reference_tests():
    check_reference_tests(lambda psf, **kw: dao_mfilter.MatchedFilter(psf, **kw))


@pytest.mark.gpu
@pytest.mark.parametrize("shape,psf_kind", [((80, 90), "gauss"), ((33, 47), "random"), ((512, 512), "gauss"),
                                            ((64, 31), "random")])
def test_convolve_matches_oracle(shape, psf_kind):
    rng = numpy.random.default_rng(shape[0] + shape[1])
    psf = MO.gaussian_psf(shape, 8.0 if shape[0] == 512 else 1.5) if psf_kind == "gauss" else \
        rng.uniform(-1.0, 1.0, size=shape)
    flt = dao_mfilter.MatchedFilter(psf)
    orc = MO.MatchedFilterOracle(psf)
    for _ in range(3):
        image = rng.uniform(0.0, 100.0, size=shape)
        close(flt.convolve(image), orc.convolve(image))
    flt.cleanup()


@pytest.mark.gpu
def test_memo_sequence_matches_oracle():
    shape = (40, 50)
    psf = MO.gaussian_psf(shape, 2.0)
    flt = dao_mfilter.MatchedFilter(psf, memoize=True, max_diff=0.1)
    orc = MO.MatchedFilterOracle(psf, memoize=True, max_diff=0.1)
    rng = numpy.random.default_rng(9)
    base = rng.uniform(0.0, 1.0, size=shape)
    # first within max_diff of the zero start image (returns the PSF), then
    # changes above and below max_diff
    for image in (numpy.full(shape, 0.05), base, base + 0.05, base + 0.2, base + 0.2, base):
        got, want = flt.convolve(image), orc.convolve(image)
        close(got, want)
    flt.cleanup()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(128, 96), (33, 47), (512, 512)])
def test_batched_convolution_matches_oracle(shape):
    import torch
    psf = MO.gaussian_psf(shape, 2.0)
    flt = dao_mfilter.MatchedFilter(psf)
    orc = MO.MatchedFilterOracle(psf)
    rng = numpy.random.default_rng(12)
    for n in (8, 3):  # a second call with another batch size re-plans
        images = rng.uniform(0.0, 50.0, size=(n,) + shape)
        dev = torch.from_numpy(images).cuda()
        out = flt.convolveBatch(dev).cpu().numpy()
        numpy.testing.assert_array_equal(dev.cpu().numpy(), images)  # input untouched
        for i in range(n):
            close(out[i], orc.convolve(images[i]))
    flt.cleanup()
