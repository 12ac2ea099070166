"""
Batched multiplane fitting (multi_plane/mp_batch.py: many frame-sets in one
launch) against the multiplane oracle fitting each frame-set alone, with
MPFitDao2D's surface: the same channel maps, weights and calibration for
every set, different emitters per set.  Bars as for the single-set fitter
(test_gpu_mp.py): identical flags and per-peak cycle counts in every channel
of every set, x / y within 1e-3 px, height / background / widths within 1e-4
relative; also the sets' peak counts after removing ERROR groups.
"""
import numpy
import pytest

from conftest import make_mp_fitter
from test_gpu_mp import affine_maps
from storm_analysis_amd.multi_plane import mp_batch
from storm_analysis_amd.multi_plane import mp_fit_dao_c as MPD
from storm_analysis_amd.sa_library import dao_fit_c as D


def movie(seed, n_sets, n_ch, size, n_em, sigma=1.5):
    rng = numpy.random.default_rng(seed)
    shape = (size, size)
    maps = affine_maps(rng, n_ch, shape)
    xt, yt = maps[0], maps[1]
    cal = [(rng.uniform(0.8, 1.2, shape), rng.uniform(0.0, 2.0, shape)) for _ in range(n_ch)]
    yy, xx = numpy.mgrid[0:size, 0:size]
    images, starts = [], []
    for s in range(n_sets):
        ex = rng.uniform(3.0, size - 3.0, n_em)
        ey = rng.uniform(3.0, size - 3.0, n_em)
        h0 = rng.uniform(300.0, 1500.0, n_em)
        per = []
        for c in range(n_ch):
            rqe, scmos = cal[c]
            cx = xt[c, 0] + xt[c, 1] * ex + xt[c, 2] * ey
            cy = yt[c, 0] + yt[c, 1] * ex + yt[c, 2] * ey
            hc = h0 * rng.uniform(0.3, 1.0, n_em)
            mean = numpy.full(shape, rng.uniform(10.0, 40.0))
            for k in range(n_em):
                mean += hc[k] * numpy.exp(-((xx - cx[k]) ** 2 + (yy - cy[k]) ** 2) / (2 * sigma ** 2))
            x = numpy.maximum(rng.poisson(mean * rqe + scmos).astype(float), 1.0)
            per.append((x - scmos) / rqe)
        images.append(per)
        starts.append({"x": ex + rng.normal(0, 0.3, n_em), "y": ey + rng.normal(0, 0.3, n_em), "z": numpy.zeros(n_em),
                       "height": h0 * 0.6, "background": numpy.full(n_em, 25.0),
                       "xsigma": numpy.full(n_em, 1.6), "ysigma": numpy.full(n_em, 1.6)})
    return shape, maps, cal, numpy.array(images), starts


def oracle_set(shape, maps, cal, images, start, roi):
    m = make_mp_fitter(MPD.MPFitDao2D, "oracle", n_channels=len(cal), roi_size=roi, sigma_range=[0.8, 3.0])
    for c, (rqe, scmos) in enumerate(cal):
        m.initializeChannel(rqe, scmos, c)
    m.setMapping(*maps)
    m.setWeights()
    m.newImage(list(images))
    m.newBackground([numpy.zeros(shape) for _ in cal])
    m.newPeaks(start, "text")
    m.doFit()
    out = []
    for c in range(len(cal)):
        out.append({k: m.getPeakProperty(k, channel=c) for k in ("x", "y", "height", "background", "xsigma",
                                                                   "ysigma", "status", "iterations")})
    m.removeErrorPeaks()
    n_left = m.getNFit()
    m.cleanup(verbose=False)
    return out, n_left


@pytest.mark.gpu
@pytest.mark.parametrize("n_sets,n_ch,size,n_em,roi", [(6, 3, 128, 60, 12), (3, 2, 160, 120, 12)])
def test_batched_sets_match_the_oracle_set_by_set(n_sets, n_ch, size, n_em, roi):
    shape, maps, cal, images, starts = movie(200 + n_sets, n_sets, n_ch, size, n_em)
    bf = mp_batch.MPBatchFitter(n_sets, n_ch, shape, roi_size=roi, sigma_range=[0.8, 3.0], peak_cap=2 * n_em)
    try:
        for c, (rqe, scmos) in enumerate(cal):
            bf.initializeChannel(rqe, scmos, c)
        bf.setMapping(*maps)
        bf.setWeights()
        bf.newImages(images)
        bf.newBackgrounds(numpy.zeros(images.shape))
        bf.newPeaks(starts, "text")
        bf.doFit()
        got = [[bf.getPeaks(s, c) for c in range(n_ch)] for s in range(n_sets)]
        bf.removeErrorPeaks()
        left = [bf.getNFit(s) for s in range(n_sets)]
    finally:
        bf.cleanup()
    for s in range(n_sets):
        want, n_left = oracle_set(shape, maps, cal, images[s], starts[s], roi)
        for c in range(n_ch):
            _assert_peaks(got[s][c], want[c], "set %d channel %d" % (s, c))
        assert left[s] == n_left, (s, left[s], n_left)


def _assert_peaks(g, r, what):
    assert g["x"].size == r["x"].size, what
    numpy.testing.assert_array_equal(g["status"], r["status"], err_msg=what)
    numpy.testing.assert_array_equal(g["iterations"], r["iterations"], err_msg=what)
    ok = r["status"] != D.ERROR
    for k in ("x", "y"):
        d = numpy.abs(g[k] - r[k])[ok]
        assert d.size == 0 or d.max() < 1e-3, (what, k, d.max())
    for k in ("height", "background", "xsigma", "ysigma"):
        d = (numpy.abs(g[k] - r[k]) / numpy.abs(r[k]))[ok]
        assert d.size == 0 or d.max() < 1e-4, (what, k, d.max())
