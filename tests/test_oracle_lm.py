"""
Pins of the oracle's LM state machine (mFitIterateLM, mFitRecenterPeaks):

- sweep by sweep, the oracle agrees with tests/lm_pure.py, a second
  restatement written independently from the reference's text, on tiny frames
  with overlapping peaks for the four DAO models: status flags, per-peak cycle
  counts and every fit counter exactly, parameters and errors to 1e-9
  relative (the two sum in different orders only inside the 4-6 term
  Cholesky);
- an isolated peak's error never increases by more than the tolerance from
  one sweep to the next (an increase below tol ends the fit as CONVERGED,
  multi_fit.c:1259-1283);
- bookkeeping: the per-peak cycle counts sum to at most the fitter's
  n_iterations; they fall short only when turns were reset (mFitResetPeak,
  multi_fit.c:1788-1798, restores the turn's starting state including its
  cycle count, while n_iterations keeps every cycle).
"""
import numpy
import pytest

from conftest import make_fitter
from lm_pure import PureFitter
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.diagnostics import synth

ROI = 10
SMIN, SMAX = 0.75, 3.0
COUNTERS = ("n_dposv", "n_iterations", "n_lost", "n_margin", "n_neg_fi", "n_neg_height", "n_non_decr")
CLASSES = {"2dfixed": D.MultiFitter2DFixed, "2d": D.MultiFitter2D, "3d": D.MultiFitter3D, "Z": D.MultiFitterZ}


def tiny_frame(seed, model, n=4, edge=False):
    rng = numpy.random.default_rng(seed)
    H, W = 36, 44
    m = ROI // 2 + 3
    x = rng.uniform(m, W - m, n)
    y = rng.uniform(m, H - m, n)
    x[1], y[1] = x[0] + 2.5, y[0] + 1.0  # an overlapping pair
    if edge:
        # an emitter beyond the right margin: its start peak walks out
        x[2], y[2] = W - ROI // 2 + 1.5, 0.5 * H
    h = rng.uniform(600.0, 2000.0, n)
    if model == "Z":
        z = rng.uniform(-0.3, 0.3, n)
        sx, sy = 0.5 * synth.width_from_z(z, synth.WX_Z), 0.5 * synth.width_from_z(z, synth.WY_Z)
    else:
        z = numpy.zeros(n)
        sx = rng.uniform(1.2, 1.8, n)
        sy = sx.copy() if model != "3d" else rng.uniform(1.2, 1.8, n)
        if model == "2dfixed":
            sx[:] = sy[:] = 1.5
    img = synth.camera(rng, synth.render((H, W), x, y, h, sx, sy, 100.0))
    s0 = numpy.full(n, 1.5)
    if edge:
        x[2] = W - ROI // 2 - 1.0  # start just inside the margin
    peaks = synth.start_peaks_hdf5(rng, img, x, y, numpy.zeros(n), numpy.full(n, 100.0), s0, s0)
    if edge:
        # a start peak on bare background, far from any emitter
        far = numpy.argmax(numpy.hypot(x[:, None] - numpy.arange(m, W - m)[None, :], 0.0).min(axis=0))
        for k, v in (("x", float(m + far)), ("y", float(H - m)), ("z", 0.0), ("background", 100.0),
                     ("height", 50.0), ("xsigma", 1.5), ("ysigma", 1.5)):
            peaks[k] = numpy.append(peaks[k], v)
    return img, peaks


def kwargs(model):
    if model == "Z":
        return {"wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
    if model in ("2d", "3d"):
        return {"sigma_range": [SMIN, SMAX]}
    return {}


def pure_fitter(model, shape):
    kw = {}
    if model == "Z":
        kw = {"wx": synth.WX_Z, "wy": synth.WY_Z, "z_min": -0.5, "z_max": 0.5}
    elif model in ("2d", "3d"):
        kw = {"width_min": 1.0 / (2.0 * SMAX * SMAX), "width_max": 1.0 / (2.0 * SMIN * SMIN)}
    return PureFitter(shape, ROI, model, **kw)


def rows7(peaks):
    return numpy.stack([peaks[k] for k in ("x", "y", "z", "background", "height", "xsigma", "ysigma")], axis=1)


def compare(m, pure, where):
    pp = pure.properties()
    st = m.getPeakProperty("status")
    numpy.testing.assert_array_equal(st, pp["status"], err_msg=where)
    numpy.testing.assert_array_equal(m.getPeakProperty("iterations"), pp["iterations"], err_msg=where)
    for k in ("x", "y", "z", "height", "background", "xsigma", "ysigma", "error"):
        a = m.getPeakProperty(k)
        b = numpy.array(pp[k])
        numpy.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12, err_msg="%s %s" % (where, k))
    for c in COUNTERS:
        assert getattr(m.mfit.contents, c) == pure.counters[c], (where, c, getattr(m.mfit.contents, c),
                                                                  pure.counters[c])


@pytest.mark.parametrize("model", ["2dfixed", "2d", "3d", "Z"])
@pytest.mark.parametrize("seed,edge", [(1, False), (2, False), (3, True)])
def test_oracle_sweeps_match_independent_restatement(model, seed, edge):
    img, peaks = tiny_frame(synth.SEED_BASE + 60 + seed, model, edge=edge)
    bg = numpy.full(img.shape, 100.0)
    m = make_fitter(CLASSES[model], "oracle", roi_size=ROI, **kwargs(model))
    m.initializeC(img)
    m.newImage(img)
    m.newBackground(bg)
    m.newPeaks(peaks, "hdf5")
    pure = pure_fitter(model, img.shape)
    pure.new_image(img)
    pure.new_peaks_text(rows7(peaks))
    compare(m, pure, "after newPeaks")
    for sweep in range(300):
        m.iterate()
        pure.iterate()
        compare(m, pure, "sweep %d" % sweep)
        if m.getUnconverged() == 0:
            break
    assert pure.unconverged() == 0
    m.cleanup(verbose=False)


@pytest.mark.parametrize("model", ["2dfixed", "Z"])
def test_isolated_peak_error_never_rises_beyond_tolerance(model):
    rng = numpy.random.default_rng(synth.SEED_BASE + 70)
    H, W = 40, 40
    x, y = numpy.array([19.3]), numpy.array([20.6])
    z = numpy.array([0.2]) if model == "Z" else numpy.zeros(1)
    s = 0.5 * synth.width_from_z(z, synth.WX_Z) if model == "Z" else numpy.array([1.5])
    sy = 0.5 * synth.width_from_z(z, synth.WY_Z) if model == "Z" else s
    img = synth.camera(rng, synth.render((H, W), x, y, numpy.array([1500.0]), s, sy, 100.0))
    peaks = synth.start_peaks_hdf5(rng, img, x, y, numpy.zeros(1), numpy.full(1, 100.0), numpy.full(1, 1.5),
                                   numpy.full(1, 1.5))
    m = make_fitter(CLASSES[model], "oracle", roi_size=ROI, **kwargs(model))
    m.initializeC(img)
    m.newImage(img)
    m.newBackground(numpy.full(img.shape, 100.0))
    m.newPeaks(peaks, "hdf5")
    errs = []
    while True:
        m.iterate()
        errs.append(float(m.getPeakProperty("error")[0]))
        if m.getUnconverged() == 0 or len(errs) > 500:
            break
    assert m.getPeakProperty("status")[0] == D.CONVERGED
    for a, b in zip(errs[:-1], errs[1:]):
        assert b <= a * (1.0 + 1.0e-6), (a, b)
    m.cleanup(verbose=False)


def test_iteration_bookkeeping():
    rng = numpy.random.default_rng(synth.SEED_BASE + 71)
    fr = synth.frame_2dfixed(rng, size=64, n=12)
    m = make_fitter(D.MultiFitter2DFixed, "oracle", roi_size=ROI)
    m.initializeC(fr.image)
    m.newImage(fr.image)
    m.newBackground(numpy.full(fr.image.shape, 100.0))
    m.newPeaks(fr.peaks, "hdf5")
    m.doFit()
    per_peak = m.getPeakProperty("iterations")
    c = m.mfit.contents
    resets = c.n_non_decr + c.n_neg_fi + c.n_margin + c.n_neg_height
    assert int(per_peak.sum()) <= m.iterations
    assert (int(per_peak.sum()) < m.iterations) == (resets > 0)
    m.cleanup(verbose=False)
