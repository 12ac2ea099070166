"""
The fitter's read-outs (multi_fit.c:708-888, 1518-1632): fit image, residual,
and the "sum", "fg_sum", "fg_sum_sc", "bg_sum" and "jacobian" properties.

- Closed forms, per peak alone in its ROI (psf = the peak's shape over the
  circular ROI; x = image rqe + scmos):
    fit image   f / rqe, f = h psf rqe inside the ROI, 0 outside (:708-720)
    residual    x - f (:880-888)
    sum         h sum(psf)                                       (:1618-1632)
    fg_sum      h sum(psf^2) / sum(psf)                          (:1581-1606)
    bg_sum      sum(((x - scmos) / rqe - h psf) psf^2) / sum(psf)^2,
                1 when not positive                              (:1518-1566)
    jacobian    dD/dtheta, by complex steps (tests/lm_evidence.py)
  Plausible slips these catch: the residual without the sCMOS term, the fit
  image without the rqe division, fg_sum without the normalisation, bg_sum
  on the rqe-scaled data.
- The reference's own residual checks: an image of ones reads back as ones
  (test_3ddao.py:144-165) and as ones plus the sCMOS term with a
  calibration (test_scmos.py:78-99).
- HIP against the oracle on a fitted frame of overlapping peaks, for every
  read-out (-m gpu).
"""
import numpy
import pytest

from conftest import make_fitter
from lm_evidence import Frame
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.diagnostics import synth

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]
R = 12
SHAPE = (48, 64)
SIGMA = 1.5


def scene(seed=4):
    """Three emitters far apart (their ROIs do not overlap) on a background of 30; start peaks near them."""
    rng = numpy.random.default_rng(seed)
    rqe = rng.uniform(0.7, 1.3, SHAPE)
    scmos = rng.uniform(0.0, 3.0, SHAPE)
    yy, xx = numpy.mgrid[0:SHAPE[0], 0:SHAPE[1]]
    mean = 30.0 * rqe + scmos
    for h, px, py in ((420.0, 14.0, 15.9), (260.0, 34.1, 30.0), (500.0, 52.4, 17.1)):
        mean = mean + h * numpy.exp(-((xx - px) ** 2 + (yy - py) ** 2) / (2 * SIGMA ** 2)) * rqe
    x = numpy.maximum(rng.poisson(mean).astype(float), 1.0)
    peaks = numpy.array([[400.0, 14.3, 15.6, 28.0], [250.0, 33.8, 30.2, 31.0], [520.0, 52.1, 17.4, 27.5]])
    return x, rqe, scmos, peaks


def fitter(backend, x, rqe, scmos, peaks, model=D.MultiFitter2DFixed, **kw):
    m = make_fitter(model, backend, roi_size=R, rqe=rqe, scmos_cal=scmos, **kw)
    image = (x - scmos) / rqe
    m.initializeC(image)
    m.newImage(image)
    m.newBackground(numpy.zeros(SHAPE))
    n = peaks.shape[0]
    m.newPeaks({"x": peaks[:, 1], "y": peaks[:, 2], "z": numpy.zeros(n), "background": peaks[:, 3],
                "height": peaks[:, 0], "xsigma": numpy.full(n, SIGMA), "ysigma": numpy.full(n, SIGMA)}, "text")
    return m


def closed_forms(x, rqe, scmos, peaks):
    fr = Frame(x, rqe, scmos, R, width=1.0 / (2 * SIGMA ** 2))
    f = numpy.zeros(SHAPE)
    out = {"sum": [], "fg_sum": [], "bg_sum": [], "jacobian": []}
    for h, px, py, bg in peaks:
        th = numpy.array([h, px - fr.xoff, py - fr.xoff, bg])
        xi, yi = fr.origin(th)
        rows, cols = yi + fr.ry, xi + fr.rx
        psf = numpy.exp(-((cols - px) ** 2 + (rows - py) ** 2) / (2 * SIGMA ** 2))
        f[rows, cols] += h * psf * rqe[rows, cols]
        out["sum"].append(h * psf.sum())
        out["fg_sum"].append(h * (psf ** 2).sum() / psf.sum())
        b = numpy.sum(((x[rows, cols] - scmos[rows, cols]) / rqe[rows, cols] - h * psf) * psf ** 2)
        out["bg_sum"].append(b / psf.sum() ** 2 if b > 0 else 1.0)
        out["jacobian"].append(fr.jh(th, xi, yi)[0])
    out["fit"] = f / rqe
    out["residual"] = x - f
    return out


@pytest.mark.parametrize("backend", BACKENDS)
def test_read_outs_match_their_closed_forms(backend):
    x, rqe, scmos, peaks = scene()
    want = closed_forms(x, rqe, scmos, peaks)
    m = fitter(backend, x, rqe, scmos, peaks)
    numpy.testing.assert_allclose(m.getFitImage(), want["fit"], rtol=1e-12, atol=1e-12)
    numpy.testing.assert_allclose(m.getResidual(), want["residual"], rtol=1e-12, atol=1e-10)
    for k in ("sum", "fg_sum", "bg_sum"):
        numpy.testing.assert_allclose(m.getPeakProperty(k), want[k], rtol=1e-10, err_msg=k)
    # fg_sum_sc is the C library's alias of fg_sum (multi_fit.c:780-785); the
    # reference's Python does not list it, so it is read through the C-ABI
    sc = numpy.zeros(peaks.shape[0])
    m.clib.mFitGetPeakPropertyDouble(m.mfit, sc, b"fg_sum_sc")
    numpy.testing.assert_allclose(sc, want["fg_sum"], rtol=1e-10)
    jac = m.getPeakProperty("jacobian").reshape(peaks.shape[0], 4)
    numpy.testing.assert_allclose(jac, numpy.array(want["jacobian"]), rtol=1e-8, atol=1e-8)
    sig = m.getPeakProperty("significance")
    numpy.testing.assert_allclose(sig, numpy.array(want["fg_sum"]) / numpy.sqrt(want["bg_sum"]), rtol=1e-10)
    m.cleanup(verbose=False)


@pytest.mark.parametrize("backend", BACKENDS)
def test_bg_sum_falls_back_to_one(backend):
    """A peak far brighter than its pixels: the weighted background sum is negative -> 1.0 (:1558-1565)."""
    x, rqe, scmos, peaks = scene()
    peaks = peaks.copy()
    peaks[1, 0] = 5.0e4
    want = closed_forms(x, rqe, scmos, peaks)
    assert want["bg_sum"][1] == 1.0 and want["bg_sum"][0] != 1.0
    m = fitter(backend, x, rqe, scmos, peaks)
    numpy.testing.assert_allclose(m.getPeakProperty("bg_sum"), want["bg_sum"], rtol=1e-10)
    m.cleanup(verbose=False)


@pytest.mark.parametrize("backend", BACKENDS)
def test_residual_of_a_new_image_is_the_image(backend):
    """test_3ddao.py:144-165 (no calibration) and test_scmos.py:78-99 (x + sCMOS term)."""
    ones = numpy.ones((100, 100))
    m = make_fitter(D.MultiFitter2DFixed, backend, roi_size=R)
    m.initializeC(ones)
    m.newImage(ones)
    assert numpy.max(numpy.abs(m.getResidual() - 1.0)) < 1.0e-6
    m.cleanup(verbose=False)
    cal = numpy.random.default_rng(2).uniform(0.5, 2.0, (100, 100))
    m = make_fitter(D.MultiFitter2DFixed, backend, roi_size=R, scmos_cal=cal)
    m.initializeC(ones)
    m.newImage(ones)
    assert numpy.max(numpy.abs(m.getResidual() - (1.0 + cal))) < 1.0e-6
    m.cleanup(verbose=False)


READ_OUTS = ("sum", "fg_sum", "bg_sum", "significance", "jacobian", "error", "x", "y", "z", "height",
             "background", "xsigma", "ysigma")


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["2dfixed", "3d", "Z"])
def test_hip_read_outs_match_the_oracle_after_a_fit(model):
    """Overlapping peaks fitted to the end, then every read-out: HIP vs oracle."""
    rng = numpy.random.default_rng(synth.SEED_BASE + 77)
    if model == "Z":
        fr = synth.frame_z(rng, size=128, n=90)
        kw = {"wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
        cls, r, bg = D.MultiFitterZ, 18, None
    elif model == "3d":
        fr = synth.frame_3d_scmos(rng, size=96, density=1.5)
        kw = {"sigma_range": [0.6, 5.0]}
        cls, r, bg = D.MultiFitter3D, 18, 100.0
    else:
        fr = synth.frame_2dfixed(rng, size=128, n=120)
        kw = {}
        cls, r, bg = D.MultiFitter2DFixed, 12, 100.0
    out = {}
    for backend in ("oracle", "hip"):
        m = make_fitter(cls, backend, roi_size=r, rqe=fr.rqe, scmos_cal=fr.scmos_cal, **kw)
        m.initializeC(fr.image)
        m.newImage(fr.image)
        if bg is None:
            m.newBackground(synth.z_background(fr.image.shape))
        else:
            m.newBackground(numpy.full(fr.image.shape, bg))
        m.newPeaks(fr.peaks, "hdf5")
        m.doFit()
        o = {k: m.getPeakProperty(k) for k in READ_OUTS}
        o["status"] = m.getPeakProperty("status")
        o["fit"] = m.getFitImage()
        o["residual"] = m.getResidual()
        out[backend] = o
        m.cleanup(verbose=False)
    g, r_ = out["hip"], out["oracle"]
    numpy.testing.assert_array_equal(g["status"], r_["status"])
    ok = r_["status"] != D.ERROR
    for k in ("x", "y", "z"):
        assert numpy.max(numpy.abs(g[k] - r_[k])[ok]) < 1e-3, k
    for k in ("sum", "fg_sum", "bg_sum", "significance", "height", "background", "xsigma", "ysigma", "error"):
        d = numpy.abs(g[k] - r_[k]) / numpy.maximum(numpy.abs(r_[k]), 1e-300)
        assert numpy.max(d[ok]) < 1e-4, (k, numpy.max(d[ok]))
    jg = g["jacobian"].reshape(ok.size, -1)[ok]
    jr = r_["jacobian"].reshape(ok.size, -1)[ok]
    scale = numpy.maximum(numpy.abs(jr).max(axis=1, keepdims=True), 1.0)
    assert numpy.max(numpy.abs(jg - jr) / scale) < 1e-4
    for k in ("fit", "residual"):
        scale = numpy.max(numpy.abs(r_[k]))
        assert numpy.max(numpy.abs(g[k] - r_[k])) <= 1e-6 * scale, k
