"""
Mixed precision (BatchFitter precision="mixed"): FP32 deviance terms,
Jacobian and Hessian; FP64 peak shape, accumulators, fi - x, deviance sums,
Cholesky solve and accept/reject.  Checked against the FP64 oracle on the
same frames and start peaks:

- the same peak count; converged flags equal for at least 99% of the peaks;
- over peaks CONVERGED in both, the 99th percentile of |dx|, |dy| within
  6.25e-3 px (1 nm at 160 nm pixels) and of |dz| within 1e-3 um, and of the
  relative height, background and width differences within 1e-2 (measured:
  ~1e-4 for isolated Z peaks, ~2e-3 in dense free-width 2D, where the
  background of overlapping peaks is weakly determined).
Percentiles rather than maxima: a deviance change that lands within the
FP32 rounding of the 1e-6 tolerance ends a peak's fit one cycle earlier or
later (about 1% of peaks), and a few peaks in overlapping groups then settle
up to ~1 px apart -- both answers are converged fits.  The medians are
~1e-8 px.  The measured percentiles and maxima are printed.
"""
import numpy
import pytest

from conftest import make_fitter
from storm_analysis_amd.sa_library import dao_batch
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.diagnostics import synth

TOL_XY = 6.25e-3
TOL_Z = 1e-3
TOL_REL = 1e-2
MIN_FLAG_AGREEMENT = 0.99
QUANTILE = 0.99


def oracle_fit(cls, kw, fr, background):
    m = make_fitter(cls, "oracle", rqe=fr.rqe, scmos_cal=fr.scmos_cal, **kw)
    m.initializeC(fr.image)
    m.newImage(fr.image)
    m.newBackground(background)
    m.newPeaks(fr.peaks, "hdf5")
    m.doFit()
    out = {k: m.getPeakProperty(k) for k in ("x", "y", "z", "height", "background", "xsigma", "ysigma")}
    out["status"] = m.getPeakProperty("status")
    m.cleanup(verbose=False)
    return out


def mixed_fit(model, kw, frames, backgrounds):
    bf = dao_batch.BatchFitter([fr.image.shape for fr in frames], [len(fr.peaks["x"]) for fr in frames],
                               model=model, precision="mixed", **kw)
    for k, fr in enumerate(frames):
        if fr.scmos_cal is not None:
            bf.setCalibration(k, rqe=fr.rqe, scmos_cal=fr.scmos_cal)
    bf.newImages([fr.image for fr in frames])
    bf.newBackgrounds(backgrounds)
    bf.newPeaks([fr.peaks for fr in frames], "hdf5")
    bf.doFit()
    out = [bf.getPeaks(k) for k in range(len(frames))]
    bf.cleanup()
    return out


def differences(g, r):
    """Flag agreement and, over peaks CONVERGED in both, |dxy|, |dz| and relative differences."""
    assert g["x"].size == r["x"].size
    agree = g["status"] == r["status"]
    both = (g["status"] == D.CONVERGED) & (r["status"] == D.CONVERGED)
    dxy = numpy.maximum(numpy.abs(g["x"] - r["x"]), numpy.abs(g["y"] - r["y"]))[both]
    dz = numpy.abs(g["z"] - r["z"])[both]
    drel = numpy.max([(numpy.abs(g[k] - r[k]) / numpy.abs(r[k]))[both]
                      for k in ("height", "background", "xsigma", "ysigma")], axis=0)
    return agree, dxy, dz, drel


def check(results, refs):
    parts = [differences(g, r) for g, r in zip(results, refs)]
    agree, dxy, dz, drel = (numpy.concatenate([p[i] for p in parts]) for i in range(4))
    q = lambda v: (numpy.quantile(v, 0.5), numpy.quantile(v, QUANTILE), v.max())  # noqa: E731
    print("mixed vs FP64 oracle: %d peaks, flags equal %.4f; within 1 nm %.4f; "
          "dxy px p50/p99/max %.1e/%.1e/%.1e; dz um %.1e/%.1e/%.1e; rel %.1e/%.1e/%.1e" %
          ((agree.size, agree.mean(), (dxy < TOL_XY).mean()) + q(dxy) + q(dz) + q(drel)))
    assert agree.mean() >= MIN_FLAG_AGREEMENT
    assert numpy.quantile(dxy, QUANTILE) < TOL_XY
    assert numpy.quantile(dz, QUANTILE) < TOL_Z
    assert numpy.quantile(drel, QUANTILE) < TOL_REL


@pytest.mark.gpu
def test_mixed_z_against_oracle():
    rng = numpy.random.default_rng(synth.SEED_BASE + 40)
    frames = [synth.frame_z(rng, size=256, n=250) for _ in range(3)]
    kw = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
    bgs = [numpy.zeros(fr.image.shape) for fr in frames]
    out = mixed_fit("Z", kw, frames, bgs)
    refs = [oracle_fit(D.MultiFitterZ, kw, fr, bg) for fr, bg in zip(frames, bgs)]
    check(out, refs)


@pytest.mark.gpu
def test_mixed_2d_dense_against_oracle():
    rng = numpy.random.default_rng(synth.SEED_BASE + 41)
    fr = synth.frame_2d_dense(rng, size=192, n=500)
    kw = {"roi_size": 18, "sigma_range": [0.6, 5.0]}
    bg = numpy.full(fr.image.shape, 100.0)
    out = mixed_fit("2d", kw, [fr], [bg])
    ref = oracle_fit(D.MultiFitter2D, kw, fr, bg)
    check(out, [ref])


@pytest.mark.gpu
def test_mixed_3d_scmos_dense_against_truth():
    """
    Dense free-width 3D with sCMOS noise (up to 2.5 emitters/um^2): here the
    FP64 fits themselves crawl (~220 LM cycles per peak) and stop on flat
    valleys, so mixed and FP64 stop at different points (median ~0.03 px
    apart) and parity to FP64 is not the measure.  The bar instead: the same
    flags for >= 99% of peaks, and localization error against the true
    emitter positions no worse than FP64's (RMS within 5%).
    """
    rng = numpy.random.default_rng(synth.SEED_BASE + 42)
    frames = [synth.frame_3d_scmos(rng, size=128, density=d) for d in (1.0, 2.5)]
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    out = mixed_fit("3d", kw, frames, bgs)
    refs = [oracle_fit(D.MultiFitter3D, kw, fr, bg) for fr, bg in zip(frames, bgs)]
    agree = numpy.concatenate([g["status"] == r["status"] for g, r in zip(out, refs)])
    err_m, err_f = [], []
    for g, r, fr in zip(out, refs, frames):
        both = (g["status"] == D.CONVERGED) & (r["status"] == D.CONVERGED)
        tx, ty = fr.truth["x"], fr.truth["y"]
        err_m.append(numpy.hypot(g["x"] - tx, g["y"] - ty)[both])
        err_f.append(numpy.hypot(r["x"] - tx, r["y"] - ty)[both])
    err_m, err_f = numpy.concatenate(err_m), numpy.concatenate(err_f)
    rms_m, rms_f = float(numpy.sqrt(numpy.mean(err_m ** 2))), float(numpy.sqrt(numpy.mean(err_f ** 2)))
    print("dense 3D sCMOS: %d peaks, flags equal %.4f; RMS error vs truth: mixed %.4f px, FP64 %.4f px; "
          "median: mixed %.4f, FP64 %.4f" % (agree.size, agree.mean(), rms_m, rms_f,
                                             numpy.median(err_m), numpy.median(err_f)))
    assert agree.mean() >= MIN_FLAG_AGREEMENT
    assert rms_m <= 1.05 * rms_f
