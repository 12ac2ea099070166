"""
Mixed precision (BatchFitter precision="mixed"): FP32 deviance terms,
Jacobian and Hessian; FP64 peak shape, accumulators, fi - x, solve; each
LM candidate judged on the summed per-pixel deviance change (exact at the
scale of the 1e-6 tolerance).  Checked against the FP64 oracle on the same
frames and start peaks, with the FP64 HIP fitter's LM cycle count beside:

- Z frames at 512^2 with 1000 emitters: identical flags; over peaks
  CONVERGED in both, the 99th percentile of |dx|, |dy| within 1e-5 px and
  at least 99.8% within 6.25e-3 px (1 nm at 160 nm pixels); every peak
  beyond 1 nm has another fitted peak within 2 px (an unresolved pair,
  whose split is weakly determined, so the FP32 Jacobian's slightly
  different steps stop it elsewhere; measured: 5 of 4000, neighbours
  0.26-1.76 px away, while 88% of all peaks have none within 3 px); LM
  cycles within 2% of FP64's.
- Dense free-width 2D: identical flags, every peak within 1e-5 px and
  1e-4 relative, LM cycles within 1% of FP64's (measured: equal).
- Dense free-width 3D with sCMOS noise, against the true positions (FP64
  itself crawls there; see its docstring), LM cycles within 1.2x of FP64's.
The measured figures are printed.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy
import pytest

from conftest import make_fitter
from storm_analysis_amd.sa_library import dao_batch
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.diagnostics import synth

NM_PX = 6.25e-3      # 1 nm at 160 nm pixels
PAIR_PX = 2.0        # an exception's neighbour lies this close
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


def batch_fit(model, kw, frames, backgrounds, precision):
    bf = dao_batch.BatchFitter([fr.image.shape for fr in frames], [len(fr.peaks["x"]) for fr in frames],
                               model=model, precision=precision, **kw)
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


def mixed_fit(model, kw, frames, backgrounds):
    return batch_fit(model, kw, frames, backgrounds, "mixed")


def cycles(results):
    return int(sum(int(r["iterations"].sum()) for r in results))


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


def summary(results, refs):
    parts = [differences(g, r) for g, r in zip(results, refs)]
    agree, dxy, dz, drel = (numpy.concatenate([p[i] for p in parts]) for i in range(4))
    q = lambda v: (numpy.quantile(v, 0.5), numpy.quantile(v, QUANTILE), v.max())  # noqa: E731
    print("mixed vs FP64 oracle: %d peaks, flags equal %.4f; within 1 nm %.4f; "
          "dxy px p50/p99/max %.1e/%.1e/%.1e; dz um %.1e/%.1e/%.1e; rel %.1e/%.1e/%.1e" %
          ((agree.size, agree.mean(), (dxy < NM_PX).mean()) + q(dxy) + q(dz) + q(drel)))
    return agree, dxy, dz, drel


def unresolved_pairs_only(results, refs):
    """Every peak beyond 1 nm has another fitted (oracle) peak within PAIR_PX."""
    for g, r in zip(results, refs):
        both = (g["status"] == D.CONVERGED) & (r["status"] == D.CONVERGED)
        dxy = numpy.maximum(numpy.abs(g["x"] - r["x"]), numpy.abs(g["y"] - r["y"]))
        far = numpy.nonzero(both & (dxy >= NM_PX))[0]
        for i in far:
            d = numpy.hypot(r["x"] - r["x"][i], r["y"] - r["y"][i])
            d[i] = numpy.inf
            print("  beyond 1 nm: peak %d, %.2e px off, nearest neighbour %.2f px" % (i, dxy[i], d.min()))
            assert d.min() < PAIR_PX, (i, dxy[i], d.min())


@pytest.mark.gpu
def test_mixed_z_512_1000_against_oracle():
    rng = numpy.random.default_rng(synth.SEED_BASE + 2)
    frames = [synth.frame_z(rng, size=512, n=1000) for _ in range(2)]
    kw = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
    bgs = [synth.z_background(fr.image.shape) for fr in frames]
    out = mixed_fit("Z", kw, frames, bgs)
    fp64 = batch_fit("Z", kw, frames, bgs, "fp64")
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        refs = list(ex.map(lambda a: oracle_fit(D.MultiFitterZ, kw, a[0], a[1]), zip(frames, bgs)))
    agree, dxy, dz, drel = summary(out, refs)
    print("LM cycles: mixed %d, FP64 %d" % (cycles(out), cycles(fp64)))
    assert agree.all()
    assert numpy.quantile(dxy, QUANTILE) < 1e-5
    assert (dxy < NM_PX).mean() >= 0.998
    unresolved_pairs_only(out, refs)
    assert cycles(out) <= 1.02 * cycles(fp64)


@pytest.mark.gpu
def test_mixed_2d_dense_against_oracle():
    rng = numpy.random.default_rng(synth.SEED_BASE + 41)
    fr = synth.frame_2d_dense(rng, size=192, n=500)
    kw = {"roi_size": 18, "sigma_range": [0.6, 5.0]}
    bg = numpy.full(fr.image.shape, 100.0)
    out = mixed_fit("2d", kw, [fr], [bg])
    fp64 = batch_fit("2d", kw, [fr], [bg], "fp64")
    ref = oracle_fit(D.MultiFitter2D, kw, fr, bg)
    agree, dxy, dz, drel = summary(out, [ref])
    print("LM cycles: mixed %d, FP64 %d" % (cycles(out), cycles(fp64)))
    assert agree.all()
    assert dxy.max() < 1e-5 and drel.max() < 1e-4
    assert abs(cycles(out) - cycles(fp64)) <= 0.01 * cycles(fp64)


@pytest.mark.gpu
def test_mixed_3d_scmos_dense_against_truth():
    """
    Dense free-width 3D with sCMOS noise (up to 2.5 emitters/um^2): here the
    FP64 fits themselves crawl (~220 LM cycles per peak) and stop on flat
    valleys, so mixed and FP64 stop at different points (median ~0.03 px
    apart) and parity to FP64 is not the measure.  The bar instead: the same
    flags for >= 99% of peaks, localization error against the true emitter
    positions no worse than FP64's (RMS within 5%), and LM cycles within
    1.2x of FP64's.
    """
    rng = numpy.random.default_rng(synth.SEED_BASE + 42)
    frames = [synth.frame_3d_scmos(rng, size=128, density=d) for d in (1.0, 2.5)]
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    out = mixed_fit("3d", kw, frames, bgs)
    fp64 = batch_fit("3d", kw, frames, bgs, "fp64")
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
          "median: mixed %.4f, FP64 %.4f; LM cycles: mixed %d, FP64 %d" %
          (agree.size, agree.mean(), rms_m, rms_f, numpy.median(err_m), numpy.median(err_f), cycles(out),
           cycles(fp64)))
    assert agree.mean() >= MIN_FLAG_AGREEMENT
    assert rms_m <= 1.05 * rms_f
    assert cycles(out) <= 1.2 * cycles(fp64)
