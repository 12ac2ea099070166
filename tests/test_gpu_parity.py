"""
GPU parity: the HIP library against the oracle on identical frames, start
peaks and calibration.  The bar: identical peak count and status flags;
x, y, z within 1e-3 (px / um); height, background and widths within 1e-4
relative; identical fit counters.
"""
import numpy
import pytest

from conftest import make_fitter
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.diagnostics import synth

TOL_POS = 1e-3
TOL_REL = 1e-4
COUNTERS = ("n_dposv", "n_lost", "n_margin", "n_neg_fi", "n_neg_height", "n_non_decr")


def fit_frame(backend, cls, kwds, image, peaks, peaks_type="hdf5", background=None, rqe=None, scmos=None,
              remove_error=False):
    m = make_fitter(cls, backend, rqe=rqe, scmos_cal=scmos, **kwds)
    m.initializeC(image)
    m.newImage(image)
    m.newBackground(background if background is not None else numpy.full(image.shape, 100.0))
    m.newPeaks(peaks, peaks_type)
    m.doFit()
    if remove_error:
        m.removeErrorPeaks()
    out = {"nfit": m.getNFit(), "iterations_total": m.iterations}
    for k in ("x", "y", "z", "height", "background", "xsigma", "ysigma", "error"):
        out[k] = m.getPeakProperty(k)
    out["status"] = m.getPeakProperty("status")
    out["iterations"] = m.getPeakProperty("iterations")
    out["fit_image"] = m.getFitImage()
    for c in COUNTERS:
        out[c] = getattr(m.mfit.contents, c)
    m.cleanup(verbose=False)
    return out


def assert_parity(g, r, tol_pos=TOL_POS, tol_rel=TOL_REL, counters=True):
    assert g["nfit"] == r["nfit"]
    numpy.testing.assert_array_equal(g["status"], r["status"])
    ok = r["status"] != D.ERROR
    for k in ("x", "y", "z"):
        d = numpy.abs(g[k] - r[k])[ok]
        assert d.size == 0 or d.max() < tol_pos, (k, d.max())
    for k in ("height", "background", "xsigma", "ysigma"):
        d = (numpy.abs(g[k] - r[k]) / numpy.abs(r[k]))[ok]
        assert d.size == 0 or d.max() < tol_rel, (k, d.max())
    if counters:
        numpy.testing.assert_array_equal(g["iterations"], r["iterations"])
        assert g["iterations_total"] == r["iterations_total"]
        for c in COUNTERS:
            assert g[c] == r[c], (c, g[c], r[c])
    # A frame so dense that rejected candidates overflow the fit image holds
    # inf / nan pixels: the same ones, with the same values, on both sides.
    gi, ri = g["fit_image"], r["fit_image"]
    fin = numpy.isfinite(ri)
    numpy.testing.assert_array_equal(gi[~fin], ri[~fin])
    if fin.any():
        scale = max(1.0, numpy.abs(ri[fin]).max())
        assert numpy.abs(gi[fin] - ri[fin]).max() < 1e-6 * scale


def both(*args, **kwds):
    return fit_frame("hip", *args, **kwds), fit_frame("oracle", *args, **kwds)


@pytest.mark.gpu
def test_2dfixed_mle_256():
    rng = numpy.random.default_rng(synth.SEED_BASE + 0)
    fr = synth.frame_2dfixed(rng, size=256, n=200)
    g, r = both(D.MultiFitter2DFixed, {"roi_size": 12}, fr.image, fr.peaks)
    assert_parity(g, r)
    assert (r["status"] == D.CONVERGED).sum() > 190


@pytest.mark.gpu
def test_z_mle_256():
    rng = numpy.random.default_rng(synth.SEED_BASE + 1)
    fr = synth.frame_z(rng, size=256, n=250)
    kw = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
    bg = numpy.zeros(fr.image.shape)
    g, r = both(D.MultiFitterZ, kw, fr.image, fr.peaks, background=bg)
    assert_parity(g, r)


@pytest.mark.gpu
def test_2d_dense_overlap():
    rng = numpy.random.default_rng(synth.SEED_BASE + 4)
    fr = synth.frame_2d_dense(rng, size=256, n=600)
    kw = {"roi_size": 18, "sigma_range": [0.96, 8.0]}
    g, r = both(D.MultiFitter2D, kw, fr.image, fr.peaks)
    assert_parity(g, r)


@pytest.mark.gpu
def test_3d_scmos():
    rng = numpy.random.default_rng(synth.SEED_BASE + 2)
    fr = synth.frame_3d_scmos(rng, size=128, density=1.0)
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    g, r = both(D.MultiFitter3D, kw, fr.image, fr.peaks, rqe=fr.rqe, scmos=fr.scmos_cal)
    assert_parity(g, r)


@pytest.mark.gpu
@pytest.mark.parametrize("cls", [D.MultiFitter2DFixedALS, D.MultiFitter2DFixedLS, D.MultiFitter2DFixedDWLS,
                                 D.MultiFitter2DFixedFWLS])
def test_2dfixed_error_models(cls):
    rng = numpy.random.default_rng(synth.SEED_BASE + 10)
    fr = synth.frame_2dfixed(rng, size=96, n=30)
    g, r = both(cls, {"roi_size": 12}, fr.image, fr.peaks)
    assert_parity(g, r)


@pytest.mark.gpu
@pytest.mark.parametrize("cls", [D.MultiFitter2DALS, D.MultiFitter2DLS, D.MultiFitter2DDWLS,
                                 D.MultiFitter2DFWLS])
def test_2d_error_models(cls):
    rng = numpy.random.default_rng(synth.SEED_BASE + 11)
    fr = synth.frame_2dfixed(rng, size=96, n=30)
    g, r = both(cls, {"roi_size": 18, "sigma_range": [0.75, 7.5]}, fr.image, fr.peaks)
    assert_parity(g, r)


@pytest.mark.gpu
def test_finder_start_and_removal():
    rng = numpy.random.default_rng(synth.SEED_BASE + 12)
    fr = synth.frame_2dfixed(rng, size=128, n=40)
    peaks = {"x": fr.peaks["x"], "y": fr.peaks["y"], "z": fr.peaks["z"],
             "sigma": numpy.full(fr.peaks["x"].size, 1.5)}
    kw = {"roi_size": 18, "sigma_range": [0.75, 7.5]}
    g, r = both(D.MultiFitter2D, kw, fr.image, peaks, peaks_type="finder", remove_error=True)
    assert_parity(g, r)


@pytest.mark.gpu
@pytest.mark.parametrize("frame", [315, 318])
def test_z_headline_frames_deviations_are_within_the_oracles_own_conditioning(frame):
    """
    Two frames of the benchmark workload (512 x 512, 1000 emitters, Z model,
    the bench's seeded generator) that hold the only peaks of its 500 frames
    where the FP64 GPU fit and the oracle differ by more than the parity bar:
    pairs of emitters closer than a pixel.  Every other peak meets the bar.
    For the deviating ones, the oracle run on the same image with half of its
    pixels moved by one ulp moves those peaks by as much as the GPU result
    differs: the fit there is decided by last-bit rounding (exp() and
    reduction order differ between the GPU and the host's libm), not by the
    algorithm.
    """
    import torch
    from storm_analysis_amd.diagnostics import synth_gpu
    images, bgs, rows, counts, _ = synth_gpu.frames_z(frame + 1, size=512, n=1000, device="cuda:0")
    torch.cuda.synchronize()
    img = images[frame].cpu().numpy()
    bg = bgs[frame].cpu().numpy()
    r7 = rows[frame * 1000:(frame + 1) * 1000].cpu().numpy()
    peaks = {k: r7[:, i] for i, k in enumerate(("x", "y", "z", "background", "height", "xsigma", "ysigma"))}
    kw = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
    g, r = both(D.MultiFitterZ, kw, img, peaks, background=bg)
    numpy.testing.assert_array_equal(g["status"], r["status"])
    d = numpy.hypot(g["x"] - r["x"], g["y"] - r["y"])
    off = numpy.nonzero(d > TOL_POS)[0]
    assert off.size <= 3, off
    if off.size == 0:
        return
    spread = numpy.zeros(off.size)
    rng = numpy.random.default_rng(1)
    for t in range(4):
        p = img.copy()
        m = rng.random(p.shape) < 0.5
        p[m] = numpy.nextafter(p[m], numpy.inf if t % 2 == 0 else -numpy.inf)
        r2 = fit_frame("oracle", D.MultiFitterZ, kw, p, peaks, background=bg)
        spread = numpy.maximum(spread, numpy.hypot(r2["x"] - r["x"], r2["y"] - r["y"])[off])
    print("deviating peaks", off.tolist(), "gpu-oracle", d[off].round(4).tolist(), "oracle spread", spread.round(4).tolist())
    assert numpy.all(spread >= 0.5 * d[off]), (d[off], spread)
