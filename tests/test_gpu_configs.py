"""
The benchmark configurations at their real sizes, FP64 GPU against the
oracle run on the host's cores (one oracle fitter per frame, in a thread
pool: ctypes releases the GIL).  Bars as in test_gpu_parity: identical
peak counts, status flags and per-peak iteration counts; x, y, z within
1e-3 px; height, background and widths within 1e-4 relative.

- Uneven sCMOS batch ('3d' model, free widths): 128^2 frames at 3 and 0.5
  emitters/um^2, a 256^2 frame at 1/um^2, a 512^2 frame at 0.3/um^2 and a
  2048^2 frame at 0.1/um^2, with per-pixel offset, gain and read variance,
  in one persistent launch (and once more with teams of workgroups).
- Serving: Z-model frames of 256^2 to 1024^2 with 200 to 4000 emitters
  submitted together to a FitServer (shared launches, slots re-sized).
- High-density overlap stress: one 2048^2 frame with 20,000 emitters, 2D
  model with free width, sigma U(1.2, 2.0), by one workgroup and by a team.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy
import pytest

from conftest import make_fitter
from storm_analysis_amd.sa_library import dao_batch
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.sa_library import dao_serve
from storm_analysis_amd.diagnostics import synth

TOL_POS = 1e-3
TOL_REL = 1e-4
POOL = min(16, os.cpu_count() or 1)


def oracle_fit(cls, kw, fr, background):
    m = make_fitter(cls, "oracle", rqe=fr.rqe, scmos_cal=fr.scmos_cal, **kw)
    m.initializeC(fr.image)
    m.newImage(fr.image)
    m.newBackground(background)
    m.newPeaks(fr.peaks, "hdf5")
    m.doFit()
    out = {k: m.getPeakProperty(k) for k in ("x", "y", "z", "height", "background", "xsigma", "ysigma")}
    out["status"] = m.getPeakProperty("status")
    out["iterations"] = m.getPeakProperty("iterations")
    m.cleanup(verbose=False)
    return out


def oracle_many(cls, kw, frames, backgrounds):
    with ThreadPoolExecutor(POOL) as ex:
        return list(ex.map(lambda a: oracle_fit(cls, kw, a[0], a[1]), zip(frames, backgrounds)))


def assert_parity(g, r, what=""):
    assert g["x"].size == r["x"].size, what
    numpy.testing.assert_array_equal(g["status"], r["status"], err_msg=what)
    numpy.testing.assert_array_equal(g["iterations"], r["iterations"], err_msg=what)
    ok = r["status"] != D.ERROR
    for k in ("x", "y", "z"):
        d = numpy.abs(g[k] - r[k])[ok]
        assert d.size == 0 or d.max() < TOL_POS, (what, k, d.max())
    for k in ("height", "background", "xsigma", "ysigma"):
        d = (numpy.abs(g[k] - r[k]) / numpy.abs(r[k]))[ok]
        assert d.size == 0 or d.max() < TOL_REL, (what, k, d.max())


def uneven_frames():
    rng = numpy.random.default_rng(synth.SEED_BASE + 3)
    spec = [(128, 3.0), (128, 0.5), (256, 1.0), (512, 0.3), (2048, 0.1)]
    return [synth.frame_3d_scmos(rng, size=s, density=d) for s, d in spec]


@pytest.mark.gpu
@pytest.mark.parametrize("team", [0, 500])
def test_uneven_scmos_batch_with_a_2048_frame_matches_the_oracle(team):
    frames = uneven_frames()
    kw = {"roi_size": 18, "sigma_range": [0.6, 5.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    bf = dao_batch.BatchFitter([fr.image.shape for fr in frames], [fr.peaks["x"].size for fr in frames],
                               model="3d", **kw)
    bf.setTuning(0, -1, team)
    for k, fr in enumerate(frames):
        bf.setCalibration(k, rqe=fr.rqe, scmos_cal=fr.scmos_cal)
    bf.newImages([fr.image for fr in frames])
    bf.newBackgrounds(bgs)
    bf.newPeaks([fr.peaks for fr in frames], "hdf5")
    bf.doFit()
    got = [bf.getPeaks(k) for k in range(len(frames))]
    bf.cleanup()
    want = oracle_many(D.MultiFitter3D, kw, frames, bgs)
    print("peaks per frame", [fr.peaks["x"].size for fr in frames])
    for k in range(len(frames)):
        assert_parity(got[k], want[k], "frame %d (%d^2)" % (k, frames[k].image.shape[0]))


@pytest.mark.gpu
def test_served_z_frames_256_to_1024_match_the_oracle():
    rng = numpy.random.default_rng(synth.SEED_BASE + 4)
    spec = [(256, 200), (512, 1000), (1024, 4000), (256, 400), (512, 2000), (1024, 2500)]
    frames = [synth.frame_z(rng, size=s, n=n) for s, n in spec]
    kw = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
    bgs = [synth.z_background(fr.image.shape) for fr in frames]
    srv = dao_serve.FitServer(workers=2, slots=4, max_shape=(1024, 1024), max_peaks=4000, model="Z", **kw)
    try:
        reqs = [srv.submit(fr.image, bg, dao_batch.formatPeaksGaussianPSF(fr.peaks, "hdf5"))
                for fr, bg in zip(frames, bgs)]
        got = [r.wait(110) for r in reqs]
    finally:
        srv.close()
    want = oracle_many(D.MultiFitterZ, kw, frames, bgs)
    for k in range(len(frames)):
        assert_parity(got[k], want[k], "request %d (%d^2, %d emitters)" % ((k,) + spec[k]))


@pytest.mark.gpu
def test_dense_2048_frame_with_20000_emitters_matches_the_oracle():
    rng = numpy.random.default_rng(synth.SEED_BASE + 5)
    fr = synth.frame_2d_dense(rng, size=2048, n=20000)
    kw = {"roi_size": 18, "sigma_range": [0.6 * 1.6, 5.0 * 1.6]}
    bg = numpy.full(fr.image.shape, 100.0)
    with ThreadPoolExecutor(1) as ex:
        pending = ex.submit(oracle_fit, D.MultiFitter2D, kw, fr, bg)
        got = {}
        for team in (0, 500):
            m = D.MultiFitter2D(team=team, **kw)
            m.initializeC(fr.image)
            m.newImage(fr.image)
            m.newBackground(bg)
            m.newPeaks(fr.peaks, "hdf5")
            m.doFit()
            got[team] = {k: m.getPeakProperty(k) for k in ("x", "y", "z", "height", "background", "xsigma",
                                                           "ysigma", "status", "iterations")}
            m.cleanup(verbose=False)
        want = pending.result()
    for team, g in got.items():
        assert_parity(g, want, "team %d" % team)
