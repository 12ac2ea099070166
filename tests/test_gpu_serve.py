"""
Serving (FitServer): frames of different sizes submitted together are fitted
in shared launches; each result equals the frame fitted alone (FP64: the
single-frame C-ABI path; mixed: a one-frame batch), to the parity bar of
test_gpu_parity (flags and iteration counts equal, x/y/z within 1e-3,
height/background/widths within 1e-4 relative).  Slot re-sizing,
admission limits and the HBM budget are covered too.
"""
import numpy
import pytest

from storm_analysis_amd.sa_library import dao_batch
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.sa_library import dao_serve
from storm_analysis_amd.diagnostics import synth

TOL_POS = 1e-3
TOL_REL = 1e-4
ZKW = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}


def z_frames():
    rng = numpy.random.default_rng(synth.SEED_BASE + 50)
    spec = [(96, 20), (128, 40), (160, 60), (96, 15), (128, 30), (160, 70), (64, 5)]
    return [synth.frame_z(rng, size=s, n=n) for s, n in spec]


def rows7(fr):
    p = fr.peaks
    return numpy.stack([p["x"], p["y"], p["z"], p["background"], p["height"], p["xsigma"], p["ysigma"]], axis=1)


def background(fr):
    return numpy.full(fr.image.shape, 125.0)


def alone_plain(fr):
    m = D.MultiFitterZ(**ZKW)
    m.initializeC(fr.image)
    m.newImage(fr.image)
    m.newBackground(background(fr))
    m.newPeaks(fr.peaks, "hdf5")
    m.doFit()
    out = {k: m.getPeakProperty(k) for k in ("x", "y", "z", "height", "background", "xsigma", "ysigma")}
    out["status"] = m.getPeakProperty("status")
    out["iterations"] = m.getPeakProperty("iterations")
    m.cleanup(verbose=False)
    return out


def alone_batch(fr, precision):
    bf = dao_batch.BatchFitter([fr.image.shape], [len(fr.peaks["x"])], model="Z", precision=precision, **ZKW)
    bf.newImages([fr.image])
    bf.newBackgrounds([background(fr)])
    bf.newPeaks([fr.peaks], "hdf5")
    bf.doFit()
    out = bf.getPeaks(0)
    bf.cleanup()
    return out


def assert_same(g, r):
    numpy.testing.assert_array_equal(g["status"], r["status"])
    numpy.testing.assert_array_equal(g["iterations"], r["iterations"])
    ok = r["status"] != D.ERROR
    for k in ("x", "y", "z"):
        d = numpy.abs(g[k] - r[k])[ok]
        assert d.size == 0 or d.max() < TOL_POS, (k, d.max())
    for k in ("height", "background", "xsigma", "ysigma"):
        d = (numpy.abs(g[k] - r[k]) / numpy.abs(r[k]))[ok]
        assert d.size == 0 or d.max() < TOL_REL, (k, d.max())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_equals_alone(precision):
    frames = z_frames()
    srv = dao_serve.FitServer(workers=2, slots=3, max_shape=(160, 160), max_peaks=80, model="Z",
                              precision=precision, **ZKW)
    try:
        reqs = [srv.submit(fr.image, background(fr), rows7(fr)) for fr in frames]
        # the same frames again, now sharing launches with different neighbours
        reqs += [srv.submit(fr.image, background(fr), rows7(fr)) for fr in frames[::-1]]
        res = [r.wait(timeout=120) for r in reqs]
        summ = srv.summary()
    finally:
        srv.close()
    print("served %d frames in %d launches (mean batch %.2f), p50 %.1f ms, p99 %.1f ms" %
          (summ["requests"], summ["launches"], summ["mean_batch"], summ["p50_ms"], summ["p99_ms"]))
    assert summ["errors"] == 0 and summ["requests"] == 2 * len(frames)
    n = len(frames)
    for k, fr in enumerate(frames):
        ref = alone_plain(fr) if precision == "fp64" else alone_batch(fr, precision)
        assert_same(res[k], ref)
        assert_same(res[2 * n - 1 - k], ref)


@pytest.mark.gpu
def test_server_admission_limits():
    srv = dao_serve.FitServer(workers=1, slots=2, max_shape=(64, 64), max_peaks=10, model="Z", **ZKW)
    try:
        with pytest.raises(D.MultiFitterException):
            srv.submit(numpy.ones((65, 64)), numpy.ones((65, 64)), numpy.zeros((1, 7)))
        with pytest.raises(D.MultiFitterException):
            srv.submit(numpy.ones((64, 64)), numpy.ones((64, 64)), numpy.zeros((11, 7)))
    finally:
        srv.close()


def test_hbm_budget_sizes_the_pool():
    # an 18-pixel fitting area holds a round ROI of 256 pixels: 4 per lane,
    # so a journal snapshot is (3 * 4 + 1) * 64 words
    assert dao_serve.roi_pixels(18) == 256
    assert dao_serve.journal_bytes(18) == dao_serve.JOURNAL_SNAPSHOTS * 13 * 64 * 8 == 26624
    per_slot = dao_serve.slot_bytes((1024, 1024), 4000)
    assert per_slot == 1024 * 1024 * dao_serve.BYTES_PER_PIXEL + 4000 * (dao_serve.BYTES_PER_PEAK + 26624)
    # no budget: as asked
    assert dao_serve.pool_size(4, 32, (1024, 1024), 4000) == (4, 32)
    # 2 GB holds 10 such slots: fewer slots per worker first, then fewer workers
    w, s = dao_serve.pool_size(4, 32, (1024, 1024), 4000, hbm_budget_gb=2.0)
    assert w * s <= int(2e9 // per_slot) and (w, s) == (4, 2)
    w, s = dao_serve.pool_size(4, 32, (1024, 1024), 4000, hbm_budget_gb=0.2)
    assert (w, s) == (1, 1)
    # smaller ROIs need smaller journals
    assert dao_serve.slot_bytes((1024, 1024), 4000, roi_size=10) < per_slot
    with pytest.raises(D.MultiFitterException):
        dao_serve.pool_size(4, 32, (1024, 1024), 4000, hbm_budget_gb=0.01)
