"""
Failure handling: no call of the library aborts the process.

CPU (no GPU needed): a call that fails inside the C-ABI (here: creating a
fitter where no GPU exists) returns, leaves its code and message for
daoGetLastError(), and the Python surface raises MultiFitterException; the
host-side checks of BatchFitter.newPeaks (capacity counting the peaks a frame
already holds, one count per frame) and of FitServer's admission (non-finite
or non-positive pixels) refuse bad input before any device work.

GPU (-m gpu): the C side refuses a capacity overflow with nothing changed;
a scheduler wait made to give up from the host (a tiny wait limit, so every
kernel still exits normally) fails only the frames it happened in -- the
other frames of the launch equal a normal run exactly -- on both fitters
(one workgroup per frame, and teams); the single-frame handle raises and
stays usable; FitServer fails only the affected requests.
"""
import ctypes

import numpy
import pytest

from conftest import HIP_SO
from storm_analysis_amd.sa_library import dao_batch
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.sa_library import dao_serve
from storm_analysis_amd.diagnostics import synth


def _no_gpu():
    try:
        import torch
        return not torch.cuda.is_available()
    except ImportError:
        return True


@pytest.mark.skipif(not _no_gpu(), reason="exercises the no-device error path")
def test_failed_call_raises_with_the_library_message():
    """daoBatchCreate / daoInitialize without a device: NULL, last error set, Python raises."""
    lib = D.setDaoFitArgtypes(ctypes.cdll.LoadLibrary(HIP_SO))
    code, _ = D.lastLibraryError(lib)
    assert code == 0
    with pytest.raises(D.MultiFitterException, match="no GPU"):
        dao_batch.BatchFitter([(64, 64)], [10], model="2dfixed", roi_size=12)
    m = D.MultiFitter2DFixed(roi_size=12)
    with pytest.raises(D.MultiFitterException, match="no GPU"):
        m.initializeC(numpy.ones((32, 32)))
    # the error was consumed by the raise
    assert D.lastLibraryError(lib)[0] == 0


class _FakeBatchLib(object):
    """Records daoBatchNewPeaks calls; frames hold `held` peaks."""
    def __init__(self, held):
        self.held = list(held)
        self.calls = 0

    def daoBatchNFit(self, handle, f):
        return self.held[f]

    def daoBatchNewPeaks(self, *args):
        self.calls += 1


def _fake_fitter(caps, held):
    bf = dao_batch.BatchFitter.__new__(dao_batch.BatchFitter)
    bf.clib = _FakeBatchLib(held)
    bf.handle = None
    bf.n_frames = len(caps)
    bf.caps = numpy.asarray(caps, dtype=numpy.int32)
    return bf


def _rows(n):
    return numpy.zeros((n, 7))


def test_new_peaks_counts_the_peaks_a_frame_holds():
    """The find-fit loop adds peaks on every pass: held + new must fit the capacity."""
    bf = _fake_fitter([10, 10], held=[8, 0])
    bf.newPeaks(_rows(2 + 10), "hdf5", counts=[2, 10])   # 8 + 2 and 0 + 10 fit
    assert bf.clib.calls == 1
    with pytest.raises(D.MultiFitterException, match="capacity"):
        bf.newPeaks(_rows(3), "hdf5", counts=[3, 0])     # 8 + 3 > 10
    with pytest.raises(D.MultiFitterException, match="one entry per frame"):
        bf.newPeaks(_rows(3), "hdf5", counts=[3])
    with pytest.raises(D.MultiFitterException, match="capacity"):
        bf.newPeaks(_rows(0), "hdf5", counts=[-1, 1])
    assert bf.clib.calls == 1


def test_server_admission_refuses_bad_frames():
    ok = numpy.full((8, 8), 10.0)
    assert dao_serve._invalid_input(ok, ok, numpy.ones((3, 7))) is None
    bad = ok.copy()
    bad[2, 3] = numpy.nan
    assert "pixels" in dao_serve._invalid_input(bad, ok, None)
    neg = ok.copy()
    neg[0, 0] = -1.0
    assert "pixels" in dao_serve._invalid_input(neg, ok, None)
    zero = ok.copy()
    zero[0, 0] = 0.0
    assert "pixels" in dao_serve._invalid_input(zero, ok, None)
    assert "background" in dao_serve._invalid_input(ok, bad, None)
    rows = numpy.ones((2, 7))
    rows[1, 4] = numpy.inf
    assert "start peak" in dao_serve._invalid_input(ok, ok, rows)


# ---------------------------------------------------------------------------
# GPU


def _frames(seed=31):
    """Two dense 3D frames (many overlapping ops: waits certain) and three sparse ones."""
    rng = numpy.random.default_rng(synth.SEED_BASE + seed)
    spec = [(96, 3.0), (64, 0.05), (96, 2.5), (64, 0.05), (64, 0.1)]
    return [synth.frame_3d_scmos(rng, size=s, density=d) for s, d in spec]


def _batch(frames, team=0):
    bf = dao_batch.BatchFitter([f.image.shape for f in frames], [f.peaks["x"].size + 8 for f in frames],
                               model="3d", roi_size=18, sigma_range=[0.6, 5.0])
    for k, f in enumerate(frames):
        bf.setCalibration(k, f.rqe, f.scmos_cal)
    bf.newImages([f.image for f in frames])
    bf.newBackgrounds([numpy.full(f.image.shape, 100.0) for f in frames])
    bf.newPeaks([f.peaks for f in frames], "hdf5")
    bf.setTuning(0, -1, team)
    return bf


def _results(bf):
    return [bf.getPeaks(f) for f in range(bf.n_frames)]


def _same(a, b):
    for k in ("x", "y", "z", "height", "background", "xsigma", "ysigma", "error", "iterations", "status"):
        numpy.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.gpu
def test_capacity_overflow_is_refused_on_the_device_side():
    frames = _frames()[:2]
    bf = _batch(frames)
    held = [bf.getNFit(f) for f in range(2)]
    spare = [int(bf.caps[f]) - held[f] for f in range(2)]
    # bypass the Python check: the C side refuses with nothing changed
    counts = numpy.array([spare[0] + 1, 0], dtype=numpy.int32)
    rows = numpy.ascontiguousarray(numpy.repeat(dao_batch.formatPeaksGaussianPSF(frames[0].peaks, "hdf5")[:1],
                                                counts[0], axis=0))
    with pytest.raises(D.MultiFitterException, match="capacity"):
        bf.clib.daoBatchNewPeaks(bf.handle, rows.ctypes.data_as(ctypes.c_void_p), counts, b"hdf5", 0)
    assert [bf.getNFit(f) for f in range(2)] == held
    with pytest.raises(D.MultiFitterException, match="capacity"):
        bf.newPeaks(rows, "hdf5", counts=counts)
    # the batch still fits normally
    bf.doFit()
    assert all(bf.getNFit(f) == held[f] for f in range(2))
    bf.cleanup()


@pytest.mark.gpu
@pytest.mark.parametrize("team", [0, 8])
def test_forced_wait_give_up_fails_only_its_frames(team):
    frames = _frames()
    ref = _batch(frames, team)
    ref.doFit()
    want = _results(ref)
    ref.cleanup()

    bf = _batch(frames, team)
    bf.setWaitLimit(50)
    with pytest.raises(dao_batch.FrameFitError) as info:
        bf.doFit()
    failed = info.value.frames
    assert 0 in failed and 2 in failed, failed    # the dense frames had to wait
    got = _results(bf)
    for f in range(len(frames)):
        if f in failed:
            assert bf.getNFit(f) == 0                 # emptied, not half-fitted
        else:
            _same(got[f], want[f])
    # the batch is usable again: reload the failed frames and fit normally
    bf.setWaitLimit(0)
    bf.newImages([f.image for f in frames])
    bf.newBackgrounds([numpy.full(f.image.shape, 100.0) for f in frames])
    bf.newPeaks([f.peaks for f in frames], "hdf5")
    bf.doFit()
    for f, r in enumerate(_results(bf)):
        _same(r, want[f])
    bf.cleanup()


@pytest.mark.gpu
def test_single_frame_handle_raises_and_stays_usable():
    fr = _frames()[0]
    def fitter():
        m = D.MultiFitter3D(roi_size=18, sigma_range=[0.6, 5.0], rqe=fr.rqe, scmos_cal=fr.scmos_cal)
        m.initializeC(fr.image)
        m.newImage(fr.image)
        m.newBackground(numpy.full(fr.image.shape, 100.0))
        m.newPeaks(fr.peaks, "hdf5")
        return m
    ref = fitter()
    ref.doFit()
    want = {k: ref.getPeakProperty(k) for k in ("x", "y", "status", "iterations")}
    ref.cleanup(verbose=False)

    m = fitter()
    m.setWaitLimit(50)
    with pytest.raises(D.MultiFitterException, match="bound"):
        m.doFit()
    assert m.getNFit() == 0
    m.setWaitLimit(0)
    m.newImage(fr.image)
    m.newPeaks(fr.peaks, "hdf5")
    m.doFit()
    for k, v in want.items():
        numpy.testing.assert_array_equal(m.getPeakProperty(k), v, err_msg=k)
    m.cleanup(verbose=False)


@pytest.mark.gpu
def test_server_fails_only_the_affected_requests():
    frames = _frames()
    kw = dict(model="3d", roi_size=18, sigma_range=[0.6, 5.0])
    srv = dao_serve.FitServer(workers=1, slots=8, max_shape=(96, 96), max_peaks=800, **kw)
    try:
        bg = numpy.full((64, 64), 100.0)
        bad = frames[1].image.copy()
        bad[5, 5] = numpy.nan
        rows = dao_batch.formatPeaksGaussianPSF(frames[1].peaks, "hdf5")
        r_bad = srv.submit(bad, bg, rows)
        assert r_bad.done.is_set()
        with pytest.raises(D.MultiFitterException, match="non-finite"):
            r_bad.wait(1)
        # healthy frames alone
        alone = [srv.submit(f.image, numpy.full(f.image.shape, 100.0),
                            dao_batch.formatPeaksGaussianPSF(f.peaks, "hdf5")).wait(60) for f in frames]
        # a tiny wait limit: the dense frames fail, the others return their results
        srv.setWaitLimit(50)
        reqs = [srv.submit(f.image, numpy.full(f.image.shape, 100.0),
                           dao_batch.formatPeaksGaussianPSF(f.peaks, "hdf5")) for f in frames]
        for r in reqs:
            r.done.wait(60)
        errors = [k for k, r in enumerate(reqs) if r.error is not None]
        for k in errors:
            assert isinstance(reqs[k].error, dao_batch.FrameFitError)
        for k, r in enumerate(reqs):
            if r.error is None:
                _same(r.result, alone[k])
        srv.setWaitLimit(0)
        again = srv.submit(frames[0].image, numpy.full(frames[0].image.shape, 100.0),
                           dao_batch.formatPeaksGaussianPSF(frames[0].peaks, "hdf5")).wait(60)
        _same(again, alone[0])
    finally:
        srv.close()
