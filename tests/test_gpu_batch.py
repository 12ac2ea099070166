"""
Batched fitting (BatchFitter / daoBatch*): frames of different sizes and peak
counts fitted in one executor give, frame by frame, the results of the
single-frame oracle; a frame gives the same result alone and in a mix; frames
already resident in HBM (torch tensors) are accepted.  Both sweep drivers
(the frame-per-workgroup fitter and the multi-kernel rounds) are covered.
Bars as in test_gpu_parity: identical status flags and counts; x, y, z within
1e-3; height, background and widths within 1e-4 relative.
"""
import os
from contextlib import contextmanager

import numpy
import pytest

from conftest import make_fitter
from storm_analysis_amd.sa_library import dao_batch
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.diagnostics import synth

TOL_POS = 1e-3
TOL_REL = 1e-4


@contextmanager
def executor(kind):
    """Select the sweep driver for fitters created inside the block."""
    old = os.environ.get("DAO_EXECUTOR")
    if kind == "rounds":
        os.environ["DAO_EXECUTOR"] = "rounds"
    else:
        os.environ.pop("DAO_EXECUTOR", None)
    try:
        yield
    finally:
        if old is None:
            os.environ.pop("DAO_EXECUTOR", None)
        else:
            os.environ["DAO_EXECUTOR"] = old


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
    out["n_iterations"] = m.iterations
    m.cleanup(verbose=False)
    return out


def assert_frame_parity(g, r, counters=None):
    numpy.testing.assert_array_equal(g["status"], r["status"])
    ok = r["status"] != D.ERROR
    for k in ("x", "y", "z"):
        d = numpy.abs(g[k] - r[k])[ok]
        assert d.size == 0 or d.max() < TOL_POS, (k, d.max())
    for k in ("height", "background", "xsigma", "ysigma"):
        d = (numpy.abs(g[k] - r[k]) / numpy.abs(r[k]))[ok]
        assert d.size == 0 or d.max() < TOL_REL, (k, d.max())
    numpy.testing.assert_array_equal(g["iterations"], r["iterations"])
    if counters is not None:
        assert counters["n_iterations"] == r["n_iterations"]


def scmos_frames():
    rng = numpy.random.default_rng(synth.SEED_BASE + 30)
    spec = [(64, 2.0), (128, 0.5), (96, 3.0), (128, 1.5), (64, 0.0)]
    frames = []
    for size, dens in spec:
        if dens == 0.0:
            fr = synth.frame_3d_scmos(rng, size=size, density=0.5)
            fr.peaks = {k: v[:0] for k, v in fr.peaks.items()}  # a frame with no peaks
        else:
            fr = synth.frame_3d_scmos(rng, size=size, density=dens)
        frames.append(fr)
    return frames


def batch_fit(frames, kw, backgrounds):
    shapes = [fr.image.shape for fr in frames]
    caps = [max(1, len(fr.peaks["x"])) for fr in frames]
    bf = dao_batch.BatchFitter(shapes, caps, model="3d", **kw)
    for k, fr in enumerate(frames):
        bf.setCalibration(k, rqe=fr.rqe, scmos_cal=fr.scmos_cal)
    bf.newImages([fr.image for fr in frames])
    bf.newBackgrounds(backgrounds)
    bf.newPeaks([fr.peaks for fr in frames], "hdf5")
    bf.doFit()
    out = [bf.getPeaks(k) for k in range(len(frames))]
    counters = [bf.getCounters(k) for k in range(len(frames))]
    allp = bf.getAllPeaks()
    bf.cleanup()
    return out, counters, allp


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["frames", "rounds"])
def test_batch_uneven_scmos_3d_matches_oracle(kind):
    frames = scmos_frames()
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    with executor(kind):
        out, counters, (rows, status, counts) = batch_fit(frames, kw, bgs)
    at = 0
    for k, fr in enumerate(frames):
        r = oracle_fit(D.MultiFitter3D, kw, fr, bgs[k])
        assert out[k]["x"].size == r["x"].size
        assert_frame_parity(out[k], r, counters[k])
        # the all-frames read-back is the per-frame one, concatenated
        n = int(counts[k])
        numpy.testing.assert_array_equal(status[at:at + n], out[k]["status"])
        numpy.testing.assert_array_equal(rows[at:at + n, 0], out[k]["x"])
        at += n


@pytest.mark.gpu
def test_batch_same_alone_and_in_mix():
    """
    Each frame's result in a mixed batch is the frame's result alone, to the
    parity bar.  (Bit-identical unless a reserved box overflowed: the rollback
    that follows undoes the ops that had finished by then with arithmetic
    subtraction, and which ops had finished depends on timing.)
    """
    frames = scmos_frames()
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    mixed, mixed_c, _ = batch_fit(frames, kw, bgs)
    for k in (0, 2, 3):
        alone, alone_c, _ = batch_fit([frames[k]], kw, [bgs[k]])
        assert alone[0]["x"].size == mixed[k]["x"].size
        assert_frame_parity(alone[0], mixed[k])
        assert alone_c[0]["n_iterations"] == mixed_c[k]["n_iterations"]


@pytest.mark.gpu
def test_batch_device_resident_z_frames():
    """Frames generated in HBM (torch) go in without a host copy and match the oracle."""
    import torch
    from storm_analysis_amd.diagnostics import synth_gpu
    F, S, N = 3, 128, 40
    images, background, rows, counts, truth = synth_gpu.frames_z(F, size=S, n=N, device="cuda:0")
    kw = {"wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
    bf = dao_batch.BatchFitter([(S, S)] * F, [N] * F, model="Z", roi_size=18, **kw)
    bf.newImages(images)
    bf.newBackgrounds(background)
    bf.newPeaks(rows, "hdf5", counts=counts.numpy())
    bf.doFit()
    out = [bf.getPeaks(k) for k in range(F)]
    bf.cleanup()
    torch.cuda.synchronize()
    host_rows = rows.cpu().numpy()
    c0 = 0
    for k in range(F):
        r7 = host_rows[c0:c0 + N]
        c0 += N
        fr = synth.Frame(images[k].cpu().numpy(), {"x": r7[:, 0], "y": r7[:, 1], "z": r7[:, 2],
                                                   "background": r7[:, 3], "height": r7[:, 4],
                                                   "xsigma": r7[:, 5], "ysigma": r7[:, 6]}, None)
        r = oracle_fit(D.MultiFitterZ, dict(kw, roi_size=18), fr, background[k].cpu().numpy())
        assert_frame_parity(out[k], r)


@pytest.mark.gpu
def test_tuning_keeps_results_and_autotune_picks_a_candidate():
    """Waves per workgroup, the reserved-box margin and the team size change speed, not results."""
    from storm_analysis_amd.sa_library import dao_tune
    frames = scmos_frames()
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    shapes = [fr.image.shape for fr in frames]
    caps = [max(1, len(fr.peaks["x"])) for fr in frames]
    bf = dao_batch.BatchFitter(shapes, caps, model="3d", **kw)

    def prepare():
        for k, fr in enumerate(frames):
            bf.setCalibration(k, rqe=fr.rqe, scmos_cal=fr.scmos_cal)
        bf.newImages([fr.image for fr in frames])
        bf.newBackgrounds(bgs)
        bf.newPeaks([fr.peaks for fr in frames], "hdf5")

    results = {}
    for waves, margin, team in ((0, 2, 0), (2, 1, 0), (4, 3, 0), (8, 2, 0), (4, 1, 30), (8, 2, 60)):
        bf.setTuning(waves, margin, team)
        prepare()
        bf.doFit()
        results[(waves, margin, team)] = [bf.getPeaks(k) for k in range(len(frames))]
    ref = results[(0, 2, 0)]
    for key, res in results.items():
        for g, r in zip(res, ref):
            assert_frame_parity(g, r)
    best, trials = dao_tune.autotune(bf, prepare, waves=(2, 8), margins=(1, 2), teams=(0, 40))
    print("autotune trials:", trials)
    assert best in [(t["waves"], t["margin"], t["team"]) for t in trials] and len(trials) == 8
    bf.cleanup()


@pytest.mark.gpu
def test_new_frames_in_one_pass_equal_image_then_background():
    """daoBatchNewFrames (image and background in one pass) gives the fit of newImages + newBackgrounds,
    from host and from device arrays; getAllPeaks fills caller buffers (pinned host memory) as it fills
    its own arrays."""
    import torch
    frames = scmos_frames()
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    shapes = [fr.image.shape for fr in frames]
    caps = [max(1, len(fr.peaks["x"])) for fr in frames]

    def fit(load):
        bf = dao_batch.BatchFitter(shapes, caps, model="3d", **kw)
        for k, fr in enumerate(frames):
            bf.setCalibration(k, rqe=fr.rqe, scmos_cal=fr.scmos_cal)
        load(bf)
        bf.newPeaks([fr.peaks for fr in frames], "hdf5")
        bf.doFit()
        own = bf.getAllPeaks()
        n = own[0].shape[0]
        rows = torch.empty((n + 5, 9), dtype=torch.float64, pin_memory=True).numpy()
        status = torch.empty(n + 5, dtype=torch.int32, pin_memory=True).numpy()
        given = bf.getAllPeaks(rows, status)
        bf.cleanup()
        numpy.testing.assert_array_equal(own[0], given[0])
        numpy.testing.assert_array_equal(own[1], given[1])
        return own

    def two_calls(bf):
        bf.newImages([fr.image for fr in frames])
        bf.newBackgrounds(bgs)

    def one_call_host(bf):
        bf.newFrames([fr.image for fr in frames], bgs)

    def one_call_device(bf):
        img = torch.from_numpy(numpy.concatenate([fr.image.ravel() for fr in frames])).cuda()
        bg = torch.from_numpy(numpy.concatenate([b.ravel() for b in bgs])).cuda()
        bf.newFrames(img, bg)

    ref = fit(two_calls)
    for load in (one_call_host, one_call_device):
        got = fit(load)
        numpy.testing.assert_array_equal(got[0], ref[0])
        numpy.testing.assert_array_equal(got[1], ref[1])
