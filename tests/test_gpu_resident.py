"""
The resident fitter (ResidentServer over include/dao_resident_amd.h):
frames admitted into a kernel that is already running.

- Every served frame equals the same frame fitted alone by BatchFitter bit
  for bit (the resident kernel runs the same ordered executor and LM turns),
  in FP64 and mixed, with more requests than slots (slots are reused and
  requests wait for them), host and device inputs.
- FP64 frames of the serving configuration (Z model, 256^2 to 1024^2, 200
  to 4000 emitters) match the oracle to the parity bar of test_gpu_parity:
  identical status flags and per-peak iteration counts, x/y/z within 1e-3
  px, height/background/widths within 1e-4 relative.
- Teams: with team > 0 a frame of n peaks is fitted by ceil(n / team)
  workgroups (the team fitter inside the resident kernel), with the same
  results.
- The kernel's workgroups leave when idle and the next request starts it
  again; bad requests fail alone.
"""
import time

import numpy
import pytest

from storm_analysis_amd.sa_library import dao_batch
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.sa_library import dao_serve
from storm_analysis_amd.diagnostics import synth

from test_gpu_configs import assert_parity, oracle_many

ZKW = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
FIELDS = ("x", "y", "z", "height", "background", "xsigma", "ysigma", "status", "iterations")


def small_frames():
    rng = numpy.random.default_rng(synth.SEED_BASE + 60)
    spec = [(96, 20), (128, 40), (160, 60), (96, 15), (128, 30), (160, 70), (64, 5)]
    return [synth.frame_z(rng, size=s, n=n) for s, n in spec]


def batch_alone(fr, bg, precision, counters=False):
    bf = dao_batch.BatchFitter([fr.image.shape], [len(fr.peaks["x"])], model="Z", precision=precision, **ZKW)
    bf.newImages([fr.image])
    bf.newBackgrounds([bg])
    bf.newPeaks([fr.peaks], "hdf5")
    bf.doFit()
    out = bf.getPeaks(0)
    cnt = bf.getCounters(0)
    bf.cleanup()
    return (out, cnt) if counters else out


def assert_same(got, want, what):
    for k in FIELDS:
        numpy.testing.assert_array_equal(got[k], want[k], err_msg="%s: %s" % (what, k))


@pytest.mark.gpu
@pytest.mark.parametrize("team", [0, 16])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_frames_equal_the_frames_fitted_alone(precision, team):
    frames = small_frames()
    bgs = [synth.z_background(fr.image.shape) for fr in frames]
    srv = dao_serve.ResidentServer(slots=3, max_shape=(160, 160), max_peaks=80, model="Z", precision=precision,
                                   team=team, **ZKW)
    try:
        reqs = [srv.submit(fr.image, bg, dao_batch.formatPeaksGaussianPSF(fr.peaks, "hdf5"))
                for fr, bg in zip(frames, bgs)]
        got = [r.wait(120) for r in reqs]
        print("kernel launches", srv.stats())
    finally:
        srv.close()
    for k, (fr, bg) in enumerate(zip(frames, bgs)):
        want, cnt = batch_alone(fr, bg, precision, counters=True)
        assert_same(got[k], want, "request %d (%d^2)" % (k, fr.image.shape[0]))
        assert reqs[k].counters == cnt, (reqs[k].counters, cnt)


@pytest.mark.gpu
def test_device_inputs_give_the_host_inputs_result():
    import torch
    frames = small_frames()[:3]
    bgs = [synth.z_background(fr.image.shape) for fr in frames]
    srv = dao_serve.ResidentServer(slots=2, max_shape=(160, 160), max_peaks=80, model="Z", **ZKW)
    try:
        host = [srv.submit(fr.image, bg, dao_batch.formatPeaksGaussianPSF(fr.peaks, "hdf5")).wait(120)
                for fr, bg in zip(frames, bgs)]
        dev = []
        for fr, bg in zip(frames, bgs):
            rows = dao_batch.formatPeaksGaussianPSF(fr.peaks, "hdf5")
            dev.append(srv.submit(torch.from_numpy(fr.image).cuda(), torch.from_numpy(bg).cuda(),
                                  torch.from_numpy(rows).cuda()))
        dev = [r.wait(120) for r in dev]
    finally:
        srv.close()
    for k in range(len(frames)):
        assert_same(dev[k], host[k], "frame %d" % k)


@pytest.mark.gpu
@pytest.mark.parametrize("team", [0, 100])
def test_served_z_frames_256_to_1024_match_the_oracle(team):
    rng = numpy.random.default_rng(synth.SEED_BASE + 61)
    spec = [(256, 200), (512, 1000), (1024, 4000), (256, 400), (512, 2000), (1024, 2500)]
    frames = [synth.frame_z(rng, size=s, n=n) for s, n in spec]
    bgs = [synth.z_background(fr.image.shape) for fr in frames]
    srv = dao_serve.ResidentServer(slots=4, max_shape=(1024, 1024), max_peaks=4000, model="Z", team=team, **ZKW)
    try:
        reqs = [srv.submit(fr.image, bg, dao_batch.formatPeaksGaussianPSF(fr.peaks, "hdf5"))
                for fr, bg in zip(frames, bgs)]
        got = [r.wait(110) for r in reqs]
        print("device fit ms", [round(r.device_seconds * 1e3, 2) for r in reqs])
    finally:
        srv.close()
    want = oracle_many(D.MultiFitterZ, ZKW, frames, bgs)
    for k in range(len(frames)):
        assert_parity(got[k], want[k], "request %d (%d^2, %d emitters)" % ((k,) + spec[k]))


@pytest.mark.gpu
def test_kernel_restarts_after_its_workgroups_leave():
    frames = small_frames()[:2]
    bgs = [synth.z_background(fr.image.shape) for fr in frames]
    srv = dao_serve.ResidentServer(slots=2, max_shape=(160, 160), max_peaks=80, model="Z", idle_seconds=0.05, **ZKW)
    try:
        a = srv.submit(frames[0].image, bgs[0], dao_batch.formatPeaksGaussianPSF(frames[0].peaks, "hdf5")).wait(60)
        time.sleep(0.5)  # ten idle bounds: every workgroup has left
        b = srv.submit(frames[1].image, bgs[1], dao_batch.formatPeaksGaussianPSF(frames[1].peaks, "hdf5")).wait(60)
        launches = srv.stats()["kernel_launches"]
    finally:
        srv.close()
    assert launches >= 2
    assert_same(a, batch_alone(frames[0], bgs[0], "fp64"), "before")
    assert_same(b, batch_alone(frames[1], bgs[1], "fp64"), "after")


@pytest.mark.gpu
def test_bad_requests_fail_alone():
    frames = small_frames()[:2]
    bgs = [synth.z_background(fr.image.shape) for fr in frames]
    srv = dao_serve.ResidentServer(slots=2, max_shape=(160, 160), max_peaks=80, model="Z", **ZKW)
    try:
        bad = frames[0].image.copy()
        bad[3, 4] = numpy.nan
        r_bad = srv.submit(bad, bgs[0], dao_batch.formatPeaksGaussianPSF(frames[0].peaks, "hdf5"))
        with pytest.raises(D.MultiFitterException):
            srv.submit(numpy.ones((200, 200)), numpy.ones((200, 200)), None)
        r_ok = srv.submit(frames[1].image, bgs[1], dao_batch.formatPeaksGaussianPSF(frames[1].peaks, "hdf5"))
        with pytest.raises(D.MultiFitterException):
            r_bad.wait(60)
        got = r_ok.wait(60)
    finally:
        srv.close()
    assert_same(got, batch_alone(frames[1], bgs[1], "fp64"), "good request")


@pytest.mark.gpu
def test_device_frames_with_bad_pixels_are_refused_by_the_kernel():
    import torch
    frames = small_frames()[:2]
    bgs = [synth.z_background(fr.image.shape) for fr in frames]
    srv = dao_serve.ResidentServer(slots=2, max_shape=(160, 160), max_peaks=80, model="Z", **ZKW)
    try:
        reqs = []
        for k, fr in enumerate(frames):
            img = torch.from_numpy(fr.image).cuda()
            if k == 0:
                img[5, 7] = -1.0
            rows = torch.from_numpy(dao_batch.formatPeaksGaussianPSF(fr.peaks, "hdf5")).cuda()
            reqs.append(srv.submit(img, torch.from_numpy(bgs[k]).cuda(), rows))
        with pytest.raises(D.MultiFitterException):
            reqs[0].wait(60)
        got = reqs[1].wait(60)
    finally:
        srv.close()
    assert_same(got, batch_alone(frames[1], bgs[1], "fp64"), "good request")


@pytest.mark.gpu
def test_other_error_models_are_refused():
    with pytest.raises(D.MultiFitterException):
        dao_serve.ResidentServer(slots=1, max_shape=(64, 64), max_peaks=10, model="2d", error_model="LS",
                                 sigma_range=[0.75, 7.5], roi_size=18)
