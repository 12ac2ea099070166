"""
The frame-per-team fitter: several workgroups, on as many CUs, fit one frame
(daoSetTeam / MultiFitter.setTeam for one frame, setTuning(team=...) for a
batch).  Its scheduling state and the frame's pixel and peak records are
shared through device-coherent accesses; the op order rule is the one of
the frame-per-workgroup fitter.  Checked: the oracle's results to the parity
bar (identical flags and counts; x, y, z within 1e-3; height, background and
widths within 1e-4 relative; identical fit counters), on sparse, dense and
sCMOS frames, with zero reserved-box margin (many escapes, so the team's
rollback runs), in FP64 and in mixed precision (against the one-workgroup
mixed fitter), and for a batch of uneven frames.
"""
import numpy
import pytest

from test_gpu_parity import TOL_POS, assert_parity, fit_frame
from test_gpu_batch import assert_frame_parity, executor, oracle_fit, scmos_frames
from storm_analysis_amd.sa_library import dao_batch
from storm_analysis_amd.sa_library import dao_fit_c as D
from storm_analysis_amd.diagnostics import synth

KW_Z = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}


def fit_team(team, cls, kwds, image, peaks, **kw):
    print("team fit: %s team=%d nfit=%d" % (cls.__name__, team, len(peaks["x"])), flush=True)
    return fit_frame("hip", cls, dict(kwds, team=team), image, peaks, **kw)


@pytest.mark.gpu
@pytest.mark.parametrize("team", [8, 40, 250])
def test_team_z_matches_oracle(team):
    rng = numpy.random.default_rng(synth.SEED_BASE + 1)
    fr = synth.frame_z(rng, size=256, n=250)
    bg = numpy.zeros(fr.image.shape)
    g = fit_team(team, D.MultiFitterZ, KW_Z, fr.image, fr.peaks, background=bg)
    r = fit_frame("oracle", D.MultiFitterZ, KW_Z, fr.image, fr.peaks, background=bg)
    assert_parity(g, r)


@pytest.mark.gpu
def test_team_2d_dense_matches_oracle():
    rng = numpy.random.default_rng(synth.SEED_BASE + 4)
    fr = synth.frame_2d_dense(rng, size=256, n=600)
    kw = {"roi_size": 18, "sigma_range": [0.96, 8.0]}
    g = fit_team(30, D.MultiFitter2D, kw, fr.image, fr.peaks)
    r = fit_frame("oracle", D.MultiFitter2D, kw, fr.image, fr.peaks)
    assert_parity(g, r)


@pytest.mark.gpu
def test_team_3d_scmos_matches_oracle():
    rng = numpy.random.default_rng(synth.SEED_BASE + 2)
    fr = synth.frame_3d_scmos(rng, size=128, density=1.0)
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    g = fit_team(16, D.MultiFitter3D, kw, fr.image, fr.peaks, rqe=fr.rqe, scmos=fr.scmos_cal)
    r = fit_frame("oracle", D.MultiFitter3D, kw, fr.image, fr.peaks, rqe=fr.rqe, scmos=fr.scmos_cal)
    assert_parity(g, r)


def batch_run(frames, kw, bgs, waves=0, margin=-1, team=0, precision="fp64", model="3d", handoff=-1):
    shapes = [fr.image.shape for fr in frames]
    caps = [max(1, len(fr.peaks["x"])) for fr in frames]
    bf = dao_batch.BatchFitter(shapes, caps, model=model, precision=precision, **kw)
    bf.setTuning(waves, margin, team)
    bf.setTeamHandoff(handoff)
    for k, fr in enumerate(frames):
        if fr.rqe is not None:
            bf.setCalibration(k, rqe=fr.rqe, scmos_cal=fr.scmos_cal)
    bf.newImages([fr.image for fr in frames])
    bf.newBackgrounds(bgs)
    bf.newPeaks([fr.peaks for fr in frames], "hdf5")
    print("batch fit: %d frames team=%d margin=%d %s" % (len(frames), team, margin, precision), flush=True)
    bf.doFit()
    out = [bf.getPeaks(k) for k in range(len(frames))]
    counters = [bf.getCounters(k) for k in range(len(frames))]
    bf.cleanup()
    return out, counters


@pytest.mark.gpu
@pytest.mark.parametrize("margin", [0, 2])
def test_team_batch_uneven_scmos_matches_oracle(margin):
    """Margin 0: reserved boxes are the bare ROIs, so many turns escape and roll back."""
    frames = scmos_frames()
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    out, counters = batch_run(frames, kw, bgs, margin=margin, team=12)
    for k, fr in enumerate(frames):
        r = oracle_fit(D.MultiFitter3D, kw, fr, bgs[k])
        assert out[k]["x"].size == r["x"].size
        assert_frame_parity(out[k], r, counters[k])


@pytest.mark.gpu
@pytest.mark.parametrize("handoff", [0, 1, 10 ** 6])
def test_team_handoff_keeps_fits(handoff):
    """
    A team hands a frame's tail -- its LM runs of at most `handoff` ops -- to
    one workgroup (0: never; 10**6: from the first sweep on, so the resumed
    fit starts at iteration 0).  FP64 fits are the default's to the bit and
    the oracle's to the parity bar, counters included.
    """
    frames = scmos_frames()
    kw = {"roi_size": 18, "sigma_range": [0.8, 6.0]}
    bgs = [numpy.full(fr.image.shape, 100.0) for fr in frames]
    out, counters = batch_run(frames, kw, bgs, margin=2, team=12, handoff=handoff)
    ref, ref_c = batch_run(frames, kw, bgs, margin=2, team=12)
    for k, fr in enumerate(frames):
        for key in ref[k]:
            numpy.testing.assert_array_equal(out[k][key], ref[k][key])
        assert counters[k] == ref_c[k]
        r = oracle_fit(D.MultiFitter3D, kw, fr, bgs[k])
        assert_frame_parity(out[k], r, counters[k])


@pytest.mark.gpu
def test_team_mixed_matches_one_workgroup_mixed():
    """Mixed precision: the team gives the one-workgroup mixed fitter's results."""
    rng = numpy.random.default_rng(synth.SEED_BASE + 40)
    frames = [synth.frame_z(rng, size=256, n=250) for _ in range(3)]
    for fr in frames:
        fr.rqe = None
    kw = {"roi_size": 18, "wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}
    bgs = [numpy.zeros(fr.image.shape) for fr in frames]
    one, one_c = batch_run(frames, kw, bgs, team=0, precision="mixed", model="Z")
    team, team_c = batch_run(frames, kw, bgs, team=20, precision="mixed", model="Z")
    for g, r, gc, rc in zip(team, one, team_c, one_c):
        assert_frame_parity(g, r)
        assert gc["n_iterations"] == rc["n_iterations"]


@pytest.mark.gpu
@pytest.mark.parametrize("cls,kw", [
    (D.MultiFitter2DFixed, {"roi_size": 12}),
    (D.MultiFitter2DFixedALS, {"roi_size": 12}),
    (D.MultiFitter2DLS, {"roi_size": 12, "sigma_range": [0.96, 8.0]}),
    (D.MultiFitter2DFWLS, {"roi_size": 14, "sigma_range": [0.96, 8.0]}),
])
def test_team_other_models_match_oracle(cls, kw):
    """Other shapes, ROI sizes (slots per lane) and error models through the team fitter."""
    rng = numpy.random.default_rng(synth.SEED_BASE + 0)
    fr = synth.frame_2dfixed(rng, size=256, n=200)
    g = fit_team(10, cls, kw, fr.image, fr.peaks)
    r = fit_frame("oracle", cls, kw, fr.image, fr.peaks)
    assert_parity(g, r)


@pytest.mark.gpu
@pytest.mark.parametrize("size,density,team", [(128, 1.2, 0), (128, 1.2, 50), (128, 2.0, 0), (96, 4.6, 0),
                                               (96, 4.6, 100)])
def test_high_density_overlap_matches_oracle(size, density, team):
    """
    The overlap stress of the high-density configuration at sizes the oracle
    runs in seconds: 2D model, free width, sigma U(1.2, 2.0), up to 4.6
    emitters/um^2 at 160 nm pixels (a start peak for every emitter; at 4.6
    nearly all peaks end as ERROR, which exercises the lost / negative-height
    paths in order), one workgroup and the team fitter.
    """
    n = int(round(size * size * 0.0256 * density))
    rng = numpy.random.default_rng(synth.SEED_BASE + (50 if density > 4 else 51))
    fr = synth.frame_2d_dense(rng, size=size, n=n)
    kw = {"roi_size": 18, "sigma_range": [0.96, 8.0]}
    g = fit_team(team, D.MultiFitter2D, kw, fr.image, fr.peaks)
    r = fit_frame("oracle", D.MultiFitter2D, kw, fr.image, fr.peaks)
    assert_parity(g, r)


def dense_rollback_frame():
    size, density = 128, 2.0
    n = int(round(size * size * 0.0256 * density))
    rng = numpy.random.default_rng(synth.SEED_BASE + 51)
    return synth.frame_2d_dense(rng, size=size, n=n), n


def batch_fit_2d(fr, n, kw, waves=0, margin=-1):
    bf = dao_batch.BatchFitter([fr.image.shape], [n], model="2d", **kw)
    try:
        if waves or margin >= 0:
            bf.setTuning(waves, margin, 0)
        bf.newImages([fr.image])
        bf.newBackgrounds([numpy.full(fr.image.shape, 100.0)])
        bf.newPeaks([fr.peaks], "hdf5")
        bf.doFit()
        return bf.getPeaks(0)
    finally:
        bf.cleanup()


def assert_flags_and_positions(g, r):
    numpy.testing.assert_array_equal(g["status"], r["status"])
    numpy.testing.assert_array_equal(g["iterations"], r["iterations"])
    ok = r["status"] != D.ERROR
    for k in ("x", "y"):
        assert numpy.abs(g[k] - r[k])[ok].max() < TOL_POS, k


@pytest.mark.gpu
def test_high_density_rounds_executor_matches_oracle():
    """
    The multi-kernel executor on the 2.0 emitters/um^2 frame: a frame whose
    turn leaves its box returns to the copy of its state taken at the
    sweep's start and reruns the sweep with that turn alone in the frame.
    """
    fr, n = dense_rollback_frame()
    kw = {"roi_size": 18, "sigma_range": [0.96, 8.0]}
    r = fit_frame("oracle", D.MultiFitter2D, kw, fr.image, fr.peaks)
    with executor("rounds"):
        g = batch_fit_2d(fr, n, kw)
    assert_flags_and_positions(g, r)


@pytest.mark.gpu
@pytest.mark.parametrize("waves,margin", [(1, 20), (8, 2), (8, 20), (2, 2)])
def test_high_density_rollbacks_match_oracle(waves, margin):
    """
    Rollbacks at 2.0 emitters/um^2: with a small margin turns leave their
    boxes after moving, with a large one they fill their journals (more ROI
    positions than snapshots); either way the turn and every later turn that
    ran beside it are undone and run again.  Rejected candidates here stamp
    contributions large enough to wipe out a pixel's other contributions when
    removed again, so only an exact restore of the pixels keeps the
    reference's results -- whatever the waves per workgroup and the margin.
    """
    fr, n = dense_rollback_frame()
    kw = {"roi_size": 18, "sigma_range": [0.96, 8.0]}
    r = fit_frame("oracle", D.MultiFitter2D, kw, fr.image, fr.peaks)
    g = batch_fit_2d(fr, n, kw, waves, margin)
    assert_flags_and_positions(g, r)
