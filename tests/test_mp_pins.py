"""
Pins of the multiplane DAO fitter (storm_analysis/multi_plane/mp_fit.c
mpIterateLM :180-487, mpUpdate :604-690, mpCheckError :44-67;
mp_fit_dao.c mpDaoUpdate :173-209, mpDaoNewPeaks :82-153), run against the
oracle (oracle/mp_oracle.c) here and against the HIP library with -m gpu.

The expected values come from evidence that shares no code with either
(tests/mp_evidence.py: per-channel J, H by complex steps, numpy solves,
closed-form deviances), from invariants, and from noise-free data:

- the first coupled step: each channel's own damped step; heights and x
  widths per channel; x, y of channel 0 moved by the mean of the channel
  steps brought to channel 0's axes (xt_Nto0 / yt_Nto0), weighted by the
  z-interpolated w_x / w_y and the new heights; the other channels placed
  by xt_0toN / yt_0toN; the y widths left alone;
- the floor of 0.01 on heights and the x width clip, per channel;
- the accept / reject decision on the channels' summed deviance;
- a failed solve in channel 0 stops the channel loop (channel 1 does not
  count that cycle);
- a group whose start lies off one channel's image is an ERROR group in
  every channel, its foreground removed everywhere;
- a second channel that is the first shifted by whole pixels changes
  nothing (against a one-channel fit at the same tolerance);
- three channels of noise-free data converge to the truth.

Plausible slips each one catches are named in its docstring;
test_oracle_mutations.py plants them in copies of the oracle.
"""
import numpy
import pytest

from conftest import make_mp_fitter
from lm_evidence import Frame
from mp_evidence import IDENTITY, Affine, WFrame, coupled_step, group_deviance, render, tables
from storm_analysis_amd.multi_plane import mp_fit_c as mpFitC
from storm_analysis_amd.multi_plane import mp_fit_dao_c as MPD
from storm_analysis_amd.sa_library import dao_fit_c as D

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]
R = 12
SHAPE = (40, 44)
REL = 1e-9
SIGMA_RANGE = [1.0, 3.0]
WMIN, WMAX = 1.0 / (2 * SIGMA_RANGE[1] ** 2), 1.0 / (2 * SIGMA_RANGE[0] ** 2)

# channel 1 rotated by 90 degrees, channel 2 mirrored and shifted
ROT = Affine([40, 0, -1], [-3, 1, 0], [3, 0, 1], [40, -1, 0])
MIRROR = Affine([43, -1, 0], [1.5, 0, 1], [43, -1, 0], [-1.5, 0, 1])
SHIFT = Affine([-3, 1, 0], [2, 0, 1], [3, 1, 0], [-2, 0, 1])


def channel_frames(maps, emitters, heights, bgs, sigmas, seed, noise=True, offsets=None, shape=SHAPE,
                   qscale=None):
    """
    Per channel a calibrated frame: emitters (x0, y0) placed in channel c by
    its map (plus offsets[c], to make a channel disagree on purpose),
    heights[c][k], background bgs[c], psf sigma sigmas[c], rqe around
    qscale[c] (default 1).
    """
    rng = numpy.random.default_rng(seed)
    frs = []
    for c, m in enumerate(maps):
        rqe = rng.uniform(0.7, 1.3, shape) * (1.0 if qscale is None else qscale[c])
        scmos = rng.uniform(0.0, 3.0, shape)
        em = []
        for k, (x, y) in enumerate(emitters):
            xc, yc = m(x, y)
            if offsets is not None:
                xc, yc = xc + offsets[c][0], yc + offsets[c][1]
            em.append((xc, yc, heights[c][k]))
        mean = render(shape, em, rqe, scmos, sigmas[c]) + bgs[c] * rqe + scmos
        img = numpy.maximum(rng.poisson(mean).astype(float), 1.0) if noise else mean
        frs.append(WFrame(img, rqe, scmos, R, model="2d"))
    return frs


class MPRun(object):
    """A multiplane fitter on the channel frames; start peaks in channel 0's coordinates."""
    def __init__(self, backend, frs, maps, peaks, weights=None, tol=1e-6, sigma_range=SIGMA_RANGE):
        self.frs, self.maps = frs, maps
        self.n = len(frs)
        self.m = make_mp_fitter(MPD.MPFitDao2D, backend, n_channels=self.n, roi_size=R, sigma_range=sigma_range)
        self.m.default_tol = tol
        for c, fr in enumerate(frs):
            self.m.initializeChannel(fr.rqe, fr.scmos, c)
        self.m.setMapping(*tables(maps))
        if weights is None:
            self.m.setWeights()
        else:
            w, off, scale = weights
            # MPFitDao2D.setWeights() takes equal weights; the base class takes a table
            mpFitC.MPFit.setWeights(self.m, w, off, scale, verbose=False)
        self.m.newImage([(fr.x - fr.scmos) / fr.rqe for fr in frs])
        self.m.newBackground([numpy.zeros(fr.x.shape) for fr in frs])
        self.m.newPeaks(peaks, "text")

    def sweep(self):
        self.m.iterate()
        return self.state()

    def state(self):
        out = []
        for c in range(self.n):
            m, xo = self.m, self.frs[c].xoff
            p = {k: m.getPeakProperty(k, channel=c) for k in ("height", "x", "y", "background", "xsigma", "ysigma",
                                                             "iterations", "status")}
            fd = m.mfit.contents.fit_data[c].contents
            p["counters"] = {k: getattr(fd, k) for k in ("n_dposv", "n_iterations", "n_lost", "n_margin",
                                                         "n_neg_fi", "n_neg_height", "n_non_decr")}
            p["theta"] = numpy.stack([p["height"], p["x"] - xo, p["y"] - xo, 1.0 / (2.0 * p["xsigma"] ** 2),
                                      p["background"]], axis=1)
            out.append(p)
        return out

    def close(self):
        self.m.cleanup(verbose=False)


def start_rows(x, y, h, bg, sigma):
    x, y = numpy.atleast_1d(x), numpy.atleast_1d(y)
    n = x.size
    return {"x": x, "y": y, "z": numpy.zeros(n), "background": numpy.full(n, bg), "height": numpy.full(n, h),
            "xsigma": numpy.full(n, sigma), "ysigma": numpy.full(n, sigma)}


def counters(**nonzero):
    c = {k: 0 for k in ("n_dposv", "n_iterations", "n_lost", "n_margin", "n_neg_fi", "n_neg_height", "n_non_decr")}
    c.update(nonzero)
    return c


def assert_theta(got, want, rel=REL):
    got, want = numpy.asarray(got, dtype=float), numpy.asarray(want, dtype=float)
    assert numpy.all(numpy.abs(got - want) <= rel * numpy.maximum(1.0, numpy.abs(want))), (got, want)


def distinct(a, b, gap=1e-5):
    return numpy.max(numpy.abs(numpy.asarray(a) - numpy.asarray(b)) / numpy.maximum(1.0, numpy.abs(b))) > gap


def step_outcome(frs, maps, start, lam, wx, wy, wmin=WMIN, wmax=WMAX):
    """Predicted thetas and the start / candidate summed deviances of one coupled step."""
    ths, new, deltas = coupled_step(frs, maps, start, lam, wx, wy, wmin, wmax)
    o0 = [fr.origin(t) for fr, t in zip(frs, ths)]
    o1 = [(Frame.moved_origin(t[1], o[0]), Frame.moved_origin(t[2], o[1])) for t, o in zip(new, o0)]
    e_start = [group_deviance([f], [t], [o], start["W"]) for f, t, o in zip(frs, ths, o0)]
    e_new = [group_deviance([f], [t], [o], start["W"]) for f, t, o in zip(frs, new, o1)]
    return ths, new, deltas, e_start, e_new


# ---------------------------------------------------------------------------


@pytest.mark.parametrize("backend", BACKENDS)
def test_first_coupled_step(backend):
    """
    Two channels (channel 1 rotated by 90 degrees), heights 900 and 250,
    w_x / w_y interpolated half way between two z rows (z_scale 2, offset
    -0.25 at z = 0: row 0, dz = 0.5).  The lambda = 1 step lowers the
    summed deviance by more than the tolerance: accepted, RUNNING, one
    cycle per channel.  Slips: unweighted or height-free mean, the 0 -> N
    tables for the way back (they differ for a rotation), a transposed
    coefficient, the weight table read channel-major, the y width updated
    with the x width, the map applied without the ROI offset.
    """
    maps = [IDENTITY, ROT]
    frs = channel_frames(maps, [(20.3, 18.6)], [[900.0], [250.0]], [20.0, 35.0], [1.5, 1.5], seed=3)
    start = {"x": 20.05, "y": 18.85, "h": [600.0, 600.0], "W": 1.0 / (2 * 1.6 ** 2), "bg": [25.0, 25.0]}
    w = {"x": numpy.array([[0.2, 0.8], [0.6, 0.4]]), "y": numpy.array([[0.9, 0.1], [0.5, 0.5]])}
    for k in ("bg", "h", "z"):
        w[k] = numpy.ones((2, 2))
    wx, wy = [0.4, 0.6], [0.7, 0.3]
    ths, want, deltas, e0, e1 = step_outcome(frs, maps, start, 1.0, wx, wy)
    # premises: accepted, nothing clamped, and the slips lead elsewhere
    assert sum(e1) < sum(e0) and (sum(e0) - sum(e1)) / sum(e0) > 2e-6
    assert all(t[0] > 0.01 for t in want) and all(WMIN < t[3] < WMAX for t in want)
    _, swapped, _ = coupled_step(frs, maps, start, 1.0, [0.6, 0.4], [0.3, 0.7], WMIN, WMAX)
    _, flat, _ = coupled_step(frs, maps, start, 1.0, [0.5, 0.5], [0.5, 0.5], WMIN, WMAX)
    assert distinct(swapped[0], want[0]) and distinct(flat[0], want[0])
    run = MPRun(backend, frs, maps, start_rows(20.05, 18.85, 600.0, 25.0, 1.6), weights=(w, -0.25, 2.0))
    s = run.sweep()
    for c in range(2):
        assert s[c]["status"][0] == D.RUNNING and s[c]["iterations"][0] == 1, c
        assert s[c]["counters"] == counters(n_iterations=1), c
        assert_theta(s[c]["theta"][0], want[c])
        assert s[c]["ysigma"][0] == 1.6   # the y width is not a parameter here
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_heights_floor_at_a_hundredth(backend):
    """
    Channel 1 holds no emitter: its height step goes below zero and is
    floored at 0.01 (mp_fit_dao.c:184-186) -- no reset, no n_neg_height --
    and that 0.01 weights channel 1's x, y step.  Slip: rejecting the step
    as the single-channel fitter does.
    """
    maps = [IDENTITY, SHIFT]
    frs = channel_frames(maps, [(20.3, 18.6)], [[900.0], [0.0]], [20.0, 30.0], [1.5, 1.5], seed=5)
    start = {"x": 20.1, "y": 18.8, "h": [500.0, 500.0], "W": 1.0 / (2 * 1.5 ** 2), "bg": [22.0, 22.0]}
    ths, want, deltas, e0, e1 = step_outcome(frs, maps, start, 1.0, [0.5, 0.5], [0.5, 0.5])
    assert ths[1][0] - deltas[1][0] < 0.01 and want[1][0] == 0.01
    assert sum(e1) < sum(e0) and (sum(e0) - sum(e1)) / sum(e0) > 2e-6
    run = MPRun(backend, frs, maps, start_rows(20.1, 18.8, 500.0, 22.0, 1.5))
    s = run.sweep()
    for c in range(2):
        assert s[c]["status"][0] == D.RUNNING and s[c]["counters"] == counters(n_iterations=1), c
        assert_theta(s[c]["theta"][0], want[c])
    assert s[1]["height"][0] == 0.01
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_x_width_clips_per_channel(backend):
    """
    Channel 1's emitter is wider (sigma 2.4) than the range allows (at most
    2.0): its x width step is clipped to width_min (daoCheck2D inside
    mpDaoUpdate, mp_fit_dao.c:193-204), channel 0's is not.  Slips:
    resetting instead of clipping, clipping every channel alike.
    """
    maps = [IDENTITY, SHIFT]
    frs = channel_frames(maps, [(20.3, 18.6)], [[900.0], [700.0]], [20.0, 20.0], [1.5, 2.4], seed=6)
    start = {"x": 20.2, "y": 18.7, "h": [800.0, 800.0], "W": 1.0 / (2 * 1.98 ** 2), "bg": [20.0, 20.0]}
    wmin, wmax = 1.0 / (2 * 2.0 ** 2), 1.0 / (2 * 1.0 ** 2)
    ths, want, deltas, e0, e1 = step_outcome(frs, maps, start, 1.0, [0.5, 0.5], [0.5, 0.5], wmin, wmax)
    assert ths[1][3] - deltas[1][3] < wmin and want[1][3] == wmin and want[0][3] > wmin
    assert sum(e1) < sum(e0) and (sum(e0) - sum(e1)) / sum(e0) > 2e-6
    run = MPRun(backend, frs, maps, start_rows(20.2, 18.7, 800.0, 20.0, 1.98), sigma_range=[1.0, 2.0])
    s = run.sweep()
    for c in range(2):
        assert s[c]["status"][0] == D.RUNNING and s[c]["counters"] == counters(n_iterations=1), c
        assert_theta(s[c]["theta"][0], want[c])
    assert abs(s[1]["xsigma"][0] - 2.0) < 1e-12
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_the_summed_deviance_decides(backend):
    """
    Channel 1's emitter sits where the start maps (it disagrees with
    channel 0 on purpose) and its rqe is six times channel 0's; x, y follow
    channel 0 (w_x = w_y = 0.98 / 0.02).  The lambda = 1 step lowers
    channel 0's deviance but raises the sum: rejected in every channel
    (n_non_decr + 1 each), lambda x 4, and the lambda = 4 step, which
    lowers the sum, is kept: two cycles per channel, while the peak's own
    count is 1 (the reset restores the turn-start record, iterations
    included, mp_fit.c:496-516).  Slips: deciding on
    channel 0's deviance alone (keeps the lambda = 1 step), counting the
    rejection in channel 0 only.
    """
    maps = [IDENTITY, ROT]
    sx, sy = 19.9, 18.6
    t1, s1 = ROT(20.3, 18.6), ROT(sx, sy)
    frs = channel_frames(maps, [(20.3, 18.6)], [[900.0], [900.0]], [20.0, 20.0], [1.5, 1.5], seed=2,
                         offsets=[(0.0, 0.0), (s1[0] - t1[0], s1[1] - t1[1])], qscale=[1.0, 6.0])
    start = {"x": sx, "y": sy, "h": [900.0, 900.0], "W": 1.0 / (2 * 1.5 ** 2), "bg": [20.0, 20.0]}
    wx = wy = [0.98, 0.02]
    w = {k: numpy.ones((2, 2)) for k in ("bg", "h", "z")}
    w["x"] = w["y"] = numpy.array([[0.98, 0.02], [0.98, 0.02]])
    _, want1, _, e0, e1 = step_outcome(frs, maps, start, 1.0, wx, wy)
    assert e1[0] < e0[0] and sum(e1) > sum(e0) and (sum(e1) - sum(e0)) / sum(e0) > 2e-6
    _, want4, _, f0, f1 = step_outcome(frs, maps, start, 4.0, wx, wy)
    assert sum(f1) < sum(f0) and (sum(f0) - sum(f1)) / sum(f0) > 2e-6
    assert distinct(want1[0], want4[0])
    run = MPRun(backend, frs, maps, start_rows(sx, sy, 900.0, 20.0, 1.5), weights=(w, 0.0, 0.0))
    s = run.sweep()
    for c in range(2):
        assert s[c]["status"][0] == D.RUNNING and s[c]["iterations"][0] == 1, c
        assert s[c]["counters"] == counters(n_iterations=2, n_non_decr=1), c
        assert_theta(s[c]["theta"][0], want4[c])
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_failed_solve_stops_the_channel_loop(backend):
    """
    Height 0: every channel's H is singular.  Channel 0's solve fails
    first and the loop over channels stops there (mp_fit.c:243-269), so
    channel 1 counts no cycle; lambda x 4 per cycle until lambda > 1e20:
    34 cycles, then every channel n_lost + 1 and ERROR.  Slips: solving all
    channels before giving up (channel 1 would count 34 cycles and 34
    n_dposv), another lambda factor.
    """
    maps = [IDENTITY, SHIFT]
    frs = channel_frames(maps, [(20.3, 18.6)], [[300.0], [300.0]], [20.0, 20.0], [1.5, 1.5], seed=1)
    run = MPRun(backend, frs, maps, start_rows(20.3, 18.6, 0.0, 20.0, 1.5))
    s = run.sweep()
    assert s[0]["status"][0] == D.ERROR and s[1]["status"][0] == D.ERROR
    assert s[0]["iterations"][0] == 34 and s[1]["iterations"][0] == 0
    assert s[0]["counters"] == counters(n_iterations=34, n_dposv=34, n_lost=1)
    assert s[1]["counters"] == counters(n_lost=1)
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_a_group_off_one_channel_is_an_error_everywhere(backend):
    """
    Peak 0 maps to x = 4.5 in channel 1 (x1 = x0 - 3), where its ROI leaves
    the image: ERROR there at newPeaks, so mpCheckError (mp_fit_dao.c:148-151)
    subtracts it from channel 0 and makes it ERROR there too; channel 0's
    fit image then holds peak 1 only.  Slip: no synchronisation after
    newPeaks (channel 0 keeps fitting it).
    """
    maps = [IDENTITY, SHIFT]
    frs = channel_frames(maps, [(7.5, 20.0), (25.0, 18.0)], [[500.0, 600.0], [500.0, 600.0]], [20.0, 20.0],
                         [1.5, 1.5], seed=8)
    rows = start_rows(numpy.array([7.5, 25.0]), numpy.array([20.0, 18.0]), 500.0, 20.0, 1.5)
    run = MPRun(backend, frs, maps, rows)
    assert list(run.m.getPeakProperty("status", channel=0)) == [D.ERROR, D.RUNNING]
    assert list(run.m.getPeakProperty("status", channel=1)) == [D.ERROR, D.RUNNING]
    alone = MPRun(backend, frs[:1], maps[:1], start_rows(25.0, 18.0, 500.0, 20.0, 1.5))
    numpy.testing.assert_allclose(run.m.getFitImage()[0], alone.m.getFitImage()[0], rtol=0, atol=1e-9)
    alone.close()
    run.m.doFit()
    assert list(run.m.getPeakProperty("status", channel=0))[0] == D.ERROR
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_new_error_groups_after_storage_growth(backend):
    """
    Not the reference's behaviour (a deliberate fix): newPeaks checks the
    groups it just added.  When adding peaks grows the storage, the k ERROR
    peaks already held are dropped first, so the reference's check of
    [old nfit, old nfit + n) misses the first k new groups (and reads k
    past the live peaks); here the 600 new groups are checked and the
    first, off channel 1's image, is an ERROR group in both.
    """
    maps = [IDENTITY, SHIFT]
    frs = channel_frames(maps, [(20.0, 20.0)], [[300.0], [300.0]], [20.0, 20.0], [1.5, 1.5], seed=9)
    run = MPRun(backend, frs, maps, start_rows(numpy.array([7.5, 25.0]), numpy.array([20.0, 18.0]), 500.0,
                                               20.0, 1.5))
    assert run.m.getNFit() == 2
    rng = numpy.random.default_rng(1)
    x = rng.uniform(14.0, 30.0, 600)
    y = rng.uniform(10.0, 28.0, 600)
    x[0] = 7.5   # the first new group leaves channel 1
    run.m.newPeaks(start_rows(x, y, 100.0, 20.0, 1.5), "text")
    n = run.m.getNFit(0)
    assert n == run.m.getNFit(1)
    for c in range(2):
        st = run.m.getPeakProperty("status", channel=c)
        assert st[n - 600] == D.ERROR and numpy.all(st[n - 599:] == D.RUNNING), c
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_a_whole_pixel_shifted_copy_changes_nothing(backend):
    """
    Channel 1 is channel 0 shifted by (-3, +2) whole pixels, with the same
    noise and calibration: every channel step is the same, so the group
    fits as one channel does at the same tolerance (2e-6 = the pair's
    1e-6 x 2), and channel 1 reads channel 0 shifted.  Slips: a mean
    without its normalisation, the map applied without the ROI offset,
    the offset of the N -> 0 table added to a step.
    """
    big = (SHAPE[0] + 2, SHAPE[1] + 3)
    emit = [(12.3, 14.6, 700.0), (15.1, 16.2, 400.0), (24.7, 22.4, 900.0), (30.2, 12.9, 300.0),
            (21.0, 28.8, 500.0)]
    rng = numpy.random.default_rng(11)
    rqe = rng.uniform(0.7, 1.3, big)
    scmos = rng.uniform(0.0, 3.0, big)
    mean = render(big, emit, rqe, scmos, 1.5) + 20.0 * rqe + scmos
    x = numpy.maximum(rng.poisson(mean).astype(float), 1.0)
    # channel 0 = rows 2.., cols 0..; channel 1 = rows 0.., cols 3.. -> x1 = x0 - 3, y1 = y0 + 2
    crop = [(slice(2, 2 + SHAPE[0]), slice(0, SHAPE[1])), (slice(0, SHAPE[0]), slice(3, 3 + SHAPE[1]))]
    frs = [WFrame(x[cr], rqe[cr], scmos[cr], R, model="2d") for cr in crop]
    starts = [(e[0] + 0.2, e[1] - 2.0 - 0.15) for e in emit]
    rows = start_rows(numpy.array([s[0] for s in starts]), numpy.array([s[1] for s in starts]), 500.0, 22.0, 1.6)
    pair = MPRun(backend, frs, [IDENTITY, SHIFT], rows)
    one = MPRun(backend, frs[:1], [IDENTITY], rows, tol=2e-6)
    pair.m.doFit()
    one.m.doFit()
    p, o = pair.state(), one.state()
    assert numpy.all(o[0]["status"] == D.CONVERGED)
    for c in range(2):
        numpy.testing.assert_array_equal(p[c]["status"], o[0]["status"])
        numpy.testing.assert_array_equal(p[c]["iterations"], o[0]["iterations"])
    for k in ("height", "background", "xsigma", "ysigma"):
        for c in range(2):
            numpy.testing.assert_allclose(p[c][k], o[0][k], rtol=1e-9, err_msg=k)
    numpy.testing.assert_allclose(p[0]["x"], o[0]["x"], rtol=0, atol=1e-9)
    numpy.testing.assert_allclose(p[0]["y"], o[0]["y"], rtol=0, atol=1e-9)
    numpy.testing.assert_allclose(p[1]["x"], o[0]["x"] - 3.0, rtol=0, atol=1e-9)
    numpy.testing.assert_allclose(p[1]["y"], o[0]["y"] + 2.0, rtol=0, atol=1e-9)
    pair.close()
    one.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_noise_free_channels_converge_to_the_truth(backend):
    """
    Three channels (identity, rotated, mirrored and shifted) of noise-free
    data with their own heights and backgrounds; the start is off by 0.35
    and 0.3 px.  The truth is where every channel's model equals its data,
    and the fit ends there in every channel.  Slips: a wrong map or a
    transposed coefficient leaves some channel's peak elsewhere.
    """
    maps = [IDENTITY, ROT, MIRROR]
    truth = (20.3, 18.6)
    hs, bgs = [900.0, 250.0, 500.0], [20.0, 35.0, 10.0]
    frs = channel_frames(maps, [truth], [[h] for h in hs], bgs, [1.5] * 3, seed=4, noise=False)
    run = MPRun(backend, frs, maps, start_rows(truth[0] + 0.35, truth[1] - 0.3, 600.0, 25.0, 1.5))
    run.m.doFit()
    s = run.state()
    for c in range(3):
        assert s[c]["status"][0] == D.CONVERGED, c
        xc, yc = maps[c](*truth)
        assert abs(s[c]["x"][0] - xc) < 1e-5 and abs(s[c]["y"][0] - yc) < 1e-5, (c, s[c]["x"], xc)
        assert abs(s[c]["height"][0] / hs[c] - 1.0) < 1e-5, c
        assert abs(s[c]["background"][0] - bgs[c]) < 1e-4, c
        assert abs(s[c]["xsigma"][0] - 1.5) < 1e-5, c
    run.close()
