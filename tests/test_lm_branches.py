"""
Branch pins for the LM state machine (mFitIterateLM, multi_fit.c:1072-1349;
mFitResetPeak :1788-1798; mFitCheck :470-502; mFitUpdate :1933-1946;
mFitRecenterPeaks :1644-1690; daoUpdate2D / daoCheck2D, dao_fit.c:1061-1085,
1798-1815), with evidence that neither restatement of the fitter
(oracle/dao_oracle.c, tests/lm_pure.py) supplies.

Each test builds a frame whose first sweep (or first two) must take one
branch, and states that branch's outcome from the reference's text: the
counters it moves, the cycles it takes, the parameters and status it leaves.
The numbers come from tests/lm_evidence.py (closed-form deviance,
complex-step gradient, Gauss-Newton matrix, numpy.linalg.solve).  Each test
first checks its premise with that evidence (for example: the lambda = 1
step leaves the image and the lambda = 4 step does not) and that the
plausible misreading it targets would end elsewhere, by more than the
tolerance of the comparison.

Parametrised over the oracle (CPU) and the HIP library (-m gpu).
"""
import numpy
import pytest

from conftest import make_fitter
from lm_evidence import Frame, Scene
from storm_analysis_amd.sa_library import dao_fit_c as D

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]
R = 12
SHAPE = (40, 44)
REL = 1e-9       # parameters vs the evidence (different summation order, solver)


def make_frame(truth, sigma=1.5, seed=1, model="2dfixed", sigma_range=None):
    """A Poisson frame of one Gaussian (x, y, h, bg) with a varying rqe / scmos."""
    rng = numpy.random.default_rng(seed)
    rqe = rng.uniform(0.7, 1.3, SHAPE)
    scmos = rng.uniform(0.0, 3.0, SHAPE)
    w = 1.0 / (2 * sigma ** 2)
    yy, xx = numpy.mgrid[0:SHAPE[0], 0:SHAPE[1]]
    x, y, h, bg = truth
    mean = h * numpy.exp(-((xx - x) ** 2 + (yy - y) ** 2) * w) * rqe + bg * rqe + scmos
    img = numpy.maximum(rng.poisson(mean).astype(float), 1.0)
    if model == "2dfixed":
        return Frame(img, rqe, scmos, R, width=w)
    wmax = 1.0 / (2 * sigma_range[0] ** 2)
    wmin = 1.0 / (2 * sigma_range[1] ** 2)
    return Frame(img, rqe, scmos, R, model="2d", wmin=wmin, wmax=wmax)


class Run(object):
    """One fitter on the frame, peaks given in the fitter's parameters."""
    def __init__(self, backend, fr, peaks, tol=1e-6, sigma=1.5, sigma_range=None):
        if fr.model == "2dfixed":
            self.m = make_fitter(D.MultiFitter2DFixed, backend, roi_size=fr.r, rqe=fr.rqe, scmos_cal=fr.scmos)
        else:
            self.m = make_fitter(D.MultiFitter2D, backend, roi_size=fr.r, rqe=fr.rqe, scmos_cal=fr.scmos,
                                 sigma_range=sigma_range)
        self.m.default_tol = tol
        self.fr = fr
        image = (fr.x - fr.scmos) / fr.rqe
        self.m.initializeC(image)
        self.m.newImage(image)
        self.m.newBackground(numpy.zeros(image.shape))
        n = len(peaks)
        sig = [sigma if fr.model == "2dfixed" else numpy.sqrt(1.0 / (2.0 * th[3])) for th in peaks]
        self.m.newPeaks({"x": numpy.array([th[1] + fr.xoff for th in peaks]),
                         "y": numpy.array([th[2] + fr.xoff for th in peaks]), "z": numpy.zeros(n),
                         "background": numpy.array([th[-1] for th in peaks]),
                         "height": numpy.array([th[0] for th in peaks]),
                         "xsigma": numpy.array(sig), "ysigma": numpy.array(sig)}, "text")

    def sweep(self):
        self.m.iterate()
        return self.state()

    def state(self):
        m, xo = self.m, self.fr.xoff
        p = {k: m.getPeakProperty(k) for k in ("height", "x", "y", "background", "xsigma", "ysigma", "error",
                                               "iterations", "status")}
        c = m.mfit.contents
        p["counters"] = {k: getattr(c, k) for k in ("n_dposv", "n_iterations", "n_lost", "n_margin", "n_neg_fi",
                                                    "n_neg_height", "n_non_decr")}
        if self.fr.model == "2dfixed":
            p["theta"] = numpy.stack([p["height"], p["x"] - xo, p["y"] - xo, p["background"]], axis=1)
        else:
            p["theta"] = numpy.stack([p["height"], p["x"] - xo, p["y"] - xo, 1.0 / (2.0 * p["xsigma"] ** 2),
                                      p["background"]], axis=1)
        p["fit"] = m.getFitImage()
        return p

    def close(self):
        self.m.cleanup(verbose=False)


def counters(**nonzero):
    c = {k: 0 for k in ("n_dposv", "n_iterations", "n_lost", "n_margin", "n_neg_fi", "n_neg_height", "n_non_decr")}
    c.update(nonzero)
    return c


def assert_theta(got, want, rel=REL):
    got, want = numpy.asarray(got, dtype=float), numpy.asarray(want, dtype=float)
    assert numpy.all(numpy.abs(got - want) <= rel * numpy.maximum(1.0, numpy.abs(want))), (got, want)


def distinct(a, b, gap=1e-5):
    """Two predicted outcomes differ by far more than the comparison's tolerance."""
    return numpy.max(numpy.abs(numpy.asarray(a) - numpy.asarray(b)) / numpy.maximum(1.0, numpy.abs(b))) > gap


def footprint(fr, xi, yi):
    m = numpy.zeros(SHAPE, dtype=bool)
    m[yi + fr.ry, xi + fr.rx] = True
    return m


# ---------------------------------------------------------------------------
# accepted steps and the lambda schedule


@pytest.mark.parametrize("backend", BACKENDS)
def test_accepted_decrease_then_lambda_three_quarters(backend):
    """
    A decreasing step is accepted at lambda = 1 (the start value) and the
    peak stays RUNNING; lambda then shrinks by 0.75 (multi_fit.c:1318-1320),
    which fixes the second sweep's step.  Misreadings: lambda kept at 1, or
    shrunk by 0.5.
    """
    fr = make_frame((20.3, 18.6, 300.0, 20.0))
    th0 = numpy.array([270.0, 20.6 - fr.xoff, 18.3 - fr.xoff, 22.0])
    xi, yi = fr.origin(th0)
    th1 = fr.step(th0, 1.0, xi, yi)
    d0, d1 = fr.deviance(th0, xi, yi), fr.deviance(th1, xi, yi)
    assert (fr.moved_origin(th1[1], xi), fr.moved_origin(th1[2], yi)) == (xi, yi)
    assert (d0 - d1) / d0 > 1e-3
    th2 = fr.step(th1, 0.75, xi, yi)
    assert fr.deviance(th2, xi, yi) < d1
    assert distinct(th2, fr.step(th1, 1.0, xi, yi)) and distinct(th2, fr.step(th1, 0.5, xi, yi))

    run = Run(backend, fr, [th0])
    s1 = run.sweep()
    assert s1["status"][0] == D.RUNNING and s1["iterations"][0] == 1
    assert s1["counters"] == counters(n_iterations=1)
    assert_theta(s1["theta"][0], th1)
    assert abs(s1["error"][0] - d1) <= 1e-9 * d1
    s2 = run.sweep()
    assert s2["iterations"][0] == 2 and s2["counters"] == counters(n_iterations=2)
    assert_theta(s2["theta"][0], th2)
    run.close()


def overshoot_frame():
    """Height three times the truth: the lambda = 1 step raises the error, lambda = 4 lowers it."""
    fr = make_frame((20.3, 18.6, 300.0, 20.0))
    th0 = numpy.array([848.6, 15.59, 13.02, 33.2])
    xi, yi = fr.origin(th0)
    t1, t4 = fr.step(th0, 1.0, xi, yi), fr.step(th0, 4.0, xi, yi)
    for t in (t1, t4):
        assert (fr.moved_origin(t[1], xi), fr.moved_origin(t[2], yi)) == (xi, yi) and t[0] > 0
    d0, d1, d4 = fr.deviance(th0, xi, yi), fr.deviance(t1, xi, yi), fr.deviance(t4, xi, yi)
    assert 0.05 < (d1 - d0) / d0 < 0.5 and (d0 - d4) / d0 > 1e-3
    return fr, th0, (xi, yi), (t1, t4), (d0, d1, d4)


@pytest.mark.parametrize("backend", BACKENDS)
def test_rejected_increase_retries_with_lambda_four(backend):
    """
    An error increase above tol is rejected: n_non_decr + 1, the peak reset
    to its turn-start parameters and lambda x 4 (mFitResetPeak), and the
    next cycle's step is accepted; afterwards lambda = 4 x 0.75 = 3.
    Misreadings: the reset forgetting lambda (next step at 1), or lambda
    restarting at 1 after acceptance (second sweep at 0.75).  The reset
    also copies back the turn-start cycle count (it restores the whole
    record but `added` and lambda, multi_fit.c:1788-1798): the peak's
    iterations read 1 while n_iterations counts both cycles.
    """
    fr, th0, (xi, yi), (t1, t4), _ = overshoot_frame()
    th2 = fr.step(t4, 3.0, xi, yi)
    assert distinct(th2, fr.step(t4, 0.75, xi, yi)) and distinct(th2, fr.step(t4, 4.0, xi, yi))
    run = Run(backend, fr, [th0])
    s1 = run.sweep()
    assert s1["status"][0] == D.RUNNING and s1["iterations"][0] == 1
    assert s1["counters"] == counters(n_iterations=2, n_non_decr=1)
    assert_theta(s1["theta"][0], t4)
    s2 = run.sweep()
    assert_theta(s2["theta"][0], th2)
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_accepted_increase_below_tol_keeps_the_new_parameters(backend):
    """
    An increase whose relative size (err - start) / start is below tol ends
    the fit as CONVERGED with the increased-error parameters kept
    (multi_fit.c:1259-1283).  Misreading: converging at the turn-start
    parameters.
    """
    fr, th0, _, (t1, t4), (d0, d1, d4) = overshoot_frame()
    tol = 1.1 * (d1 - d0) / d0
    run = Run(backend, fr, [th0], tol=tol)
    s1 = run.sweep()
    assert s1["status"][0] == D.CONVERGED and s1["iterations"][0] == 1
    assert s1["counters"] == counters(n_iterations=1)
    assert_theta(s1["theta"][0], t1)
    assert abs(s1["error"][0] - d1) <= 1e-9 * d1 and s1["error"][0] > d0
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_increase_tolerance_is_relative_to_the_start_error(backend):
    """
    (err - start) / start, not / err: with tol between the two ratios the
    increase is rejected (the / err reading would converge).
    """
    fr, th0, _, (t1, t4), (d0, d1, d4) = overshoot_frame()
    by_start, by_err = (d1 - d0) / d0, (d1 - d0) / d1
    tol = 0.5 * (by_start + by_err)
    run = Run(backend, fr, [th0], tol=tol)
    s1 = run.sweep()
    assert s1["counters"]["n_non_decr"] == 1 and s1["iterations"][0] == 1
    assert_theta(s1["theta"][0], t4)
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_decrease_below_tol_converges_relative_to_the_start_error(backend):
    """
    A decrease with (start - err) / start < tol converges with the new
    parameters; the / err reading would not (tol between the two ratios).
    """
    fr = make_frame((20.3, 18.6, 300.0, 20.0))
    th0 = numpy.array([270.0, 20.6 - fr.xoff, 18.3 - fr.xoff, 22.0])
    xi, yi = fr.origin(th0)
    th1 = fr.step(th0, 1.0, xi, yi)
    d0, d1 = fr.deviance(th0, xi, yi), fr.deviance(th1, xi, yi)
    tol = 0.5 * ((d0 - d1) / d0 + (d0 - d1) / d1)
    run = Run(backend, fr, [th0], tol=tol)
    s1 = run.sweep()
    assert s1["status"][0] == D.CONVERGED and s1["iterations"][0] == 1
    assert_theta(s1["theta"][0], th1)
    run.close()


# ---------------------------------------------------------------------------
# rejected candidates


@pytest.mark.parametrize("backend", BACKENDS)
def test_step_onto_the_far_margin_is_rejected(backend):
    """
    The lambda = 1 step moves the ROI origin (X - xi > 1.5) to exactly
    W - r, which mFitCheck refuses (xi >= W - r, multi_fit.c:482):
    n_margin + 1, reset, lambda x 4, and the smaller step stays.
    Misreading: an exclusive bound (xi > W - r) would accept the first step.
    """
    fr = make_frame((42.5, 18.6, 600.0, 20.0), sigma=2.0)
    th0 = numpy.array([600.0, 30.9, 18.6 - fr.xoff, 20.0])
    xi, yi = fr.origin(th0)
    t1, t4 = fr.step(th0, 1.0, xi, yi), fr.step(th0, 4.0, xi, yi)
    assert fr.moved_origin(t1[1], xi) == SHAPE[1] - R
    assert fr.moved_origin(t4[1], xi) == xi and fr.deviance(t4, xi, yi) < fr.deviance(th0, xi, yi)
    run = Run(backend, fr, [th0], sigma=2.0)
    s1 = run.sweep()
    assert s1["status"][0] == D.RUNNING and s1["iterations"][0] == 1   # reset restores the count
    assert s1["counters"] == counters(n_iterations=2, n_margin=1)
    assert_theta(s1["theta"][0], t4)
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_negative_height_is_rejected(backend):
    """No emitter under the peak: the lambda = 1 step drives the height below zero."""
    fr = make_frame((20.3, 18.6, 0.0, 20.0))
    th0 = numpy.array([200.0, 20.3 - fr.xoff, 18.6 - fr.xoff, 20.0])
    xi, yi = fr.origin(th0)
    t1, t4 = fr.step(th0, 1.0, xi, yi), fr.step(th0, 4.0, xi, yi)
    assert t1[0] < 0 < t4[0] and fr.deviance(t4, xi, yi) < fr.deviance(th0, xi, yi)
    run = Run(backend, fr, [th0])
    s1 = run.sweep()
    assert s1["status"][0] == D.RUNNING and s1["iterations"][0] == 1   # reset restores the count
    assert s1["counters"] == counters(n_iterations=2, n_neg_height=1)
    assert_theta(s1["theta"][0], t4)
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_negative_fit_value_is_rejected(backend):
    """
    Background near one photon, start background 20: the lambda = 1 step
    makes bg negative, so f <= 0 off the peak (multi_fit.c:148-163): the
    candidate is subtracted again, n_neg_fi + 1, reset, lambda x 4.
    """
    fr = make_frame((20.3, 18.6, 300.0, 1.5))
    th0 = numpy.array([300.0, 20.3 - fr.xoff, 18.6 - fr.xoff, 20.0])
    xi, yi = fr.origin(th0)
    t1, t4 = fr.step(th0, 1.0, xi, yi), fr.step(th0, 4.0, xi, yi)
    assert t1[0] > 0 and fr.deviance(t1, xi, yi) is None
    assert fr.deviance(t4, xi, yi) < fr.deviance(th0, xi, yi)
    run = Run(backend, fr, [th0])
    s1 = run.sweep()
    assert s1["status"][0] == D.RUNNING and s1["iterations"][0] == 1   # reset restores the count
    assert s1["counters"] == counters(n_iterations=2, n_neg_fi=1)
    assert_theta(s1["theta"][0], t4)
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_singular_system_climbs_lambda_until_lost(backend):
    """
    Height 0: the x and y rows of H vanish and every Cholesky fails
    (multi_fit.c:1194-1204): n_dposv + 1 and lambda x 4 per cycle, from
    lambda = 1 until lambda > 1e20 (checked before the cycle is counted,
    :1152-1157): 4^34 > 1e20 > 4^33, so 34 cycles, n_lost + 1, ERROR; no
    reset runs on this path, so the peak's own count is 34 as well.
    Misreadings: lambda >= 1e20 (also 34), an up factor of 10 (21 cycles),
    counting the final check (35).
    """
    fr = make_frame((20.3, 18.6, 300.0, 20.0))
    th0 = numpy.array([0.0, 20.3 - fr.xoff, 18.6 - fr.xoff, 20.0])
    _, H = fr.jh(th0, *fr.origin(th0))
    assert H[1, 1] == 0.0 and H[2, 2] == 0.0
    assert 4.0 ** 33 < 1e20 < 4.0 ** 34
    run = Run(backend, fr, [th0])
    s1 = run.sweep()
    assert s1["status"][0] == D.ERROR and s1["iterations"][0] == 34
    assert s1["counters"] == counters(n_iterations=34, n_dposv=34, n_lost=1)
    run.close()


# ---------------------------------------------------------------------------
# ROI moves: inside the cycle and after the sweep


@pytest.mark.parametrize("backend", BACKENDS)
def test_roi_moves_inside_the_cycle_past_one_and_a_half(backend):
    """
    X - xi > 1.5 after the update moves the ROI at once (mFitUpdate): the
    candidate's error is taken over the new ROI, and the fit image holds the
    peak there.  Misreading: moving only after the sweep (error over the old ROI).
    """
    fr = make_frame((22.9, 18.6, 600.0, 20.0), sigma=2.0)
    th0 = numpy.array([600.0, 20.9 - fr.xoff, 18.6 - fr.xoff, 20.0])
    xi, yi = fr.origin(th0)
    t1 = fr.step(th0, 1.0, xi, yi)
    nx = fr.moved_origin(t1[1], xi)
    assert t1[1] - xi > 1.5 and nx == int(numpy.floor(t1[1]))
    d_new, d_old = fr.deviance(t1, nx, yi), fr.deviance(t1, xi, yi)
    assert d_new < fr.deviance(th0, xi, yi) and abs(d_new - d_old) > 1e-3 * d_new
    run = Run(backend, fr, [th0], sigma=2.0)
    s1 = run.sweep()
    assert s1["iterations"][0] == 1 and s1["status"][0] == D.RUNNING
    assert_theta(s1["theta"][0], t1)
    assert abs(s1["error"][0] - d_new) <= 1e-9 * d_new
    numpy.testing.assert_array_equal(s1["fit"] != 0, footprint(fr, nx, yi))
    run.close()


@pytest.mark.parametrize("backend,xt,moves", [
    ("oracle", 24.5, True), ("oracle", 24.0, False),
    pytest.param("hip", 24.5, True, marks=pytest.mark.gpu), pytest.param("hip", 24.0, False, marks=pytest.mark.gpu)])
def test_recentering_after_the_sweep_uses_a_quarter_pixel_band(backend, xt, moves):
    """
    After the sweep a RUNNING peak whose X - xi left [-0.25, 1.25] gets its
    ROI moved to floor(X) (mFitRecenterPeaks); 1.25 < X - xi <= 1.5 is
    inside the cycle's band, so only the recentering moves it.  The error
    stays the turn's (over the old ROI).  X - xi just below 1.25 stays put.
    """
    fr = make_frame((xt, 18.6, 600.0, 20.0), sigma=2.5)
    th0 = numpy.array([600.0, 20.05 - fr.xoff, 18.6 - fr.xoff, 20.0])
    xi, yi = fr.origin(th0)
    t1 = fr.step(th0, 1.0, xi, yi)
    dx = t1[1] - xi
    assert (1.25 < dx <= 1.5) if moves else (1.0 < dx < 1.25)
    d1 = fr.deviance(t1, xi, yi)
    assert d1 < fr.deviance(th0, xi, yi)
    run = Run(backend, fr, [th0], sigma=2.5)
    s1 = run.sweep()
    assert s1["status"][0] == D.RUNNING
    assert_theta(s1["theta"][0], t1)
    assert abs(s1["error"][0] - d1) <= 1e-9 * d1
    nx = int(numpy.floor(t1[1])) if moves else xi
    numpy.testing.assert_array_equal(s1["fit"] != 0, footprint(fr, nx, yi))
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_recentering_off_the_image_loses_the_peak(backend):
    """
    A step to X - xi < -0.25 (but > -0.5) near the left edge is accepted in
    the cycle; the recentering then moves the origin to -1, mFitCheck
    refuses it (n_margin + 1) and the peak is lost (n_lost + 1, ERROR),
    already subtracted: the fit image is empty.
    """
    fr = make_frame((3.0, 18.6, 300.0, 20.0))
    th0 = numpy.array([300.0, 0.3, 18.6 - fr.xoff, 20.0])
    xi, yi = fr.origin(th0)
    t1 = fr.step(th0, 1.0, xi, yi)
    assert xi == 0 and -0.5 < t1[1] < -0.25
    assert fr.deviance(t1, xi, yi) < fr.deviance(th0, xi, yi)
    run = Run(backend, fr, [th0])
    s1 = run.sweep()
    assert s1["status"][0] == D.ERROR and s1["iterations"][0] == 1
    assert s1["counters"] == counters(n_iterations=1, n_margin=1, n_lost=1)
    assert not numpy.any(s1["fit"])
    run.close()


# ---------------------------------------------------------------------------
# the 2D width clamp


@pytest.mark.parametrize("backend", BACKENDS)
def test_2d_width_clamp_leaves_the_y_width_unclamped(backend):
    """
    daoUpdate2D copies XWIDTH to YWIDTH before daoCheck2D clamps XWIDTH
    alone (dao_fit.c:1061-1085, 1798-1815): a step past width_max leaves
    XWIDTH = width_max and YWIDTH at the unclamped value.  Misreadings:
    both clamped, or the clamp refusing the step.
    """
    sigma_range = [1.45, 5.0]
    # the emitter (sigma 1.2) is narrower than the clamp allows (1.45)
    fr = make_frame((20.3, 18.6, 300.0, 20.0), sigma=1.2, model="2d", sigma_range=sigma_range)
    th0 = numpy.array([300.0, 20.3 - fr.xoff, 18.6 - fr.xoff, 1.0 / (2 * 1.6 ** 2), 20.0])
    xi, yi = fr.origin(th0)
    t1 = fr.step(th0, 1.0, xi, yi)
    assert t1[3] > fr.wmax
    run = Run(backend, fr, [th0], sigma_range=sigma_range)
    s1 = run.sweep()
    assert s1["iterations"][0] == 1 and s1["counters"] == counters(n_iterations=1)
    assert abs(s1["xsigma"][0] - sigma_range[0]) <= 1e-12
    assert abs(s1["ysigma"][0] - numpy.sqrt(1.0 / (2.0 * t1[3]))) <= 1e-9
    assert_theta([s1["height"][0], s1["x"][0] - fr.xoff, s1["y"][0] - fr.xoff, s1["background"][0]],
                 t1[[0, 1, 2, 4]])
    run.close()


# ---------------------------------------------------------------------------
# two overlapping peaks: update order and the shared background


def two_peaks():
    fr = make_frame((20.3, 18.6, 300.0, 20.0))
    rng = numpy.random.default_rng(9)
    yy, xx = numpy.mgrid[0:SHAPE[0], 0:SHAPE[1]]
    # a second emitter 3.5 px away, 25 photons of background
    fr.x = numpy.maximum(fr.x + rng.poisson(250.0 * numpy.exp(-((xx - 23.6) ** 2 + (yy - 19.9) ** 2) /
                                                               (2 * 1.5 ** 2)) * fr.rqe), 1.0)
    p0 = numpy.array([260.0, 20.0 - fr.xoff, 18.9 - fr.xoff, 18.0])
    p1 = numpy.array([280.0, 23.9 - fr.xoff, 19.6 - fr.xoff, 26.0])
    return fr, [p0, p1]


@pytest.mark.parametrize("backend", BACKENDS)
def test_overlapping_peaks_see_the_mean_background(backend):
    """
    Where two ROIs overlap the model's background is the mean of the two
    peaks' bg rqe + scmos (multi_fit.c:54, 143): each peak's starting error.
    Misreadings: the sum, or each peak's own background.
    """
    fr, starts = two_peaks()
    sc = Scene(fr, [(th, *fr.origin(th)) for th in starts])
    want = [sc.deviance(0), sc.deviance(1)]
    # both ROIs cover a good part of each other
    rows, cols = sc.peaks[0][2] + fr.ry, sc.peaks[0][1] + fr.rx
    shared = sc._cover(sc.peaks[1][1], sc.peaks[1][2], rows, cols)
    assert numpy.count_nonzero(shared) > 20
    # the misreadings: backgrounds summed where ROIs overlap, or each peak's own
    f = sc.fit(rows, cols)
    other = starts[1][3] * fr.rqe[rows, cols] + fr.scmos[rows, cols]
    own = starts[0][3] * fr.rqe[rows, cols] + fr.scmos[rows, cols]
    f_sum = f + numpy.where(shared, 0.5 * (own + other), 0.0)   # mean -> sum
    x = fr.x[rows, cols]
    summed = numpy.sum(2.0 * ((f_sum - x) - x * numpy.log(f_sum / x)))
    solo = [Frame.deviance(fr, th, *fr.origin(th)) for th in starts]
    assert abs(summed - want[0]) > 1e-3 * want[0]
    assert all(abs(solo[k] - want[k]) > 1e-3 * want[k] for k in range(2))
    run = Run(backend, fr, starts)
    s0 = run.state()
    for k in range(2):
        assert abs(s0["error"][k] - want[k]) <= 1e-9 * want[k]
    run.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_overlapping_peaks_update_in_index_order(backend):
    """
    Peaks update one at a time in index order, each seeing the earlier
    peaks' updates of this sweep (multi_fit.c:1089).  Misreading: both
    steps taken from the sweep's start state.
    """
    fr, starts = two_peaks()
    sc = Scene(fr, [(th, *fr.origin(th)) for th in starts])
    jacobi1 = sc.step(1, 1.0)
    d0 = sc.deviance(0)
    t0 = sc.step(0, 1.0)
    sc.update(0, t0)
    assert sc.deviance(0) < d0
    d1 = sc.deviance(1)
    t1 = sc.step(1, 1.0)
    sc.update(1, t1)
    assert sc.deviance(1) < d1
    assert distinct(t1, jacobi1)
    run = Run(backend, fr, starts)
    s1 = run.sweep()
    assert list(s1["status"]) == [D.RUNNING, D.RUNNING] and list(s1["iterations"]) == [1, 1]
    assert_theta(s1["theta"][0], t0)
    assert_theta(s1["theta"][1], t1)
    run.close()
