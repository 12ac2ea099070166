"""
The multiplane fitter on the GPU against the multiplane oracle, at sizes
beyond the pins: 2, 4 and 8 channels related by general affine maps (small
rotations, scales and shifts, as a plane-to-plane registration gives),
tens to hundreds of overlapping emitters whose brightness differs per
channel, some near the edges of some channels (their groups fail
everywhere).  Bars as for the single-channel fitter: identical status flags
and per-peak cycle counts in every channel; x, y within 1e-3 px; height,
background and width within 1e-4 relative; fit images and residuals within
1e-6 of their largest value.
"""
import numpy
import pytest

from conftest import make_mp_fitter
from storm_analysis_amd.multi_plane import mp_fit_dao_c as MPD
from storm_analysis_amd.sa_library import dao_fit_c as D

TOL_POS = 1e-3
TOL_REL = 1e-4


def affine_maps(rng, n, shape):
    """Channel 0 -> c maps (and their inverses) as [n, 3] tables: rotation, scale, shift."""
    xt, yt, xi, yi = (numpy.zeros((n, 3)) for _ in range(4))
    for c in range(n):
        if c == 0:
            a, b, tx, ty = 1.0, 0.0, 0.0, 0.0
        else:
            th = rng.uniform(-0.03, 0.03)
            s = rng.uniform(0.98, 1.02)
            a, b = s * numpy.cos(th), s * numpy.sin(th)
            tx, ty = rng.uniform(-4, 4), rng.uniform(-4, 4)
        A = numpy.array([[a, -b], [b, a]])
        t = numpy.array([tx, ty])
        Ai = numpy.linalg.inv(A)
        ti = -Ai @ t
        xt[c] = [t[0], A[0, 0], A[0, 1]]
        yt[c] = [t[1], A[1, 0], A[1, 1]]
        xi[c] = [ti[0], Ai[0, 0], Ai[0, 1]]
        yi[c] = [ti[1], Ai[1, 0], Ai[1, 1]]
    return xt, yt, xi, yi


def scene(seed, n_ch, size, n_em, sigma=1.5):
    rng = numpy.random.default_rng(seed)
    shape = (size, size)
    xt, yt, xi, yi = affine_maps(rng, n_ch, shape)
    ex = rng.uniform(3.0, size - 3.0, n_em)
    ey = rng.uniform(3.0, size - 3.0, n_em)
    h0 = rng.uniform(300.0, 1500.0, n_em)
    yy, xx = numpy.mgrid[0:size, 0:size]
    chans = []
    for c in range(n_ch):
        rqe = rng.uniform(0.8, 1.2, shape)
        scmos = rng.uniform(0.0, 2.0, shape)
        cx = xt[c, 0] + xt[c, 1] * ex + xt[c, 2] * ey
        cy = yt[c, 0] + yt[c, 1] * ex + yt[c, 2] * ey
        hc = h0 * rng.uniform(0.3, 1.0, n_em)   # the planes see each emitter differently
        bg = rng.uniform(10.0, 40.0)
        mean = numpy.full(shape, bg)
        for k in range(n_em):
            r0, r1 = int(max(cy[k] - 8, 0)), int(min(cy[k] + 9, size))
            c0, c1 = int(max(cx[k] - 8, 0)), int(min(cx[k] + 9, size))
            if r0 >= r1 or c0 >= c1:
                continue
            mean[r0:r1, c0:c1] += hc[k] * numpy.exp(-((xx[r0:r1, c0:c1] - cx[k]) ** 2 +
                                                       (yy[r0:r1, c0:c1] - cy[k]) ** 2) / (2 * sigma ** 2))
        x = numpy.maximum(rng.poisson(mean * rqe + scmos).astype(float), 1.0)
        chans.append({"rqe": rqe, "scmos": scmos, "image": (x - scmos) / rqe, "bg": bg})
    start = {"x": ex + rng.normal(0, 0.3, n_em), "y": ey + rng.normal(0, 0.3, n_em), "z": numpy.zeros(n_em),
             "height": h0 * 0.6, "background": numpy.full(n_em, 25.0), "xsigma": numpy.full(n_em, 1.6),
             "ysigma": numpy.full(n_em, 1.6)}
    return shape, (xt, yt, xi, yi), chans, start


def make(backend, shape, maps, chans, start, roi):
    m = make_mp_fitter(MPD.MPFitDao2D, backend, n_channels=len(chans), roi_size=roi, sigma_range=[0.8, 3.0])
    for c, ch in enumerate(chans):
        m.initializeChannel(ch["rqe"], ch["scmos"], c)
    m.setMapping(*maps)
    m.setWeights()
    m.newImage([ch["image"] for ch in chans])
    m.newBackground([numpy.zeros(shape) for _ in chans])
    m.newPeaks(start, "text")
    return m


KEYS = ("x", "y", "height", "background", "xsigma", "ysigma", "status", "iterations")


def read(m, n_ch):
    out = []
    for c in range(n_ch):
        p = {k: m.getPeakProperty(k, channel=c) for k in KEYS}
        p["fit"] = m.getFitImage()[c]
        p["residual"] = m.getResidual()[c]
        out.append(p)
    return out


def assert_parity(g, r, what):
    assert g["x"].size == r["x"].size, what
    numpy.testing.assert_array_equal(g["status"], r["status"], err_msg=what)
    numpy.testing.assert_array_equal(g["iterations"], r["iterations"], err_msg=what)
    ok = r["status"] != D.ERROR
    for k in ("x", "y"):
        d = numpy.abs(g[k] - r[k])[ok]
        assert d.size == 0 or d.max() < TOL_POS, (what, k, d.max())
    for k in ("height", "background", "xsigma", "ysigma"):
        d = (numpy.abs(g[k] - r[k]) / numpy.abs(r[k]))[ok]
        assert d.size == 0 or d.max() < TOL_REL, (what, k, d.max())
    for k in ("fit", "residual"):
        scale = numpy.max(numpy.abs(r[k]))
        assert numpy.max(numpy.abs(g[k] - r[k])) <= 1e-6 * scale, (what, k)


CASES = [(2, 128, 60, 12), (4, 256, 400, 18), (8, 96, 30, 12)]


@pytest.mark.gpu
@pytest.mark.parametrize("n_ch,size,n_em,roi", CASES, ids=["2ch-128-roi12", "4ch-256-roi18", "8ch-96-roi12"])
def test_multiplane_fit_matches_the_oracle(n_ch, size, n_em, roi):
    shape, maps, chans, start = scene(100 + n_ch, n_ch, size, n_em)
    out, counts = {}, {}
    for backend in ("oracle", "hip"):
        m = make(backend, shape, maps, chans, start, roi)
        m.doFit()
        out[backend] = read(m, n_ch)
        counts[backend] = m.getIterations()
        m.cleanup(verbose=False)
    assert counts["hip"] == counts["oracle"]
    n_err = int(numpy.sum(out["oracle"][0]["status"] == D.ERROR))
    print("%d channels, %d groups, %d in ERROR, %d cycles" % (n_ch, start["x"].size, n_err, counts["oracle"]))
    assert n_err < start["x"].size
    for c in range(n_ch):
        assert_parity(out["hip"][c], out["oracle"][c], "channel %d" % c)


@pytest.mark.gpu
def test_multiplane_sweeps_one_at_a_time_match_the_oracle():
    """iterate() (mpIterateLM) five times, compared after every sweep, then compaction of ERROR groups."""
    shape, maps, chans, start = scene(7, 3, 128, 80)
    fitters = {b: make(b, shape, maps, chans, start, 12) for b in ("oracle", "hip")}
    for k in range(5):
        got = {}
        for b, m in fitters.items():
            m.iterate()
            got[b] = read(m, 3)
        for c in range(3):
            assert_parity(got["hip"][c], got["oracle"][c], "sweep %d channel %d" % (k, c))
    for b, m in fitters.items():
        m.removeErrorPeaks()
    for c in range(3):
        assert fitters["hip"].getNFit(c) == fitters["oracle"].getNFit(c)
    for m in fitters.values():
        m.cleanup(verbose=False)
