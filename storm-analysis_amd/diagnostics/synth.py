"""
Seeded synthetic SMLM frames and starting peaks for the fitter's workloads.

Conventions (the fitter's own, dao_fit.c:968-998): pixel (row i, column j)
has its centre at x = j, y = i; an emitter at (x, y) with height h and widths
sx, sy contributes h * exp(-(j-x)^2/(2 sx^2) - (i-y)^2/(2 sy^2)) evaluated at
pixel centres (no pixel integration).  Frames go through a camera model
(Poisson photons, ADU = gain*e + offset, then (ADU - offset)/gain clamped to
>= 1, analysis_io.py:243-260) and are returned as float64 photo-electrons.

Z calibration: w(z) = wo*sqrt(1 + X^2 + A X^3 + B X^4), X = (z-c)/d
(z_calibration.py:352-358); sigma = w/2, so that the fitter's
XWIDTH = 1/(2 sigma^2) = 2/w^2 (dao_fit.c:1019-1034).
"""
import numpy

SEED_BASE = 20261203

# Z calibration of the 3D-astigmatism workload: [wo(px), c(um), d(um), A, B].
WX_Z = numpy.array([2.0, 0.15, 0.4, 0.0, 0.0])
WY_Z = numpy.array([2.0, -0.15, 0.4, 0.0, 0.0])


def roi_size_for(sigma, fixed):
    """ROI size rule of fitting.py:376-386."""
    r = int(8.0 * sigma)
    if not fixed:
        r = int(1.5 * r)
    return r


def width_from_z(z, p):
    X = (z - p[1]) / p[2]
    return p[0] * numpy.sqrt(1.0 + X * X + p[3] * X ** 3 + p[4] * X ** 4)


def render(shape, x, y, h, sx, sy, bg):
    """
    Expected photo-electrons: background image (or scalar) plus Gaussians.
    """
    H, W = shape
    img = numpy.zeros(shape, dtype=numpy.float64) + bg
    if x.size == 0:
        return img
    half = int(numpy.ceil(5.0 * max(sx.max(), sy.max()))) + 1
    k = numpy.arange(-half, half + 1)
    cx = numpy.floor(x).astype(numpy.int64)
    cy = numpy.floor(y).astype(numpy.int64)
    cols = cx[:, None] + k[None, :]                 # (N, K)
    rows = cy[:, None] + k[None, :]
    gx = numpy.exp(-((cols - x[:, None]) ** 2) / (2.0 * sx[:, None] ** 2))
    gy = numpy.exp(-((rows - y[:, None]) ** 2) / (2.0 * sy[:, None] ** 2))
    vals = h[:, None, None] * gy[:, :, None] * gx[:, None, :]   # (N, K, K)
    rr = numpy.broadcast_to(rows[:, :, None], vals.shape)
    cc = numpy.broadcast_to(cols[:, None, :], vals.shape)
    ok = (rr >= 0) & (rr < H) & (cc >= 0) & (cc < W)
    numpy.add.at(img, (rr[ok], cc[ok]), vals[ok])
    return img


def camera(rng, expected, offset=100.0, gain=1.0):
    """Poisson photons -> ADU -> photo-electrons clamped at 1 (analysis_io.py:243-260)."""
    e = rng.poisson(expected).astype(numpy.float64)
    adu = e * gain + offset
    return numpy.maximum((adu - offset) / gain, 1.0)


def scmos_camera(rng, expected, offset, gain, variance):
    """
    sCMOS pixel model: ADU = gain*Poisson(e) + offset + N(0, variance);
    returns (frame in e-, scmos_cal = variance/gain^2) (find_peaks.py:44-45).
    """
    e = rng.poisson(expected).astype(numpy.float64)
    adu = gain * e + offset + rng.normal(size=expected.shape) * numpy.sqrt(variance)
    frame = numpy.maximum((adu - offset) / gain, 1.0)
    return frame, variance / (gain * gain)


def start_peaks_hdf5(rng, img, x, y, z, bg_at, sx, sy, jitter=1.0):
    """
    "hdf5" starting peaks: truth +/- U(-jitter, jitter) px, true background,
    height = brightest pixel of the 3x3 around the start minus background.
    """
    H, W = img.shape
    xs = x + rng.uniform(-jitter, jitter, x.size)
    ys = y + rng.uniform(-jitter, jitter, y.size)
    ci = numpy.clip(numpy.round(xs).astype(int), 1, W - 2)
    ri = numpy.clip(numpy.round(ys).astype(int), 1, H - 2)
    best = numpy.full(x.size, -numpy.inf)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            best = numpy.maximum(best, img[ri + di, ci + dj])
    hgt = numpy.maximum(best - bg_at, 1.0)
    return {"x": xs, "y": ys, "z": z.astype(numpy.float64), "background": bg_at.astype(numpy.float64),
            "height": hgt, "xsigma": sx.astype(numpy.float64), "ysigma": sy.astype(numpy.float64)}


class Frame(object):
    """One synthetic frame: image, camera terms, truth and starting peaks."""
    def __init__(self, image, peaks, truth, rqe=None, scmos_cal=None):
        self.image = image
        self.peaks = peaks
        self.truth = truth
        self.rqe = rqe
        self.scmos_cal = scmos_cal


def frame_2dfixed(rng, size=256, n=200, sigma=1.5, bg=100.0, hmin=500.0, hmax=3000.0):
    """2D-fixed workload: uniform emitters, sigma 1.5, heights U(500,3000), bg 100."""
    r = roi_size_for(sigma, True)
    m = r // 2 + 2
    x = rng.uniform(m, size - m, n)
    y = rng.uniform(m, size - m, n)
    h = rng.uniform(hmin, hmax, n)
    s = numpy.full(n, sigma)
    img = camera(rng, render((size, size), x, y, h, s, s, bg))
    peaks = start_peaks_hdf5(rng, img, x, y, numpy.zeros(n), numpy.full(n, bg), s, s)
    return Frame(img, peaks, {"x": x, "y": y, "height": h, "sigma": s})


def z_background(shape):
    """The Z workload's background, 125 + 75 sin(2 pi x / W) cos(2 pi y / H), in [50, 200]."""
    h, w = shape
    jj, ii = numpy.meshgrid(numpy.arange(w), numpy.arange(h))
    return 125.0 + 75.0 * numpy.sin(2 * numpy.pi * jj / w) * numpy.cos(2 * numpy.pi * ii / h)


def frame_z(rng, size=512, n=1000):
    """
    3D-astigmatism workload: z ~ U(-0.5,0.5) um, widths from WX_Z/WY_Z,
    heights lognormal (median 1500, s 0.5), background 125 + 75 sin cos.
    """
    r = 18
    m = r // 2 + 2
    x = rng.uniform(m, size - m, n)
    y = rng.uniform(m, size - m, n)
    z = rng.uniform(-0.5, 0.5, n)
    h = numpy.exp(numpy.log(1500.0) + 0.5 * rng.normal(size=n))
    sx = 0.5 * width_from_z(z, WX_Z)
    sy = 0.5 * width_from_z(z, WY_Z)
    jj, ii = numpy.meshgrid(numpy.arange(size), numpy.arange(size))
    bgimg = 125.0 + 75.0 * numpy.sin(2 * numpy.pi * jj / size) * numpy.cos(2 * numpy.pi * ii / size)
    img = camera(rng, render((size, size), x, y, h, sx, sy, bgimg))
    bg_at = 125.0 + 75.0 * numpy.sin(2 * numpy.pi * x / size) * numpy.cos(2 * numpy.pi * y / size)
    s0 = numpy.full(n, 1.5)
    peaks = start_peaks_hdf5(rng, img, x, y, numpy.zeros(n), bg_at, s0, s0)
    return Frame(img, peaks, {"x": x, "y": y, "z": z, "height": h})


def frame_2d_dense(rng, size=2048, n=20000, smin=1.2, smax=2.0, hmin=200.0, hmax=5000.0, bg=100.0):
    """High-density 2D workload: sigma U(1.2,2.0), heights U(200,5000), bg 100."""
    r = 18
    m = r // 2 + 2
    x = rng.uniform(m, size - m, n)
    y = rng.uniform(m, size - m, n)
    s = rng.uniform(smin, smax, n)
    h = rng.uniform(hmin, hmax, n)
    img = camera(rng, render((size, size), x, y, h, s, s, bg))
    s0 = numpy.full(n, 0.5 * (smin + smax))
    peaks = start_peaks_hdf5(rng, img, x, y, numpy.zeros(n), numpy.full(n, bg), s0, s0)
    return Frame(img, peaks, {"x": x, "y": y, "height": h, "sigma": s})


def frame_3d_scmos(rng, size=256, density=1.0, pixel_um=0.16, smin=1.2, smax=2.0):
    """
    sCMOS 3D workload: emitters at `density` per um^2, free sigma in x and y,
    per-pixel offset N(100,5), gain N(2,0.1), read variance lognormal(ln 4, 0.8).
    """
    r = 18
    m = r // 2 + 2
    area = (size * pixel_um) ** 2
    n = max(1, int(round(density * area)))
    x = rng.uniform(m, size - m, n)
    y = rng.uniform(m, size - m, n)
    sx = rng.uniform(smin, smax, n)
    sy = rng.uniform(smin, smax, n)
    h = rng.uniform(500.0, 3000.0, n)
    offset = rng.normal(100.0, 5.0, (size, size))
    gain = rng.normal(2.0, 0.1, (size, size))
    var = numpy.exp(numpy.log(4.0) + 0.8 * rng.normal(size=(size, size)))
    img, scmos_cal = scmos_camera(rng, render((size, size), x, y, h, sx, sy, 100.0), offset, gain, var)
    s0 = numpy.full(n, 0.5 * (smin + smax))
    peaks = start_peaks_hdf5(rng, img, x, y, numpy.zeros(n), numpy.full(n, 100.0), s0, s0)
    return Frame(img, peaks, {"x": x, "y": y, "sx": sx, "sy": sy, "height": h},
                 rqe=numpy.ones((size, size)), scmos_cal=scmos_cal)
