"""
An evaluation of one Gaussian peak's LM step that shares no code and no
arithmetic with the two restatements of the fitter (oracle/dao_oracle.c and
tests/lm_pure.py): the model and the MLE deviance in closed form over the
reference's circular ROI, the gradient of the deviance by complex-step
differentiation (exact to rounding), the Gauss-Newton matrix
sum 2 x / f^2 jt jt^T (daoCalcJH*, dao_fit.c:48-105, 409-467) from
complex-step derivatives jt of the model, and the damped step
(H + lambda diag H) delta = J (multi_fit.c:1176-1190) from
numpy.linalg.solve.  It evaluates steps; the branch a step takes is decided
by each test, from the reference's text.

Parameterisation (the fitter's): XCENTER = x - xoff with xoff = r/2 - 1,
ROI origin xi = floor(XCENTER) (dao_fit.c:1701-1706), shape
exp(-(col - XCENTER - xoff)^2 XW) exp(-(row - YCENTER - xoff)^2 YW) at
pixel centres (dao_fit.c:968-998), model h psf rqe + bg rqe + scmos for a
peak alone (multi_fit.c:28-57).  2dfixed parameters (J order): h, X, Y, bg
with a fixed width; 2d: h, X, Y, W, bg (XW = YW = W).
"""
import numpy

NAMES = {"2dfixed": ("h", "X", "Y", "bg"), "2d": ("h", "X", "Y", "W", "bg")}
CS = 1e-30  # complex-step size


def roi_offsets(r):
    """(rows, cols) of the circular ROI, row-major (multi_fit.c:1043-1059)."""
    c = 0.5 * r
    i, j = numpy.mgrid[0:r, 0:r]
    m = ((j - c + 0.5) ** 2 + (i - c + 0.5) ** 2) <= c * c
    return i[m], j[m]


class Frame(object):
    """A fitter-space image x = image rqe + scmos with its calibration."""
    def __init__(self, x, rqe, scmos, r, model="2dfixed", width=None, wmin=None, wmax=None):
        self.x, self.rqe, self.scmos = x, rqe, scmos
        self.r, self.model = r, model
        self.xoff = 0.5 * r - 1.0
        self.width = width          # 2dfixed: XW = YW
        self.wmin, self.wmax = wmin, wmax
        self.ry, self.rx = roi_offsets(r)

    def widths(self, th):
        w = self.width if self.model == "2dfixed" else th[3]
        return w, w

    def fit(self, th, xi, yi):
        """Model over the ROI at origin (xi, yi); complex-safe."""
        rows, cols = yi + self.ry, xi + self.rx
        wx, wy = self.widths(th)
        bg = th[-1]
        e = numpy.exp(-((cols - th[1] - self.xoff) ** 2) * wx) * numpy.exp(-((rows - th[2] - self.xoff) ** 2) * wy)
        q, s = self.rqe[rows, cols], self.scmos[rows, cols]
        return th[0] * e * q + bg * q + s

    def data(self, xi, yi):
        return self.x[yi + self.ry, xi + self.rx]

    def deviance(self, th, xi, yi):
        """MLE deviance sum 2[(f - x) - x ln(f / x)] (multi_fit.c:111-195); None if some f <= 0."""
        f = self.fit(th, xi, yi)
        if numpy.any(numpy.real(f) <= 0):
            return None
        x = self.data(xi, yi)
        return numpy.sum(2.0 * ((f - x) - x * numpy.log(f / x)))

    def jh(self, th, xi, yi):
        """J = dD/dtheta and the Gauss-Newton H, both by complex steps."""
        n = len(th)
        th = numpy.asarray(th, dtype=complex)
        x = self.data(xi, yi)
        f = numpy.real(self.fit(th, xi, yi))
        J = numpy.zeros(n)
        jt = numpy.zeros((n, x.size))
        for k in range(n):
            t = th.copy()
            t[k] += 1j * CS
            J[k] = numpy.imag(self.deviance(t, xi, yi)) / CS
            jt[k] = numpy.imag(self.fit(t, xi, yi)) / CS
        H = (jt * (2.0 * x / (f * f))) @ jt.T
        return J, H

    def step(self, th, lam, xi, yi):
        """theta - delta of the damped system at lambda, before the model's checks."""
        J, H = self.jh(th, xi, yi)
        A = H.copy()
        A[numpy.diag_indices_from(A)] *= (1.0 + lam)
        return numpy.asarray(th, dtype=float) - numpy.linalg.solve(A, J)

    @staticmethod
    def moved_origin(X, xi):
        """ROI origin after an update (multi_fit.c:1933-1946: hysteresis [-0.5, 1.5])."""
        d = X - xi
        return int(numpy.floor(X)) if (d < -0.5 or d > 1.5) else xi

    def origin(self, th):
        return int(numpy.floor(th[1])), int(numpy.floor(th[2]))


class Scene(object):
    """
    Several peaks on one Frame (2dfixed): foregrounds add, and each pixel's
    background is the mean of bg rqe + scmos over the peaks whose ROI covers
    it (multi_fit.c:28-57, 143).  A peak's J and H follow the reference's
    text: jt of h, X, Y the derivatives of its own foreground, jt of bg = rqe
    (dao_fit.c:88-91), summed over its ROI with the full model's f.
    """
    def __init__(self, frame, peaks):
        self.fr = frame
        self.peaks = [(numpy.asarray(th, dtype=float), xi, yi) for th, xi, yi in peaks]

    def _cover(self, xi, yi, rows, cols):
        r = self.fr.r
        ry, rx = rows - yi, cols - xi
        inside = (ry >= 0) & (ry < r) & (rx >= 0) & (rx < r)
        c = 0.5 * r
        return inside & (((rx - c + 0.5) ** 2 + (ry - c + 0.5) ** 2) <= c * c)

    def fit(self, rows, cols, peaks=None, k=None, thk=None):
        fr = self.fr
        f = 0.0
        bgs = numpy.zeros(rows.shape)
        cnt = numpy.zeros(rows.shape)
        q, s = fr.rqe[rows, cols], fr.scmos[rows, cols]
        for j, (th, xi, yi) in enumerate(self.peaks if peaks is None else peaks):
            if j == k:
                th = thk
            m = self._cover(xi, yi, rows, cols)
            e = numpy.exp(-((cols - th[1] - fr.xoff) ** 2) * fr.width) * numpy.exp(-((rows - th[2] - fr.xoff) ** 2) * fr.width)
            f = f + numpy.where(m, th[0] * e * q, 0.0)
            bgs = bgs + numpy.where(m, th[3] * q + s, 0.0)
            cnt = cnt + m
        return f + bgs / cnt

    def deviance(self, k, thk=None):
        """Peak k's error: the deviance over its ROI (at its current origin)."""
        th, xi, yi = self.peaks[k]
        rows, cols = yi + self.fr.ry, xi + self.fr.rx
        f = self.fit(rows, cols, k=k, thk=th if thk is None else thk)
        x = self.fr.x[rows, cols]
        return numpy.sum(2.0 * ((f - x) - x * numpy.log(f / x)))

    def step(self, k, lam):
        th, xi, yi = self.peaks[k]
        fr = self.fr
        rows, cols = yi + fr.ry, xi + fr.rx
        x = fr.x[rows, cols]
        f = numpy.real(self.fit(rows, cols))
        q = fr.rqe[rows, cols]
        jt = numpy.zeros((4, x.size))
        for p in range(3):
            t = th.astype(complex)
            t[p] += 1j * CS
            e = numpy.exp(-((cols - t[1] - fr.xoff) ** 2) * fr.width) * numpy.exp(-((rows - t[2] - fr.xoff) ** 2) * fr.width)
            jt[p] = numpy.imag(t[0] * e * q) / CS
        jt[3] = q
        J = jt @ (2.0 * (1.0 - x / f))
        H = (jt * (2.0 * x / (f * f))) @ jt.T
        A = H.copy()
        A[numpy.diag_indices_from(A)] *= (1.0 + lam)
        return th - numpy.linalg.solve(A, J)

    def update(self, k, thk):
        th, xi, yi = self.peaks[k]
        self.peaks[k] = (numpy.asarray(thk, dtype=float), Frame.moved_origin(thk[1], xi), Frame.moved_origin(thk[2], yi))
