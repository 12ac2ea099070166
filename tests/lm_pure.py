"""
A second restatement of the reference's LM fitter, in plain Python, used only
to check the oracle (oracle/dao_oracle.c) on tiny frames.  It is written from
the reference's text independently of the oracle's code, keeps the
reference's own data flow (per-pixel t_fi / err_i cache with stale flags,
working peak copied in and out, sequential sums in ROI order) and supports
the MLE error model of the four DAO models.

Followed:
  storm_analysis/sa_library/multi_fit.c
    mFitInitializeROIIndexing :1031-1060, mFitAddPeak / mFitSubtractPeak
    :28-57 / :1894-1922, mFitCalcErr :111-195, mFitCheck :470-502,
    mFitIterateLM :1072-1349, mFitResetPeak :1788-1798, mFitSolve
    :1839-1883 (dposv on the upper triangle), mFitRecenterPeaks
    :1644-1690, mFitUpdate :1933-1946, mFitNewPeaks :1478-1499
  storm_analysis/sa_library/dao_fit.c
    daoCalcJH2DFixed :48-105, daoCalcJH2D :409-467, daoCalcJH3D :811-871,
    daoCalcJHZ :883-958, daoCalcPeakShape :968-998, daoCalcWidthsFromZ
    :1009-1035, daoCheck* :1047-1133, daoInitialize :1265-1270 (offsets),
    daoInitializeZ :1548-1571, daoNewPeaks "text"/"hdf5" :1691-1751,
    daoUpdate* :1777-1880
  storm_analysis/sa_library/dao_fit_c.py: doFit :332-356
"""
import math

RUNNING, CONVERGED, ERROR = 0, 1, 2
HEIGHT, XCENTER, XWIDTH, YCENTER, YWIDTH, BACKGROUND, ZCENTER = range(7)
LAMBDASTART, LAMBDAUP, LAMBDADOWN, LAMBDAMIN, LAMBDAMAX = 1.0, 4.0, 0.75, 1.0e-3, 1.0e20
JAC = {"2dfixed": 4, "2d": 5, "3d": 6, "Z": 5}


class Peak(object):
    def __init__(self):
        self.params = [0.0] * 7
        self.xi = 0
        self.yi = 0
        self.status = RUNNING
        self.lam = LAMBDASTART
        self.error = 0.0
        self.iterations = 0
        self.added = 0
        self.wx_term = 0.0
        self.wy_term = 0.0

    def copy(self):
        c = Peak()
        c.params = list(self.params)
        for k in ("xi", "yi", "status", "lam", "error", "iterations", "added", "wx_term", "wy_term"):
            setattr(c, k, getattr(self, k))
        return c


class PureFitter(object):
    def __init__(self, shape, roi, model, tol=1.0e-6, rqe=None, scmos=None, width_min=0.0, width_max=0.0,
                 wx=None, wy=None, z_min=0.0, z_max=0.0):
        self.H, self.W = shape
        self.roi = roi
        self.model = model
        self.tol = tol
        self.xoff = 0.5 * roi - 1.0
        self.zoff = 0.0
        self.wmin, self.wmax = width_min, width_max
        if model == "Z":
            self.wxz = [wx[0] * wx[0]] + list(wx[1:5])
            self.wyz = [wy[0] * wy[0]] + list(wy[1:5])
            self.zmin, self.zmax = z_min, z_max
        n = self.H * self.W
        self.rqe = [1.0] * n if rqe is None else [float(v) for v in rqe.ravel()]
        self.scmos = [0.0] * n if scmos is None else [float(v) for v in scmos.ravel()]
        # circular ROI, row-major
        self.rx, self.ry = [], []
        c = 0.5 * roi
        for i in range(roi):
            dy = i - c + 0.5
            for j in range(roi):
                dx = j - c + 0.5
                if dx * dx + dy * dy <= c * c:
                    self.rx.append(j)
                    self.ry.append(i)
        self.peaks = []
        self.counters = dict.fromkeys(("n_dposv", "n_iterations", "n_lost", "n_margin", "n_neg_fi",
                                       "n_neg_height", "n_non_decr"), 0)

    # -- image state ---------------------------------------------------------
    def new_image(self, image):
        n = self.H * self.W
        img = [float(v) for v in image.ravel()]
        self.x = [img[k] * self.rqe[k] + self.scmos[k] for k in range(n)]
        self.f = [0.0] * n
        self.bg = [0.0] * n
        self.cnt = [0] * n
        self.t_fi = [0.0] * n
        self.err_i = [0.0] * n
        self.stale = [1] * n
        self.peaks = []

    def pixels(self, p):
        base = p.yi * self.W + p.xi
        return [self.ry[j] * self.W + self.rx[j] + base for j in range(len(self.rx))]

    # -- model ---------------------------------------------------------------
    def widths_from_z(self, p):
        for cal, w, term in ((self.wxz, XWIDTH, "wx_term"), (self.wyz, YWIDTH, "wy_term")):
            z0 = (p.params[ZCENTER] - cal[1]) / cal[2]
            z1 = z0 * z0
            z2 = z1 * z0
            z3 = z2 * z0
            t = 1.0 + z1 + cal[3] * z2 + cal[4] * z3
            setattr(p, term, t * t)
            p.params[w] = 2.0 / (cal[0] * t)

    def shape(self, p):
        xo = -(p.params[XCENTER] - p.xi) - self.xoff
        yo = -(p.params[YCENTER] - p.yi) - self.xoff  # dao_fit.c:979 uses xoff
        xt = [i + xo for i in range(self.roi)]
        yt = [i + yo for i in range(self.roi)]
        ext = [math.exp(-t * t * p.params[XWIDTH]) for t in xt]
        eyt = [math.exp(-t * t * p.params[YWIDTH]) for t in yt]
        psf = [eyt[self.ry[j]] * ext[self.rx[j]] for j in range(len(self.rx))]
        return xt, ext, yt, eyt, psf

    def add(self, p, sign):
        _, _, _, _, psf = self.shape(p)
        p.added += sign
        for j, k in enumerate(self.pixels(p)):
            r = self.rqe[k]
            if sign > 0:
                self.f[k] += p.params[HEIGHT] * psf[j] * r
                self.cnt[k] += 1
                self.bg[k] += p.params[BACKGROUND] * r + self.scmos[k]
            else:
                self.f[k] -= p.params[HEIGHT] * psf[j] * r
                self.cnt[k] -= 1
                self.bg[k] -= p.params[BACKGROUND] * r + self.scmos[k]
            self.stale[k] = 1

    def error(self, p):
        """mFitCalcErr: 1 on fi <= 0 (error not updated), else 0."""
        if p.status != RUNNING:
            return 0
        err = 0.0
        for k in self.pixels(p):
            if self.stale[k]:
                fi = self.f[k] + self.bg[k] / self.cnt[k]
                if fi <= 0.0:
                    self.counters["n_neg_fi"] += 1
                    return 1
                self.t_fi[k] = fi
                xi = self.x[k]
                self.err_i[k] = 2.0 * ((fi - xi) - xi * math.log(fi / xi))
                self.stale[k] = 0
            err += self.err_i[k]
        p.error = err
        return 0

    def jh(self, p):
        n = JAC[self.model]
        jac = [0.0] * n
        hes = [[0.0] * n for _ in range(n)]
        xt, ext, yt, eyt, _ = self.shape(p)
        a1 = p.params[HEIGHT]
        if self.model == "Z":
            g = []
            for cal, term in ((self.wxz, p.wx_term), (self.wyz, p.wy_term)):
                z0 = (p.params[ZCENTER] - cal[1]) / cal[2]
                z1 = z0 * z0
                z2 = z1 * z0
                zt = 2.0 * z0 + 3.0 * cal[3] * z1 + 4.0 * cal[4] * z2
                g.append(-2.0 * zt / (cal[0] * term))
            gx, gy = g
        for j, k in enumerate(self.pixels(p)):
            fi = self.t_fi[k]
            r = self.rqe[k]
            xi = self.x[k]
            x_t, e_x = xt[self.rx[j]], ext[self.rx[j]]
            y_t, e_y = yt[self.ry[j]], eyt[self.ry[j]]
            e_t = e_x * e_y
            if self.model == "2dfixed":
                w = p.params[XWIDTH]
                jt = [r * e_t, r * 2.0 * a1 * w * x_t * e_t, r * 2.0 * a1 * w * y_t * e_t, r]
            elif self.model == "2d":
                w = p.params[XWIDTH]
                jt = [r * e_t, r * 2.0 * a1 * w * x_t * e_t, r * 2.0 * a1 * w * y_t * e_t,
                      r * (-a1 * x_t * x_t * e_t - a1 * y_t * y_t * e_t), r]
            elif self.model == "3d":
                a3, a5 = p.params[XWIDTH], p.params[YWIDTH]
                jt = [r * e_t, r * 2.0 * a1 * a3 * x_t * e_t, r * (-a1 * x_t * x_t * e_t),
                      r * 2.0 * a1 * a5 * y_t * e_t, r * (-a1 * y_t * y_t * e_t), r]
            else:
                a3, a5 = p.params[XWIDTH], p.params[YWIDTH]
                jt = [r * e_t, r * 2.0 * a1 * a3 * x_t * e_t, r * 2.0 * a1 * a5 * y_t * e_t,
                      r * (-a1 * x_t * x_t * gx * e_t - a1 * y_t * y_t * gy * e_t), r]
            t1 = 2.0 * (1.0 - xi / fi)
            t2 = 2.0 * xi / (fi * fi)
            for a in range(n):
                jac[a] += t1 * jt[a]
                for b in range(a, n):
                    hes[a][b] += t2 * jt[a] * jt[b]
        return jac, hes

    def margin_check(self, p):
        """mFitCheck."""
        if p.xi < 0 or p.xi >= self.W - self.roi or p.yi < 0 or p.yi >= self.H - self.roi:
            self.counters["n_margin"] += 1
            return 1
        if p.params[HEIGHT] < 0.0:
            self.counters["n_neg_height"] += 1
            return 1
        return 0

    def check(self, p):
        """daoCheck2DFixed / 2D / 3D / Z: 2D clamps XWIDTH, 3D both widths."""
        if self.margin_check(p):
            return 1
        clamp = {"2d": (XWIDTH,), "3d": (XWIDTH, YWIDTH)}.get(self.model, ())
        for q in clamp:
            if p.params[q] < self.wmin:
                p.params[q] = self.wmin
            elif p.params[q] > self.wmax:
                p.params[q] = self.wmax
        return 0

    @staticmethod
    def follow(p):
        dx = p.params[XCENTER] - p.xi
        if dx < -0.5 or dx > 1.5:
            p.xi = int(math.floor(p.params[XCENTER]))
        dy = p.params[YCENTER] - p.yi
        if dy < -0.5 or dy > 1.5:
            p.yi = int(math.floor(p.params[YCENTER]))

    def update(self, p, d):
        order = {"2dfixed": (HEIGHT, XCENTER, YCENTER, BACKGROUND),
                 "2d": (HEIGHT, XCENTER, YCENTER, XWIDTH, BACKGROUND),
                 "3d": (HEIGHT, XCENTER, XWIDTH, YCENTER, YWIDTH, BACKGROUND),
                 "Z": (HEIGHT, XCENTER, YCENTER, ZCENTER, BACKGROUND)}[self.model]
        for q, v in zip(order, d):
            p.params[q] -= v
        if self.model == "2d":
            p.params[YWIDTH] = p.params[XWIDTH]
        if self.model == "Z":
            if p.params[ZCENTER] < self.zmin:
                p.params[ZCENTER] = self.zmin
            if p.params[ZCENTER] > self.zmax:
                p.params[ZCENTER] = self.zmax
        self.follow(p)
        if self.model == "Z":
            self.widths_from_z(p)

    @staticmethod
    def solve(hes, lam, jac):
        """Cholesky of the symmetric matrix whose upper triangle is hes with the diagonal x (1 + lam)."""
        n = len(jac)
        a = [[0.0] * n for _ in range(n)]
        for i in range(n):
            for j in range(n):
                v = hes[min(i, j)][max(i, j)]
                a[i][j] = (1.0 + lam) * v if i == j else v
        low = [[0.0] * n for _ in range(n)]
        for j in range(n):
            s = a[j][j] - sum(low[j][k] * low[j][k] for k in range(j))
            if not s > 0.0:
                return None
            low[j][j] = math.sqrt(s)
            for i in range(j + 1, n):
                low[i][j] = (a[i][j] - sum(low[i][k] * low[j][k] for k in range(j))) / low[j][j]
        y = [0.0] * n
        for i in range(n):
            y[i] = (jac[i] - sum(low[i][k] * y[k] for k in range(i))) / low[i][i]
        x = [0.0] * n
        for i in reversed(range(n)):
            x[i] = (y[i] - sum(low[k][i] * x[k] for k in range(i + 1, n))) / low[i][i]
        if any(math.isnan(v) for v in x):
            return None
        return x

    # -- peaks ---------------------------------------------------------------
    def new_peaks_text(self, rows):
        start = len(self.peaks)
        for r in rows:
            p = Peak()
            p.params[XCENTER] = r[0] - self.xoff
            p.params[YCENTER] = r[1] - self.xoff
            p.params[ZCENTER] = r[2] - self.zoff
            p.params[BACKGROUND] = r[3]
            p.params[HEIGHT] = r[4]
            p.params[XWIDTH] = 1.0 / (2.0 * r[5] * r[5])
            p.params[YWIDTH] = 1.0 / (2.0 * r[6] * r[6])
            if self.model == "Z":
                self.widths_from_z(p)
            p.xi = int(math.floor(p.params[XCENTER]))
            p.yi = int(math.floor(p.params[YCENTER]))
            if self.check(p):
                p.status = ERROR
            else:
                self.add(p, +1)
            self.peaks.append(p)
        for p in self.peaks[start:]:
            w = p.copy()
            self.error(w)
            p.error = w.error

    # -- fitting -------------------------------------------------------------
    def iterate(self):
        for i, stored in enumerate(self.peaks):
            if stored.status != RUNNING:
                continue
            w = stored.copy()
            self.error(w)
            start = w.error
            jac, hes = self.jh(w)
            self.add(w, -1)
            while True:
                if w.lam > LAMBDAMAX:
                    self.counters["n_lost"] += 1
                    w.status = ERROR
                    break
                w.iterations += 1
                self.counters["n_iterations"] += 1
                w.status = RUNNING
                d = self.solve(hes, w.lam, jac)
                if d is None:
                    self.counters["n_dposv"] += 1
                    w.status = ERROR
                    w.lam *= LAMBDAUP
                    continue
                self.update(w, d)
                if self.check(w):
                    w = self.reset(w, stored)
                    w.status = ERROR
                    continue
                self.add(w, +1)
                if self.error(w):
                    self.add(w, -1)
                    w = self.reset(w, stored)
                    w.status = ERROR
                    continue
                if w.error > start:
                    if (w.error - start) / start < self.tol:
                        w.status = CONVERGED
                        break
                    self.counters["n_non_decr"] += 1
                    self.add(w, -1)
                    w = self.reset(w, stored)
                else:
                    if (start - w.error) / start < self.tol:
                        w.status = CONVERGED
                        break
                    if w.lam > LAMBDAMIN:
                        w.lam *= LAMBDADOWN
                    break
            self.peaks[i] = w
        self.recenter()

    @staticmethod
    def reset(w, stored):
        r = stored.copy()
        r.added = w.added
        r.lam = w.lam * LAMBDAUP
        return r

    def recenter(self):
        for i, p in enumerate(self.peaks):
            if p.status != RUNNING:
                continue
            dx = p.params[XCENTER] - p.xi
            dy = p.params[YCENTER] - p.yi
            if dx < -0.25 or dx > 1.25 or dy < -0.25 or dy > 1.25:
                w = p.copy()
                self.add(w, -1)
                w.xi = int(math.floor(p.params[XCENTER]))
                w.yi = int(math.floor(p.params[YCENTER]))
                if self.margin_check(w):
                    self.counters["n_lost"] += 1
                    w.status = ERROR
                else:
                    self.add(w, +1)
                self.peaks[i] = w

    def unconverged(self):
        return sum(1 for p in self.peaks if p.status == RUNNING)

    def properties(self):
        def sigma(w):  # sqrt(1/(2w)) as C computes it: NaN for w < 0
            v = 1.0 / (2.0 * w) if w != 0.0 else math.inf
            return math.sqrt(v) if v >= 0.0 else math.nan
        return {"x": [p.params[XCENTER] + self.xoff for p in self.peaks],
                "y": [p.params[YCENTER] + self.xoff for p in self.peaks],
                "z": [p.params[ZCENTER] + self.zoff for p in self.peaks],
                "height": [p.params[HEIGHT] for p in self.peaks],
                "background": [p.params[BACKGROUND] for p in self.peaks],
                "xsigma": [sigma(p.params[XWIDTH]) for p in self.peaks],
                "ysigma": [sigma(p.params[YWIDTH]) for p in self.peaks],
                "error": [p.error for p in self.peaks],
                "status": [p.status for p in self.peaks],
                "iterations": [p.iterations for p in self.peaks]}
