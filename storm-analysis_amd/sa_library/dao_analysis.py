#!/usr/bin/env python
"""
3D-DAOSTORM analysis of a batch of frames with every image and every peak
list kept in GPU memory: the find / fit / subtract loop of
PeakFinderFitter.analyzeImage (storm_analysis/sa_library/fitting.py:866-902)
with PeakFinderGaussian (:361-501) and PeakFitter.fitPeaks (:701-790), set
up as daostorm_3d/find_peaks.py (initFindAndFit) sets it up, for all frames
of a batch at once.  Per step of the loop:

    background   matched filter of image - fit (hipFFT), per frame memoised
                 as convolveMemo (matched_filter.c:118-146) does it
    search image finderPrepare / finderSNR kernels (fitting.py:452-494)
    new peaks    one local-maxima search over the batch (frames as
                 independent planes, ia_utilities.c:105-214)
    fit          runningIfHasNeighbors, newPeaks, doFit on the frames that
                 found peaks, removeErrorPeaks, markDimmerPeaks /
                 markLowSignificancePeaks, doFit again if needed
    fit image    read on the device for the next step

Only per-frame scalars (peak counts, done flags) reach the host between
steps.  A frame leaves the loop when the reference's would (threshold
schedule and no new peaks, fitting.py:274-302), and later steps do not
touch it.

    params = Parameters.fromFile("analysis.xml")
    ff = BatchFinderFitter(params, (height, width), n_frames)
    peaks = ff.analyze(frames)          # frames: n x height x width e-, GPU
    movie_peaks = analyzeMovie("movie.dax", "analysis.xml")
"""
import ctypes
import os
import xml.etree.ElementTree as ElementTree

import numpy

from . import dao_batch
from . import dao_filter
from . import dao_finder
from . import dao_ingest
from . import dao_mfilter
from . import dao_peakfinder
from .dao_fit_c import MultiFitterException

# PeakFinderFitter.unconverged_dist (fitting.py:844-847)
UNCONVERGED_DIST = 5.0
MEMO_MAX_DIFF = 1.0e-3


class Parameters(object):
    """
    The analysis settings of an XML file in the reference's format
    (sa_library/parameters.py): <name type="int|float|string|...">value</name>.
    """
    DEFAULTS = {"fit_error_model": "MLE", "z_value": 0.0, "no_fitting": 0, "min_z": -0.5, "max_z": 0.5,
                "start_frame": -1, "max_frame": -1}

    def __init__(self, values=None):
        self.values = dict(values or {})

    @classmethod
    def fromFile(cls, filename):
        root = ElementTree.parse(filename).getroot()
        values = {}
        for node in root:
            if not isinstance(node.tag, str):
                continue
            kind = node.attrib.get("type", "string")
            text = (node.text or "").strip()
            if kind == "int":
                values[node.tag] = int(text)
            elif kind == "float":
                values[node.tag] = float(text)
            elif kind == "int-array":
                values[node.tag] = numpy.array([int(v) for v in text.split(",") if v.strip()])
            elif kind == "float-array":
                values[node.tag] = numpy.array([float(v) for v in text.split(",") if v.strip()])
            else:
                values[node.tag] = text
        return cls(values)

    def hasAttr(self, name):
        return name in self.values

    def getAttr(self, name, default=None):
        if name in self.values:
            return self.values[name]
        if name in self.DEFAULTS:
            return self.DEFAULTS[name]
        if default is not None:
            return default
        raise MultiFitterException("parameter '%s' is not set" % name)

    def getWidthParams(self):
        """wx, wy versus z for the Z model (parameters.py:557-574, for_mu_Zfit)."""
        out = []
        for ax in ("wx", "wy"):
            p = numpy.array([float(self.values.get(ax + s, 0.0)) for s in ("_wo", "_c", "_d", "A", "B", "C", "D")])
            p[0] = p[0] / self.getAttr("pixel_size")
            p[1] = p[1] * 0.001
            p[2] = p[2] * 0.001
            out.append(p)
        return out


def gaussianPSF(shape, sigma):
    """
    gaussianPSF (fitting.py:38-45): a unit-sum Gaussian at (shape[0] / 2,
    shape[1] / 2), drawn as drawGaussiansXY draws it (draw_gaussians.c:24-80,
    one sample per pixel) -- only over the pixels within int(5 sigma) of the
    centre's integer part on each axis.
    """
    px, py = 0.5 * shape[0], 0.5 * shape[1]
    aw = max(int(5.0 * sigma), 1)
    r0, r1 = max(int(px) - aw, 0), min(int(px) + aw + 1, shape[0])
    c0, c1 = max(int(py) - aw, 0), min(int(py) + aw + 1, shape[1])
    s = 1.0 / (2.0 * sigma * sigma)
    rows = numpy.arange(r0, r1, dtype=numpy.float64)
    cols = numpy.arange(c0, c1, dtype=numpy.float64)
    psf = numpy.zeros(shape)
    psf[r0:r1, c0:c1] = numpy.exp(-1.0 * ((rows[:, None] - px) ** 2 * s + (cols[None, :] - py) ** 2 * s))
    return psf / numpy.sum(psf)


def peakMask(shape, parameters, margin):
    """peakMask (fitting.py:134-178), the square AOI kept with its last-row / last-column quirk."""
    mask = numpy.ones(shape)
    if parameters.hasAttr("x_center"):
        rr = parameters.getAttr("aoi_radius") ** 2
        xc = parameters.getAttr("x_center") + margin
        yc = parameters.getAttr("y_center") + margin
        xr = numpy.arange(mask.shape[0]) - xc
        yr = numpy.arange(mask.shape[0]) - yc
        xv, yv = numpy.meshgrid(xr, yr)
        mask[((xv * xv + yv * yv) > rr)] = 0
        return mask
    if parameters.hasAttr("x_start"):
        mask[:, 0:parameters.getAttr("x_start") + margin] = 0.0
    if parameters.hasAttr("x_stop"):
        mask[:, parameters.getAttr("x_stop") + margin:-1] = 0.0
    if parameters.hasAttr("y_start"):
        mask[0:parameters.getAttr("y_start") + margin, :] = 0.0
    if parameters.hasAttr("y_stop"):
        mask[parameters.getAttr("y_stop") + margin:-1, :] = 0.0
    return mask


def _offsets(counts, device):
    import torch
    off = numpy.zeros(len(counts) + 1, dtype=numpy.int32)
    off[1:] = numpy.cumsum(counts)
    return torch.from_numpy(off).to(device)


class _Memo(object):
    """Per-frame convolveMemo: recompute a frame's convolution only when its input moved by more than max_diff."""
    def __init__(self, mfilter, max_diff=MEMO_MAX_DIFF):
        self.mfilter = mfilter
        self.max_diff = max_diff
        self.old_in = None
        self.old_out = None

    def convolve(self, images):
        import torch
        new = self.mfilter.convolveBatch(images)
        if self.old_in is None:
            self.old_in, self.old_out = images.clone(), new
            return new
        changed = (images - self.old_in).abs().amax(dim=(1, 2)) > self.max_diff
        sel = changed[:, None, None]
        self.old_in = torch.where(sel, images, self.old_in)
        self.old_out = torch.where(sel, new, self.old_out)
        return self.old_out


class BatchFinderFitter(object):
    """
    PeakFinderGaussian + PeakFitter + PeakFinderFitter for n frames of one
    size, on the GPU.  shape: the frame size before padding.
    """
    def __init__(self, parameters, shape, n_frames, peak_cap=4096, rqe=None, camera_variance=None,
                 device="cuda"):
        import torch
        p = parameters
        self.parameters = p
        self.device = torch.device(device)
        self.n = int(n_frames)
        self.model = p.getAttr("model")
        self.sigma = float(p.getAttr("sigma"))
        self.threshold = float(p.getAttr("threshold"))
        self.iterations = int(p.getAttr("iterations"))
        self.radius = float(p.getAttr("find_max_radius"))
        self.z_value = float(p.getAttr("z_value"))
        self.no_fitting = bool(p.getAttr("no_fitting"))
        # ROI and margin (fitting.py:375-386)
        if p.hasAttr("roi_size") and p.getAttr("roi_size") != -1:
            self.roi_size = int(p.getAttr("roi_size"))
        else:
            self.roi_size = int(8.0 * self.sigma)
            if self.model != "2dfixed":
                self.roi_size = int(1.5 * self.roi_size)
        self.margin = int(self.roi_size / 2 + 2)
        self.neighborhood = self.sigma * UNCONVERGED_DIST
        self.shape = (shape[0] + 2 * self.margin, shape[1] + 2 * self.margin)
        h, w = self.shape

        self.peak_mask = torch.from_numpy(peakMask(self.shape, p, self.margin)).to(self.device)
        self.bg_memo = _Memo(dao_mfilter.MatchedFilter(gaussianPSF(self.shape, p.getAttr("background_sigma"))))
        self.fg_memo = self.var_memo = None
        if p.hasAttr("foreground_sigma") and p.getAttr("foreground_sigma") > 0.0:
            fg = gaussianPSF(self.shape, p.getAttr("foreground_sigma"))
            self.fg_memo = _Memo(dao_mfilter.MatchedFilter(fg))
            self.var_memo = _Memo(dao_mfilter.MatchedFilter(fg * fg))
        self.camera_variance = None
        if camera_variance is not None:
            self.camera_variance = dao_ingest.padArrays(numpy.asarray(camera_variance, dtype=numpy.float64),
                                                        self.margin, device=self.device)[0]

        # the fitter (find_peaks.py:15-100)
        kw = {"roi_size": self.roi_size, "error_model": p.getAttr("fit_error_model")}
        if self.model in ("2d", "3d"):
            kw["sigma_range"] = (p.getAttr("sigma_range") if p.hasAttr("sigma_range")
                                 else [0.5 * self.sigma, 5.0 * self.sigma])
        elif self.model == "Z":
            kw["wx_params"], kw["wy_params"] = p.getWidthParams()
            kw["min_z"], kw["max_z"] = p.getAttr("min_z"), p.getAttr("max_z")
        self.fitter = dao_batch.BatchFitter([self.shape] * self.n, [peak_cap] * self.n, model=self.model, **kw)
        if rqe is not None or camera_variance is not None:
            rq = None if rqe is None else dao_ingest.padArrays(numpy.asarray(rqe, dtype=numpy.float64),
                                                               self.margin, device="cpu")[0].numpy()
            cv = None if camera_variance is None else self.camera_variance.cpu().numpy()
            for f in range(self.n):
                self.fitter.setCalibration(f, rqe=rq, scmos_cal=cv)
        self.finder_lib = dao_finder.loadFinderC()

    def cleanup(self):
        self.fitter.cleanup()
        for m in (self.bg_memo, self.fg_memo, self.var_memo):
            if m is not None:
                m.mfilter.cleanup()

    # -- pieces of one step --

    def _find(self, images, background, fit, taken):
        """PeakFinderGaussian.peakFinder for every frame: (frame, y, x) of the new maxima, frame order."""
        import torch
        bg_var, fg = dao_peakfinder.finderPrepare(images, background, fit, self.camera_variance)
        if self.var_memo is not None:
            bg_var = self.var_memo.convolve(bg_var)
            fg = self.fg_memo.convolve(fg)
        snr = dao_peakfinder.finderSNR(fg.contiguous(), bg_var.contiguous(), self.peak_mask)
        n, h, w = snr.shape
        cap = n * h * w + 1
        out = torch.empty((4, cap), dtype=torch.float64, device=self.device)
        zv = numpy.arange(n, dtype=numpy.float64)   # the plane index tells the frame
        found = ctypes.c_int(0)
        torch.cuda.current_stream(self.device).synchronize()
        rc = self.finder_lib.daoFindLocalMaxima(
            ctypes.c_void_p(snr.data_ptr()), ctypes.c_void_p(taken.data_ptr()), n, h, w, self.margin + 1, 0,
            self.radius, self.threshold, zv.ctypes.data_as(ctypes.c_void_p), cap,
            ctypes.c_void_p(out[0].data_ptr()), ctypes.c_void_p(out[1].data_ptr()),
            ctypes.c_void_p(out[2].data_ptr()), ctypes.c_void_p(out[3].data_ptr()), ctypes.byref(found))
        if rc != 0:
            raise MultiFitterException("local maxima search failed (%d)" % rc)
        k = found.value
        return out[0, :k].long(), out[1, :k], out[2, :k]

    def _peaks(self):
        rows, status, counts = self.fitter.getAllPeaksDevice(self.device)
        return rows, status, counts

    def analyze(self, images):
        """
        images: n x height x width photo-electron frames (torch float64 on the
        GPU, not padded; or already padded to the analysis size).  Returns one
        dict of numpy arrays per frame (x, y in unpadded pixels).
        """
        import torch
        if tuple(images.shape[1:]) != self.shape:
            images = dao_ingest.padArrays(images, self.margin, device=self.device)
        images = images.contiguous()
        n = self.n
        fb = self.fitter
        fb.newImages(images)
        fit = torch.zeros_like(images)
        taken = torch.zeros(images.shape, dtype=torch.int32, device=self.device)
        done = numpy.zeros(n, dtype=bool)
        # newImage's threshold schedule (fitting.py:317-323)
        start = self.threshold + (4.0 if self.iterations > 4 else float(self.iterations))
        cur = numpy.full(n, start)
        frame_of = lambda counts: torch.repeat_interleave(  # noqa: E731
            torch.arange(n, device=self.device), torch.as_tensor(counts, device=self.device))
        for _ in range(self.iterations):
            active = ~done
            if not active.any():
                break
            background = self.bg_memo.convolve((images - fit).contiguous())
            f_idx, y, x = self._find(images, background, fit, taken)
            # only the frames still in the loop take their new peaks
            act_dev = torch.from_numpy(active).to(self.device)
            keep = act_dev[f_idx]
            f_idx, y, x = f_idx[keep], y[keep], x[keep]
            new_counts = torch.bincount(f_idx, minlength=n).cpu().numpy()
            # findPeaks (fitting.py:274-302)
            finished = numpy.zeros(n, dtype=bool)
            for f in numpy.flatnonzero(active):
                if cur[f] > self.threshold:
                    cur[f] -= 1.0
                else:
                    finished[f] = new_counts[f] == 0
            fb.newBackgrounds(background)
            part = active & (new_counts > 0)
            if part.any():
                self._fit_new(part, new_counts, x, y)
                fit = fb.getFitImages(device=self.device).view(images.shape)
            done |= finished
        # PeakFinderFitter.analyzeImage's end (fitting.py:898-902)
        if not self.no_fitting:
            fb.removeRunningPeaks()
        rows, status, counts = fb.getAllPeaks()
        out, at = [], 0
        for f in range(n):
            r = rows[at:at + counts[f]]
            at += counts[f]
            d = {c: r[:, k].copy() for k, c in enumerate(dao_batch.PEAK_COLUMNS)}
            d["x"] -= float(self.margin)
            d["y"] -= float(self.margin)
            out.append(d)
        return out

    def _fit_new(self, part, new_counts, x, y):
        """PeakFitter.fitPeaks (fitting.py:701-790) for the frames in `part`."""
        import torch
        fb = self.fitter
        n = self.n
        part_dev = torch.from_numpy(part).to(self.device)
        rows, status, counts = self._peaks()
        if counts.sum() > 0:
            c_off = _offsets(counts, self.device)
            n_off = _offsets(new_counts, self.device)
            st = status.clone()
            dao_filter.runningIfHasNeighborsBatch(st, rows[:, 0].contiguous(), rows[:, 1].contiguous(), c_off,
                                                  x.contiguous(), y.contiguous(), n_off, self.neighborhood)
            fb.setPeakStatus(st)
        new_rows = torch.stack([x, y, torch.full_like(x, self.z_value), torch.full_like(x, self.sigma)], dim=1)
        fb.newPeaks(new_rows.contiguous(), "finder", counts=new_counts.astype(numpy.int32))
        if not self.no_fitting:
            fb.doFitFrames(part)
            fb.removeErrorPeaks()
        # proximity and significance (fitting.py:748-779), for these frames only
        rows, status, counts = self._peaks()
        if counts.sum() > 0:
            off = _offsets(counts, self.device)
            sig = fb.getSignificanceDevice(self.device)
            st = status.clone()
            px, py, ph = rows[:, 0].contiguous(), rows[:, 1].contiguous(), rows[:, 3].contiguous()
            n_prox = dao_filter.markDimmerPeaksBatch(px, py, ph, st, off, self.sigma, self.neighborhood)
            n_sig = dao_filter.markLowSignificancePeaksBatch(px, py, sig, st, off, self.threshold,
                                                             self.neighborhood)
            owner = torch.repeat_interleave(torch.arange(n, device=self.device),
                                            torch.as_tensor(counts, device=self.device))
            st = torch.where(part_dev[owner], st, status).contiguous()
            removed = ((n_prox + n_sig).cpu().numpy() > 0) & part
            if removed.any():
                fb.setPeakStatus(st)
                fb.removeErrorPeaks()
        # more fitting where peaks are still running (fitting.py:781-785)
        if not self.no_fitting:
            rows, status, counts = self._peaks()
            if counts.sum() > 0:
                owner = torch.repeat_interleave(torch.arange(n, device=self.device),
                                                torch.as_tensor(counts, device=self.device))
                running = torch.bincount(owner[status == 0], minlength=n).cpu().numpy()
                more = part & (running > 0)
                if more.any():
                    fb.doFitFrames(more)
                    fb.removeErrorPeaks()


def readInf(dax_name):
    """(width, height, frames, big_endian) from a .dax movie's .inf file (datareader.py:190-232)."""
    inf = os.path.splitext(dax_name)[0] + ".inf"
    width = height = frames = None
    big_endian = False
    with open(inf) as fp:
        for line in fp:
            line = line.strip()
            if line.startswith("frame dimensions"):
                width, height = [int(v) for v in line.split("=")[1].split("x")]
            elif line.startswith("number of frames"):
                frames = int(line.split("=")[1])
            elif "big endian" in line:
                big_endian = True
    return width, height, frames, big_endian


def analyzeMovie(movie_name, settings_name, device="cuda"):
    """
    mufit_analysis.analyze (daostorm_3d/mufit_analysis.py:14-38) on the GPU
    for a .dax movie with a standard camera, all frames in one batch; returns
    the per-frame peak dicts (the localizations the reference would write).
    """
    params = Parameters.fromFile(settings_name)
    width, height, n_frames, big_endian = readInf(movie_name)
    first = max(params.getAttr("start_frame"), 0)
    last = n_frames if params.getAttr("max_frame") <= 0 else min(params.getAttr("max_frame"), n_frames)
    raw = dao_ingest.readDaxFrames(movie_name, width, height, list(range(first, last)), big_endian=False)
    reader = dao_ingest.FrameReaderStd(params.getAttr("camera_gain"), params.getAttr("camera_offset"), device=device)
    images = reader.loadFrames(raw, pad=0, big_endian=big_endian)
    ff = BatchFinderFitter(params, (height, width), last - first, device=device)
    try:
        return ff.analyze(images)
    finally:
        ff.cleanup()
