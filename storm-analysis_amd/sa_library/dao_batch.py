"""
Batched fitting: many frames resident on the GPU, fitted together.

The reference fits one frame per MultiFitter (dao_fit_c.py:257-549) and its
callers loop over frames.  BatchFitter holds a whole set of frames (sizes may
differ) on the device and runs MultiFitter.doFit (dao_fit_c.py:332-356) for
all of them in one executor: the per-frame results are those of the
single-frame fitter, because frames share no state.  Arrays may be numpy
(host) or torch CUDA tensors (already resident in HBM).
"""
import ctypes

import numpy
from numpy.ctypeslib import ndpointer

from . import loadclib
from .dao_fit_c import MultiFitterException, formatPeaksGaussianPSF, raiseOnLibraryError

SHAPES = {"2dfixed": 0, "2d": 1, "3d": 2, "Z": 3}
ERROR_MODELS = {"MLE": 0, "ALS": 1, "LS": 2, "DWLS": 3, "FWLS": 4}
PRECISIONS = {"fp64": 0, "mixed": 1}
COUNTER_NAMES = ("n_dposv", "n_iterations", "n_lost", "n_margin", "n_neg_fi", "n_neg_height",
                 "n_non_converged", "n_non_decr")
PEAK_COLUMNS = ("x", "y", "z", "height", "background", "xsigma", "ysigma", "error", "iterations")


class daoBatchModel(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int),
                ("error_model", ctypes.c_int),
                ("roi_size", ctypes.c_int),
                ("margin", ctypes.c_int),
                ("tolerance", ctypes.c_double),
                ("width_min", ctypes.c_double),
                ("width_max", ctypes.c_double),
                ("wx", ctypes.c_double * 5),
                ("wy", ctypes.c_double * 5),
                ("z_min", ctypes.c_double),
                ("z_max", ctypes.c_double),
                ("precision", ctypes.c_int)]


class FrameFitError(MultiFitterException):
    """
    Some frames of a batch failed in doFit (a scheduler wait of the fitter
    exceeded its bound).  `frames` lists them; they were emptied.  The other
    frames were fitted and their results can be read as usual.
    """
    def __init__(self, message, frames):
        super(FrameFitError, self).__init__(message)
        self.frames = list(frames)


_FRAME_FAILED = -5
_BATCH_ENTRY_POINTS = (
    "daoBatchCreate", "daoBatchSetCalibration", "daoBatchNewImages", "daoBatchNewBackgrounds", "daoBatchNewFrames",
    "daoBatchAnscombe", "daoBatchNewPeaks", "daoBatchFit", "daoBatchGetPeaks", "daoBatchGetAllPeaks",
    "daoBatchGetAllPeaksDevice", "daoBatchGetFitImage", "daoBatchGetFitImages", "daoBatchRemoveErrorPeaks",
    "daoBatchLoadFrame", "daoBatchClearFrame", "daoBatchLoadFrames", "daoBatchClearFrames",
    "daoBatchRemoveRunningPeaks", "daoBatchSetPeakStatus", "daoBatchGetSignificanceDevice", "daoBatchFitActive")


def _batch_error(code, msg, func, args):
    if code == _FRAME_FAILED and func.__name__ in ("daoBatchFit", "daoBatchFitActive"):
        lib = loadBatchC()
        n = lib.daoBatchNFrames(args[0])
        flags = numpy.zeros(max(n, 1), dtype=numpy.int32)
        lib.daoBatchFailedFrames(args[0], flags)
        return FrameFitError("daoBatchFit: " + msg, numpy.flatnonzero(flags[:n]).tolist())
    return MultiFitterException("%s failed (code %d): %s" % (func.__name__, code, msg))


_lib = None


def loadBatchC():
    global _lib
    if _lib is None:
        lib = loadclib.loadCLibrary("dao_fit")
        vp = ctypes.c_void_p
        i32p = ndpointer(dtype=numpy.int32)
        f64p = ndpointer(dtype=numpy.float64)
        lib.daoBatchCreate.argtypes = [ctypes.POINTER(daoBatchModel), ctypes.c_int, i32p, i32p, i32p]
        lib.daoBatchCreate.restype = vp
        lib.daoBatchDestroy.argtypes = [vp]
        lib.daoBatchNFrames.argtypes = [vp]
        lib.daoBatchNFrames.restype = ctypes.c_int
        lib.daoBatchSetCalibration.argtypes = [vp, ctypes.c_int, vp, vp]
        lib.daoBatchNewImages.argtypes = [vp, vp, ctypes.c_int]
        lib.daoBatchNewBackgrounds.argtypes = [vp, vp, ctypes.c_int]
        lib.daoBatchNewFrames.argtypes = [vp, vp, vp, ctypes.c_int]
        lib.daoBatchAnscombe.argtypes = [vp]
        lib.daoBatchNewPeaks.argtypes = [vp, vp, i32p, ctypes.c_char_p, ctypes.c_int]
        lib.daoBatchFit.argtypes = [vp, ctypes.c_int]
        lib.daoBatchFit.restype = ctypes.c_int
        lib.daoBatchFailedFrames.argtypes = [vp, i32p]
        lib.daoBatchFailedFrames.restype = ctypes.c_int
        lib.daoBatchSetWaitLimit.argtypes = [vp, ctypes.c_longlong]
        lib.daoBatchSetTeamHandoff.argtypes = [vp, ctypes.c_int]
        lib.daoBatchNFit.argtypes = [vp, ctypes.c_int]
        lib.daoBatchNFit.restype = ctypes.c_int
        lib.daoBatchGetPeaks.argtypes = [vp, ctypes.c_int, f64p, i32p]
        lib.daoBatchGetAllPeaks.argtypes = [vp, f64p, i32p]
        lib.daoBatchGetAllPeaksDevice.argtypes = [vp, vp, vp]
        lib.daoBatchCounters.argtypes = [vp, ctypes.c_int, i32p]
        lib.daoBatchGetFitImage.argtypes = [vp, ctypes.c_int, f64p]
        lib.daoBatchGetFitImages.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int]
        lib.daoBatchRemoveErrorPeaks.argtypes = [vp]
        lib.daoBatchRemoveRunningPeaks.argtypes = [vp]
        lib.daoBatchSetPeakStatus.argtypes = [vp, vp, ctypes.c_int]
        lib.daoBatchGetSignificanceDevice.argtypes = [vp, vp]
        lib.daoBatchFitActive.argtypes = [vp, ctypes.c_int, vp]
        lib.daoBatchFitActive.restype = ctypes.c_int
        lib.daoBatchLoadFrame.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp,
                                          ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        lib.daoBatchLoadFrame.restype = ctypes.c_int
        lib.daoBatchClearFrame.argtypes = [vp, ctypes.c_int]
        lib.daoBatchLoadFrames.argtypes = [vp, ctypes.c_int, i32p, i32p, i32p, vp, vp, vp, i32p, ctypes.c_char_p,
                                           ctypes.c_int]
        lib.daoBatchLoadFrames.restype = ctypes.c_int
        lib.daoBatchClearFrames.argtypes = [vp, ctypes.c_int, i32p]
        lib.daoBatchSetTiming.argtypes = [vp, ctypes.c_int]
        lib.daoBatchSetTuning.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.daoBatchStats.argtypes = [vp, f64p]
        _lib = raiseOnLibraryError(lib, _BATCH_ENTRY_POINTS, _batch_error)
    return _lib


def _ptr(a, dtype):
    """(pointer, on_device, keep-alive) for a numpy array or a torch tensor."""
    if hasattr(a, "data_ptr") and hasattr(a, "is_cuda"):
        if a.is_cuda:
            import torch
            if not a.is_contiguous() or a.dtype != torch.float64:
                raise MultiFitterException("device arrays must be contiguous float64")
            # the library works on its own streams: finish the writes queued on
            # torch's current stream (not the whole device, which would also
            # wait for other fitters' work)
            torch.cuda.current_stream(a.device).synchronize()
            return ctypes.c_void_p(a.data_ptr()), 1, a
        a = a.numpy()
    a = numpy.ascontiguousarray(a, dtype=dtype)
    return a.ctypes.data_as(ctypes.c_void_p), 0, a


class BatchFitter(object):
    """
    frame_shapes: list of (height, width); peak_caps: max peaks per frame.
    model: "2dfixed", "2d", "3d" or "Z"; error_model: "MLE", "ALS", "LS",
    "DWLS", "FWLS".  sigma_range, wx_params, wy_params, min_z, max_z as for
    the MultiFitter classes.  precision: "fp64" (the reference's arithmetic)
    or "mixed" (FP32 deviance terms, Jacobian terms and Hessian; FP64 peak
    shape, fit-image accumulators, deviance and Jacobian sums, solve and
    accept/reject; MLE only, other error models run in FP64).
    """
    def __init__(self, frame_shapes, peak_caps, model="Z", error_model="MLE", roi_size=18, sigma_range=None,
                 wx_params=None, wy_params=None, min_z=None, max_z=None, margin=2, tolerance=1.0e-6,
                 precision="fp64"):
        self.clib = loadBatchC()
        self.shapes = [tuple(s) for s in frame_shapes]
        self.n_frames = len(self.shapes)
        m = daoBatchModel()
        m.shape = SHAPES[model]
        m.error_model = ERROR_MODELS[error_model]
        m.roi_size = roi_size
        m.margin = margin
        m.tolerance = tolerance
        m.precision = PRECISIONS[precision]
        if sigma_range is not None:
            m.width_max = 1.0 / (2.0 * sigma_range[0] * sigma_range[0])
            m.width_min = 1.0 / (2.0 * sigma_range[1] * sigma_range[1])
        if model == "Z":
            for i in range(5):
                m.wx[i] = wx_params[i]
                m.wy[i] = wy_params[i]
            m.z_min, m.z_max = min_z, max_z
        self.model = model
        self.error_model = error_model
        self.roi_size = roi_size
        self.precision = precision
        h = numpy.array([s[0] for s in self.shapes], dtype=numpy.int32)
        w = numpy.array([s[1] for s in self.shapes], dtype=numpy.int32)
        self.caps = numpy.asarray(peak_caps, dtype=numpy.int32)
        self.handle = self.clib.daoBatchCreate(ctypes.byref(m), self.n_frames, w, h, self.caps)
        if not self.handle:
            raise MultiFitterException("daoBatchCreate failed (no usable GPU, or bad sizes)")
        self.npix = int(numpy.sum(h.astype(numpy.int64) * w))

    def cleanup(self):
        if self.handle:
            self.clib.daoBatchDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass

    def setCalibration(self, frame, rqe=None, scmos_cal=None):
        keep = []
        args = []
        for a in (rqe, scmos_cal):
            if a is None:
                args.append(None)
            else:
                a = numpy.ascontiguousarray(a, dtype=numpy.float64)
                if a.shape != self.shapes[frame]:
                    raise MultiFitterException("calibration shape does not match frame %d" % frame)
                keep.append(a)
                args.append(a.ctypes.data_as(ctypes.c_void_p))
        self.clib.daoBatchSetCalibration(self.handle, frame, args[0], args[1])

    def _concat(self, images):
        if isinstance(images, (list, tuple)):
            for k, im in enumerate(images):
                if tuple(im.shape) != self.shapes[k]:
                    raise MultiFitterException("image %d has the wrong shape" % k)
            return numpy.concatenate([numpy.ascontiguousarray(im, dtype=numpy.float64).ravel()
                                      for im in images])
        n = images.numel() if hasattr(images, "numel") else images.size
        if n != self.npix:
            raise MultiFitterException("concatenated images have the wrong size")
        return images

    def newImages(self, images):
        p, dev, keep = _ptr(self._concat(images), numpy.float64)
        self.clib.daoBatchNewImages(self.handle, p, dev)
        if self.error_model == "ALS":
            self.clib.daoBatchAnscombe(self.handle)

    def newBackgrounds(self, backgrounds):
        p, dev, keep = _ptr(self._concat(backgrounds), numpy.float64)
        self.clib.daoBatchNewBackgrounds(self.handle, p, dev)

    def newFrames(self, images, backgrounds):
        """newImages and newBackgrounds in one pass over the pixels (same values)."""
        pi, di, ki = _ptr(self._concat(images), numpy.float64)
        pb, db, kb = _ptr(self._concat(backgrounds), numpy.float64)
        if di != db:
            raise MultiFitterException("images and backgrounds must both be host or both be device arrays")
        self.clib.daoBatchNewFrames(self.handle, pi, pb, di)
        if self.error_model == "ALS":
            self.clib.daoBatchAnscombe(self.handle)

    def newPeaks(self, peaks, peaks_type, counts=None):
        """
        peaks: a list with one peaks dict per frame (as for
        MultiFitter.newPeaks), or concatenated rows (numpy / torch) with
        `counts` giving the rows of each frame.
        """
        if isinstance(peaks, (list, tuple)):
            rows = [formatPeaksGaussianPSF(p, peaks_type) for p in peaks]
            counts = numpy.array([r.shape[0] for r in rows], dtype=numpy.int32)
            width = 7 if peaks_type in ("text", "hdf5") else 4
            rows = numpy.concatenate(rows) if len(rows) else numpy.zeros((0, width))
        else:
            rows = peaks
            if counts is None:
                raise MultiFitterException("counts are required with concatenated rows")
        counts = numpy.ascontiguousarray(counts, dtype=numpy.int32)
        if counts.shape != (self.n_frames,):
            raise MultiFitterException("counts must have one entry per frame (%d)" % self.n_frames)
        held = numpy.array([self.getNFit(f) for f in range(self.n_frames)], dtype=numpy.int64)
        if numpy.any(counts < 0) or numpy.any(held + counts > self.caps):
            raise MultiFitterException("more peaks than the frame's capacity")
        p, dev, keep = _ptr(rows, numpy.float64)
        self.clib.daoBatchNewPeaks(self.handle, p, counts, ctypes.c_char_p(peaks_type.encode()), dev)

    def loadFrame(self, frame, image, background, rows, peaks_type="hdf5", rqe=None, scmos_cal=None):
        """
        Serving: put one frame (any size up to the slot's creation size) with
        its background, start peaks (rows in the daoNewPeaks layout) and
        optional calibration into slot `frame`.  Arrays are numpy (host) or
        torch CUDA tensors (all of one kind).
        """
        h, w = image.shape[0], image.shape[1]
        width = 7 if peaks_type in ("text", "hdf5") else 4
        n = rows.shape[0] if rows is not None else 0
        if rows is not None and rows.shape[1] != width:
            raise MultiFitterException("peak rows must be n x %d for '%s'" % (width, peaks_type))
        keep = []
        ptrs = []
        on_device = None
        for a in (image, background, rqe, scmos_cal, rows if n else None):
            if a is None:
                ptrs.append(None)
                continue
            p, dev, k = _ptr(a if hasattr(a, "data_ptr") else numpy.ascontiguousarray(a, dtype=numpy.float64),
                             numpy.float64)
            if on_device is None:
                on_device = dev
            elif dev != on_device:
                raise MultiFitterException("loadFrame arrays must all be host or all device arrays")
            keep.append(k)
            ptrs.append(p)
        ok = self.clib.daoBatchLoadFrame(self.handle, frame, w, h, ptrs[0], ptrs[1], ptrs[2], ptrs[3], ptrs[4], n,
                                         ctypes.c_char_p(peaks_type.encode()), 1 if on_device else 0)
        if not ok:
            raise MultiFitterException("frame %dx%d with %d peaks does not fit slot %d" % (h, w, n, frame))
        self.shapes[frame] = (h, w)

    def clearFrame(self, frame):
        """Serving: empty slot `frame` (doFit skips it)."""
        self.clib.daoBatchClearFrame(self.handle, frame)

    def loadFrames(self, slots, images, backgrounds, rows, peaks_type="hdf5"):
        """
        Serving: loadFrame for several slots in one call (rqe = 1, scmos = 0).
        images, backgrounds, rows: sequences (one entry per slot) of numpy
        arrays or of torch CUDA tensors (all of one kind).
        """
        n = len(slots)
        if n == 0:
            return
        width = 7 if peaks_type in ("text", "hdf5") else 4
        keep = []
        ptr = {"img": (ctypes.c_void_p * n)(), "bg": (ctypes.c_void_p * n)(), "rows": (ctypes.c_void_p * n)()}
        on_device = None
        nr = numpy.zeros(n, dtype=numpy.int32)
        w = numpy.zeros(n, dtype=numpy.int32)
        h = numpy.zeros(n, dtype=numpy.int32)
        for k in range(n):
            h[k], w[k] = images[k].shape[0], images[k].shape[1]
            nr[k] = rows[k].shape[0] if rows[k] is not None else 0
            if nr[k] and rows[k].shape[1] != width:
                raise MultiFitterException("peak rows must be n x %d for '%s'" % (width, peaks_type))
            for key, a in (("img", images[k]), ("bg", backgrounds[k]), ("rows", rows[k] if nr[k] else None)):
                if a is None:
                    continue
                p, dev, kp = _ptr(a, numpy.float64)
                if on_device is None:
                    on_device = dev
                elif dev != on_device:
                    raise MultiFitterException("loadFrames arrays must all be host or all device arrays")
                keep.append(kp)
                ptr[key][k] = p.value
        ok = self.clib.daoBatchLoadFrames(self.handle, n, numpy.asarray(slots, dtype=numpy.int32), w, h,
                                          ptr["img"], ptr["bg"], ptr["rows"], nr,
                                          ctypes.c_char_p(peaks_type.encode()), 1 if on_device else 0)
        if not ok:
            raise MultiFitterException("a frame or its peaks do not fit its slot")
        for k, f in enumerate(slots):
            self.shapes[f] = (int(h[k]), int(w[k]))

    def clearFrames(self, slots):
        """Serving: empty several slots in one call."""
        if len(slots):
            self.clib.daoBatchClearFrames(self.handle, len(slots), numpy.asarray(slots, dtype=numpy.int32))

    def doFit(self, max_iterations=2000):
        """
        MultiFitter.doFit for every frame; returns the sweeps run.  Raises
        FrameFitError (naming the frames) if some frames failed; the other
        frames' results stand.
        """
        return self.clib.daoBatchFit(self.handle, max_iterations)

    def setWaitLimit(self, polls):
        """Polls after which a scheduler wait gives up and fails its frame (0: default)."""
        self.clib.daoBatchSetWaitLimit(self.handle, int(polls))

    def setTeamHandoff(self, ops=-1):
        """
        With teams, hand a frame's tail to one workgroup once an LM sweep
        would run at most `ops` peaks (0: never; -1: the default, 32).  The
        fits do not depend on it.
        """
        self.clib.daoBatchSetTeamHandoff(self.handle, int(ops))

    def doFitFrames(self, active, max_iterations=2000):
        """
        doFit limited to the frames whose active[f] is true (the others are
        left as they are); returns the sweeps run.
        """
        act = numpy.ascontiguousarray(numpy.asarray(active, dtype=bool).astype(numpy.int32))
        if act.shape != (self.n_frames,):
            raise MultiFitterException("active must have one entry per frame (%d)" % self.n_frames)
        return self.clib.daoBatchFitActive(self.handle, max_iterations, act.ctypes.data_as(ctypes.c_void_p))

    def removeErrorPeaks(self):
        self.clib.daoBatchRemoveErrorPeaks(self.handle)

    def removeRunningPeaks(self):
        """mFitRemoveRunningPeaks of every frame (each removal counts as n_non_converged)."""
        self.clib.daoBatchRemoveRunningPeaks(self.handle)

    def setPeakStatus(self, status):
        """
        mFitSetPeakStatus of every frame: one int32 status per live peak, frames
        in order (the rows of getAllPeaks), a numpy array or a torch CUDA tensor.
        """
        n = int(sum(self.getNFit(f) for f in range(self.n_frames)))
        if hasattr(status, "is_cuda") and status.is_cuda:
            import torch
            if status.dtype != torch.int32 or not status.is_contiguous() or status.numel() != n:
                raise MultiFitterException("status must be a contiguous int32 tensor of %d entries" % n)
            torch.cuda.current_stream(status.device).synchronize()
            self.clib.daoBatchSetPeakStatus(self.handle, ctypes.c_void_p(status.data_ptr()), 1)
            return
        st = numpy.ascontiguousarray(status, dtype=numpy.int32)
        if st.size != n:
            raise MultiFitterException("status must have %d entries" % n)
        self.clib.daoBatchSetPeakStatus(self.handle, st.ctypes.data_as(ctypes.c_void_p), 0)

    def getSignificanceDevice(self, device):
        """The peaks' significance (fg_sum / sqrt(bg_sum)), frames in order, as a torch tensor on `device`."""
        import torch
        n = int(sum(self.getNFit(f) for f in range(self.n_frames)))
        out = torch.empty(n, dtype=torch.float64, device=device)
        if n:
            torch.cuda.current_stream(out.device).synchronize()
            self.clib.daoBatchGetSignificanceDevice(self.handle, ctypes.c_void_p(out.data_ptr()))
        return out

    def getNFit(self, frame):
        return self.clib.daoBatchNFit(self.handle, frame)

    def getPeaks(self, frame):
        n = self.getNFit(frame)
        out = numpy.zeros((n, 9))
        status = numpy.zeros(n, dtype=numpy.int32)
        if n:
            self.clib.daoBatchGetPeaks(self.handle, frame, out, status)
        d = {c: out[:, k].copy() for k, c in enumerate(PEAK_COLUMNS)}
        d["iterations"] = d["iterations"].astype(numpy.int32)
        d["status"] = status
        return d

    def getAllPeaks(self, out=None, status=None):
        """
        (rows n x 9 in PEAK_COLUMNS order, status, per-frame counts), frames in
        order.  out / status: optional host arrays of at least n rows to fill
        (for example in pinned memory, which the device writes directly); the
        returned arrays are their first n rows.
        """
        counts = numpy.array([self.getNFit(f) for f in range(self.n_frames)], dtype=numpy.int64)
        n = int(counts.sum())
        if out is None or status is None:
            out = numpy.zeros((n, 9))
            status = numpy.zeros(n, dtype=numpy.int32)
        else:
            if out.shape[0] < n or out.shape[1:] != (9,) or out.dtype != numpy.float64 or status.shape[0] < n or \
                    status.dtype != numpy.int32 or not out.flags.c_contiguous or not status.flags.c_contiguous:
                raise MultiFitterException("getAllPeaks: out must be n x 9 float64 and status n int32, contiguous")
            out, status = out[:n], status[:n]
        if n:
            self.clib.daoBatchGetAllPeaks(self.handle, out, status)
        return out, status, counts

    def getCounters(self, frame):
        c = numpy.zeros(8, dtype=numpy.int32)
        self.clib.daoBatchCounters(self.handle, frame, c)
        return dict(zip(COUNTER_NAMES, c.tolist()))

    def getFitImage(self, frame):
        out = numpy.zeros(self.shapes[frame])
        self.clib.daoBatchGetFitImage(self.handle, frame, out)
        return out

    def getFitImages(self, residual=False, device=None):
        """
        Fit images (f / rqe) -- or residuals x - f -- of every frame,
        concatenated in frame order (flat).  device: None for a numpy array,
        else a torch device: the images are written there without a host copy.
        """
        n = int(sum(h * w for h, w in self.shapes))
        if device is None:
            out = numpy.zeros(n)
            self.clib.daoBatchGetFitImages(self.handle, out.ctypes.data_as(ctypes.c_void_p), int(residual), 0)
            return out
        import torch
        out = torch.empty(n, dtype=torch.float64, device=device)
        torch.cuda.current_stream(out.device).synchronize()
        self.clib.daoBatchGetFitImages(self.handle, ctypes.c_void_p(out.data_ptr()), int(residual), 1)
        return out

    def getAllPeaksDevice(self, device):
        """getAllPeaks into torch tensors on `device` (rows n x 9, status n, counts)."""
        import torch
        counts = numpy.array([self.getNFit(f) for f in range(self.n_frames)], dtype=numpy.int64)
        n = int(counts.sum())
        out = torch.empty((n, 9), dtype=torch.float64, device=device)
        status = torch.empty(n, dtype=torch.int32, device=device)
        if n:
            torch.cuda.current_stream(out.device).synchronize()
            self.clib.daoBatchGetAllPeaksDevice(self.handle, ctypes.c_void_p(out.data_ptr()),
                                                ctypes.c_void_p(status.data_ptr()))
        return out, status, counts

    def setTuning(self, waves=0, margin=-1, team=-1):
        """
        Waves per workgroup (0: the kernel's maximum), reserved-box margin
        (-1: keep) and team size: peaks per workgroup when several
        workgroups share a frame (0: one workgroup per frame; -1: keep).
        """
        self.clib.daoBatchSetTuning(self.handle, waves, margin, team)

    def setTiming(self, on=True):
        self.clib.daoBatchSetTiming(self.handle, 1 if on else 0)

    def stats(self):
        s = numpy.zeros(6)
        self.clib.daoBatchStats(self.handle, s)
        return dict(zip(("lm_ms", "lm_launches", "sweeps", "rounds", "escape_reruns", "max_attempts"),
                        s.tolist()))
