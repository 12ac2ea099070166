"""
Batched multiplane fitting: many frame-sets (one frame of every channel)
fitted together on the GPU, one wavefront per frame-set, with the arithmetic
of mpIterateLM (storm_analysis/multi_plane/mp_fit.c:180-487) and the 2D
Gaussian channels of mp_fit_dao.c.  No counterpart in the reference, which
fits one frame-set at a time (MPFitDao2D, mp_fit_dao_c.py:65-94): each set
here gets the results that MPFitDao2D gives it alone.

    bf = MPBatchFitter(n_sets, n_channels, (height, width), roi_size=12, sigma_range=[1.0, 3.0])
    for c in range(n_channels):
        bf.initializeChannel(rqe[c], variance[c], c)
    bf.setMapping(xt_0toN, yt_0toN, xt_Nto0, yt_Nto0)
    bf.setWeights()
    bf.newImages(images)            # n_sets x n_channels x height x width
    bf.newBackgrounds(backgrounds)
    bf.newPeaks(peaks, "text")      # one channel-0 peaks dict per set
    bf.doFit()
    p = bf.getPeaks(set, channel)
"""
import ctypes

import numpy
from numpy.ctypeslib import ndpointer

from ..sa_library import loadclib
from ..sa_library.dao_batch import PEAK_COLUMNS, COUNTER_NAMES
from ..sa_library.dao_fit_c import MultiFitterException, formatPeaksGaussianPSF, raiseOnLibraryError

_lib = None
_ENTRY_POINTS = ("mpBatchCreate", "mpBatchInitializeChannel", "mpBatchSetTransforms", "mpBatchSetWeights",
                 "mpBatchNewImages", "mpBatchNewBackgrounds", "mpBatchNewPeaks", "mpBatchFit", "mpBatchGetAllPeaks",
                 "mpBatchRemoveErrorPeaks", "mpBatchRemoveRunningPeaks")


def loadMPBatchC():
    global _lib
    if _lib is None:
        lib = loadclib.loadCLibrary("dao_fit")
        vp = ctypes.c_void_p
        f64 = ndpointer(dtype=numpy.float64)
        i32 = ndpointer(dtype=numpy.int32)
        ci, cd = ctypes.c_int, ctypes.c_double
        lib.mpBatchCreate.argtypes = [ci, ci, ci, ci, ci, cd, ci]
        lib.mpBatchCreate.restype = vp
        lib.mpBatchDestroy.argtypes = [vp]
        lib.mpBatchInitializeChannel.argtypes = [vp, ci, f64, f64, cd, cd]
        lib.mpBatchSetTransforms.argtypes = [vp, f64, f64, f64, f64]
        lib.mpBatchSetWeights.argtypes = [vp, f64, f64, ci, cd, cd]
        lib.mpBatchNewImages.argtypes = [vp, vp, ci]
        lib.mpBatchNewBackgrounds.argtypes = [vp, vp, ci]
        lib.mpBatchNewPeaks.argtypes = [vp, vp, i32, ctypes.c_char_p, ci]
        lib.mpBatchFit.argtypes = [vp, ci]
        lib.mpBatchFit.restype = ci
        lib.mpBatchNFit.argtypes = [vp, ci]
        lib.mpBatchNFit.restype = ci
        lib.mpBatchGetAllPeaks.argtypes = [vp, vp, vp, ci]
        lib.mpBatchRemoveErrorPeaks.argtypes = [vp]
        lib.mpBatchRemoveRunningPeaks.argtypes = [vp]
        lib.mpBatchCounters.argtypes = [vp, ci, ci, i32]
        for name in _ENTRY_POINTS:
            if name != "mpBatchCreate":
                getattr(lib, name).restype = ci
        lib.mpBatchFit.restype = ci
        _lib = raiseOnLibraryError(lib, _ENTRY_POINTS)
    return _lib


def _ptr(a):
    """(pointer, on_device, keep-alive) for a float64 numpy array or torch CUDA tensor."""
    if hasattr(a, "is_cuda") and a.is_cuda:
        import torch
        a = a.to(torch.float64).contiguous()
        torch.cuda.current_stream(a.device).synchronize()
        return ctypes.c_void_p(a.data_ptr()), 1, a
    a = numpy.ascontiguousarray(a, dtype=numpy.float64)
    return a.ctypes.data_as(ctypes.c_void_p), 0, a


class MPBatchFitter(object):
    def __init__(self, n_sets, n_channels, shape, roi_size=12, sigma_range=None, peak_cap=1000, default_tol=1.0e-6):
        if sigma_range is None:
            raise MultiFitterException("sigma_range is required")
        self.n_sets, self.n_channels = int(n_sets), int(n_channels)
        self.shape = tuple(shape)
        self.sigma_range = sigma_range
        self.clib = loadMPBatchC()
        # the single-channel tolerance times the number of channels (mp_fit_c.py:262-269)
        self.handle = self.clib.mpBatchCreate(self.n_sets, self.n_channels, self.shape[1], self.shape[0],
                                              int(roi_size), default_tol * self.n_channels, int(peak_cap))
        if not self.handle:
            raise MultiFitterException("mpBatchCreate failed (no usable GPU, or bad sizes)")

    def cleanup(self):
        if self.handle:
            self.clib.mpBatchDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.cleanup()
        except Exception:
            pass

    def initializeChannel(self, rqe, variance, channel):
        """MPFitDao2D.initializeChannel (mp_fit_dao_c.py:73-81) for channel `channel` of every set."""
        width_max = 1.0 / (2.0 * self.sigma_range[0] * self.sigma_range[0])
        width_min = 1.0 / (2.0 * self.sigma_range[1] * self.sigma_range[1])
        self.clib.mpBatchInitializeChannel(self.handle, int(channel), numpy.ascontiguousarray(rqe, dtype=numpy.float64),
                                           numpy.ascontiguousarray(variance, dtype=numpy.float64), width_min, width_max)

    def setMapping(self, xt_0toN, yt_0toN, xt_Nto0, yt_Nto0):
        t = [numpy.ascontiguousarray(v, dtype=numpy.float64) for v in (xt_0toN, yt_0toN, xt_Nto0, yt_Nto0)]
        self.clib.mpBatchSetTransforms(self.handle, *t)

    def setWeights(self, weights=None, z_offset=0.0, z_scale=0.0):
        """The x / y weights (z rows x channels); None: equal weights, as MPFitDao2D.setWeights."""
        if weights is None:
            w = numpy.ones((2, self.n_channels)) / float(self.n_channels)
            wx = wy = w
        else:
            wx, wy = weights["x"], weights["y"]
        wx = numpy.ascontiguousarray(wx, dtype=numpy.float64)
        wy = numpy.ascontiguousarray(wy, dtype=numpy.float64)
        if wx.shape[1] != self.n_channels or wx.shape != wy.shape or numpy.any(wx <= 0) or numpy.any(wy <= 0):
            raise MultiFitterException("weights must be positive, z rows x channels")
        self.clib.mpBatchSetWeights(self.handle, wx, wy, wx.shape[0], z_offset, z_scale)

    def _frames(self, images, what):
        n = self.n_sets * self.n_channels * self.shape[0] * self.shape[1]
        size = images.numel() if hasattr(images, "numel") else numpy.size(images)
        if size != n:
            raise MultiFitterException("%s must hold n_sets x n_channels x %d x %d values" % ((what,) + self.shape))
        return _ptr(images)

    def newImages(self, images):
        p, dev, keep = self._frames(images, "images")
        self.clib.mpBatchNewImages(self.handle, p, dev)

    def newBackgrounds(self, backgrounds):
        p, dev, keep = self._frames(backgrounds, "backgrounds")
        self.clib.mpBatchNewBackgrounds(self.handle, p, dev)

    def newPeaks(self, peaks, peaks_type):
        """One channel-0 peaks dict per set (as MPFitDao.newPeaks takes for one set)."""
        if len(peaks) != self.n_sets:
            raise MultiFitterException("one peaks dict per set")
        rows = [formatPeaksGaussianPSF(p, peaks_type) for p in peaks]
        counts = numpy.array([r.shape[0] for r in rows], dtype=numpy.int32)
        allr = numpy.ascontiguousarray(numpy.concatenate(rows), dtype=numpy.float64)
        self.clib.mpBatchNewPeaks(self.handle, allr.ctypes.data_as(ctypes.c_void_p), counts,
                                  ctypes.c_char_p(peaks_type.encode()), 0)

    def doFit(self, max_iterations=2000):
        """MultiFitter.doFit for every set; returns the largest sweep count."""
        return self.clib.mpBatchFit(self.handle, int(max_iterations))

    def getNFit(self, s):
        return self.clib.mpBatchNFit(self.handle, int(s))

    def getAllPeaks(self):
        """(rows n x 9 in PEAK_COLUMNS order, status) of every frame in frame order, and per-set counts."""
        counts = numpy.array([self.getNFit(s) for s in range(self.n_sets)], dtype=numpy.int64)
        n = int(counts.sum()) * self.n_channels
        out = numpy.zeros((max(n, 1), 9))
        status = numpy.zeros(max(n, 1), dtype=numpy.int32)
        if n:
            self.clib.mpBatchGetAllPeaks(self.handle, out.ctypes.data_as(ctypes.c_void_p),
                                         status.ctypes.data_as(ctypes.c_void_p), 0)
        return out[:n], status[:n], counts

    def getPeaks(self, s, channel):
        out, status, counts = self.getAllPeaks()
        at = int(counts[:s].sum()) * self.n_channels + int(counts[s]) * channel
        n = int(counts[s])
        d = {c: out[at:at + n, k].copy() for k, c in enumerate(PEAK_COLUMNS)}
        d["iterations"] = d["iterations"].astype(numpy.int32)
        d["status"] = status[at:at + n].copy()
        return d

    def removeErrorPeaks(self):
        self.clib.mpBatchRemoveErrorPeaks(self.handle)

    def removeRunningPeaks(self):
        self.clib.mpBatchRemoveRunningPeaks(self.handle)

    def getCounters(self, s, channel):
        c = numpy.zeros(8, dtype=numpy.int32)
        self.clib.mpBatchCounters(self.handle, int(s), int(channel), c)
        return dict(zip(COUNTER_NAMES, c.tolist()))
