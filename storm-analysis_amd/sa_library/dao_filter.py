#!/usr/bin/env python
"""
Post-fit peak filters on the GPU.

The functions keep the names, arguments and results of the reference's
wrappers (storm_analysis/sa_library/ia_utilities_c.py:254-265, 325-336,
404-419), which fitting.py:718-782 calls between doFit passes:

    n = markDimmerPeaks(x, y, h, status, r_removal, r_neighbors)
    n = markLowSignificancePeaks(x, y, significance, status, minimum_significance, r_neighbors)
    status = runningIfHasNeighbors(status, c_x, c_y, n_x, n_y, radius)

(numpy arrays; status, int32, is updated in place).  The *Batch forms take the
peak lists of many frames as torch tensors on the GPU, back to back, with
offsets[f] .. offsets[f + 1] - 1 the peaks of frame f, and filter each frame
on its own without a host copy.

Needs the HIP library and a GPU: otherwise the calls raise
MultiFitterException (there is no CPU path).
"""
import ctypes

import numpy

from . import loadclib
from .dao_fit_c import MultiFitterException

RUNNING = 0
CONVERGED = 1
ERROR = 2

_lib = None


def loadFilterC():
    global _lib
    if _lib is None:
        lib = loadclib.loadCLibrary("dao_fit")
        vp = ctypes.c_void_p
        ci = ctypes.c_int
        cd = ctypes.c_double
        for name in ("markDimmerPeaks", "markLowSignificancePeaks"):
            fn = getattr(lib, name)
            fn.argtypes = [vp, vp, vp, vp, cd, cd, ci]
            fn.restype = ci
        lib.runningIfHasNeighbors.argtypes = [vp, vp, vp, vp, vp, cd, ci, ci]
        lib.runningIfHasNeighbors.restype = None
        for name in ("daoBatchMarkDimmerPeaks", "daoBatchMarkLowSignificancePeaks"):
            fn = getattr(lib, name)
            fn.argtypes = [vp, vp, vp, vp, vp, ci, cd, cd, vp]
            fn.restype = ci
        lib.daoBatchRunningIfHasNeighbors.argtypes = [vp, vp, vp, vp, vp, vp, vp, ci, cd]
        lib.daoBatchRunningIfHasNeighbors.restype = ci
        _lib = lib
    return _lib


def _ptr(a, what, dtype=numpy.float64):
    assert a.flags['C_CONTIGUOUS'], "%s is not C contiguous!" % what
    if a.dtype != dtype:
        raise MultiFitterException("%s must be %s" % (what, numpy.dtype(dtype).name))
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc == -1:
        raise MultiFitterException("bad peak filter arguments")
    if rc < 0:
        raise MultiFitterException("peak filter failed on the device")
    return rc


def markDimmerPeaks(x, y, h, status, r_removal, r_neighbors):
    """
    For each peak, check if it has a brighter neighbor within r_removal, and if
    it does mark the peak for removal (status ERROR) and its CONVERGED
    neighbors within r_neighbors as RUNNING.  Returns the number removed.
    """
    lib = loadFilterC()
    return _check(lib.markDimmerPeaks(_ptr(x, "X"), _ptr(y, "Y"), _ptr(h, "H"), _ptr(status, "Status", numpy.int32),
                                      float(r_removal), float(r_neighbors), x.size))


def markLowSignificancePeaks(x, y, significance, status, minimum_significance, r_neighbors):
    """
    Mark every peak with significance <= minimum_significance for removal
    (status ERROR) and its CONVERGED neighbors within r_neighbors as RUNNING.
    Returns the number removed.
    """
    lib = loadFilterC()
    return _check(lib.markLowSignificancePeaks(_ptr(x, "'X'"), _ptr(y, "'Y'"), _ptr(significance, "'Significance'"),
                                               _ptr(status, "'Status'", numpy.int32), float(minimum_significance),
                                               float(r_neighbors), x.size))


def runningIfHasNeighbors(status, c_x, c_y, n_x, n_y, radius):
    """
    Update status based on proximity of new peaks (n_x, n_y) to current peaks
    (c_x, c_y): CONVERGED peaks with a new peak within radius become RUNNING.
    """
    lib = loadFilterC()
    lib.runningIfHasNeighbors(_ptr(c_x, "C_X"), _ptr(c_y, "C_Y"), _ptr(n_x, "N_X"), _ptr(n_y, "N_Y"),
                              _ptr(status, "Status", numpy.int32), float(radius), c_x.size, n_x.size)
    return status


# ---- batched forms (torch tensors on the GPU) ----

def _dev(t, dtype, what):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()):
        raise MultiFitterException("%s must be a contiguous %s tensor on the GPU" % (what, dtype))
    return ctypes.c_void_p(t.data_ptr())


def _offsets(offsets, n, device):
    import torch
    off = torch.as_tensor(offsets, dtype=torch.int32).to(device).contiguous()
    if off.dim() != 1 or off.numel() < 2:
        raise MultiFitterException("offsets must hold nframes + 1 entries")
    host = off.cpu()
    if int(host[0]) != 0 or int(host[-1]) != n or bool((host[1:] < host[:-1]).any()):
        raise MultiFitterException("offsets must rise from 0 to the number of peaks")
    return off


def _batch_mark(fn, x, y, v, status, offsets, ra, rb):
    import torch
    off = _offsets(offsets, x.numel(), x.device)
    for t in (y, v, status):
        if t.numel() != x.numel():
            raise MultiFitterException("peak arrays differ in length")
    nframes = off.numel() - 1
    removed = torch.zeros(nframes, dtype=torch.int32, device=x.device)
    if x.numel() == 0:
        return removed
    torch.cuda.current_stream(x.device).synchronize()
    _check(fn(_dev(x, torch.float64, "x"), _dev(y, torch.float64, "y"), _dev(v, torch.float64, "h"),
              _dev(status, torch.int32, "status"), ctypes.c_void_p(off.data_ptr()), nframes, float(ra), float(rb),
              ctypes.c_void_p(removed.data_ptr())))
    return removed


def markDimmerPeaksBatch(x, y, h, status, offsets, r_removal, r_neighbors):
    """markDimmerPeaks of every frame of a batch; returns the removed count per frame (GPU int32)."""
    return _batch_mark(loadFilterC().daoBatchMarkDimmerPeaks, x, y, h, status, offsets, r_removal, r_neighbors)


def markLowSignificancePeaksBatch(x, y, significance, status, offsets, minimum_significance, r_neighbors):
    """markLowSignificancePeaks of every frame of a batch; returns the removed count per frame."""
    return _batch_mark(loadFilterC().daoBatchMarkLowSignificancePeaks, x, y, significance, status, offsets,
                       minimum_significance, r_neighbors)


def runningIfHasNeighborsBatch(status, c_x, c_y, c_offsets, n_x, n_y, n_offsets, radius):
    """runningIfHasNeighbors of every frame of a batch (status updated in place and returned)."""
    import torch
    coff = _offsets(c_offsets, c_x.numel(), c_x.device)
    noff = _offsets(n_offsets, n_x.numel(), c_x.device)
    if coff.numel() != noff.numel():
        raise MultiFitterException("current and new peaks differ in frame count")
    if c_x.numel() == 0 or n_x.numel() == 0:
        return status
    torch.cuda.current_stream(c_x.device).synchronize()
    _check(loadFilterC().daoBatchRunningIfHasNeighbors(
        _dev(c_x, torch.float64, "c_x"), _dev(c_y, torch.float64, "c_y"), ctypes.c_void_p(coff.data_ptr()),
        _dev(n_x, torch.float64, "n_x"), _dev(n_y, torch.float64, "n_y"), ctypes.c_void_p(noff.data_ptr()),
        _dev(status, torch.int32, "status"), coff.numel() - 1, float(radius)))
    return status
