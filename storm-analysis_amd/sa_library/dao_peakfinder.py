#!/usr/bin/env python
"""
PeakFinderGaussian's search image on the GPU for a batch of frames: the
per-pixel arithmetic of PeakFinderGaussian.peakFinder
(storm_analysis/sa_library/fitting.py:444-501) around its two matched
filters, so that image -> signal-to-noise -> local maxima stays in device
memory:

    snr = peakFinderSNR(images, backgrounds, fits, fg_filter, var_filter, peak_mask)
    [x, y, z] = MaximaFinder(...).findMaxima(snr)        # per frame, or as planes

images, backgrounds and fits are n x X x Y float64 torch tensors on the GPU;
fg_filter / var_filter are dao_mfilter.MatchedFilter objects (or None), the
reference's foreground filter (PSF fg_psf) and variance filter (PSF
fg_psf * fg_psf); peak_mask and camera_variance are X x Y arrays shared by
all frames.  The reference memoizes its foreground filters (max_diff 1e-3);
this batch path always convolves.
"""
import ctypes

import numpy

from . import loadclib
from .dao_fit_c import MultiFitterException

_lib = None


def loadSNRC():
    global _lib
    if _lib is None:
        lib = loadclib.loadCLibrary("dao_fit")
        vp = ctypes.c_void_p
        lib.daoFinderPrepare.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_longlong, vp, vp]
        lib.daoFinderPrepare.restype = ctypes.c_int
        lib.daoFinderSNR.argtypes = [vp, vp, vp, ctypes.c_int, ctypes.c_longlong, vp]
        lib.daoFinderSNR.restype = ctypes.c_int
        _lib = lib
    return _lib


def _dev(t, shape, what):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and
            tuple(t.shape) == tuple(shape)):
        raise MultiFitterException("%s must be a contiguous float64 GPU tensor of shape %s" % (what, shape))
    return ctypes.c_void_p(t.data_ptr())


def _shared(a, shape, device):
    """An X x Y map shared by all frames, on the GPU (None stays None)."""
    import torch
    if a is None:
        return None, ctypes.c_void_p(0)
    t = torch.as_tensor(numpy.ascontiguousarray(a, dtype=numpy.float64) if not isinstance(a, torch.Tensor) else a,
                        dtype=torch.float64).to(device).contiguous()
    if tuple(t.shape) != tuple(shape):
        raise MultiFitterException("shared map must be %s" % (shape,))
    return t, ctypes.c_void_p(t.data_ptr())


def finderPrepare(images, backgrounds, fits, camera_variance=None):
    """(bg_var, foreground) = (background + fit (+ camera variance), image - background - fit)."""
    import torch
    n, X, Y = images.shape
    lib = loadSNRC()
    cam, pcam = _shared(camera_variance, (X, Y), images.device)
    bg_var = torch.empty_like(images)
    fg = torch.empty_like(images)
    torch.cuda.current_stream(images.device).synchronize()
    rc = lib.daoFinderPrepare(_dev(images, (n, X, Y), "images"), _dev(backgrounds, (n, X, Y), "backgrounds"),
                              _dev(fits, (n, X, Y), "fits"), pcam, n, X * Y, ctypes.c_void_p(bg_var.data_ptr()),
                              ctypes.c_void_p(fg.data_ptr()))
    if rc != 0:
        raise MultiFitterException("daoFinderPrepare failed (%d)" % rc)
    return bg_var, fg


def finderSNR(foreground, bg_var, peak_mask=None):
    """foreground / sqrt(bg_var clamped to > 0.1) * peak_mask."""
    import torch
    n, X, Y = foreground.shape
    lib = loadSNRC()
    mask, pmask = _shared(peak_mask, (X, Y), foreground.device)
    snr = torch.empty_like(foreground)
    torch.cuda.current_stream(foreground.device).synchronize()
    rc = lib.daoFinderSNR(_dev(foreground, (n, X, Y), "foreground"), _dev(bg_var, (n, X, Y), "bg_var"), pmask, n,
                          X * Y, ctypes.c_void_p(snr.data_ptr()))
    if rc != 0:
        raise MultiFitterException("daoFinderSNR failed (%d)" % rc)
    return snr


def peakFinderSNR(images, backgrounds, fits, fg_filter=None, var_filter=None, peak_mask=None, camera_variance=None):
    """The search image of PeakFinderGaussian.peakFinder for n frames (fitting.py:444-498)."""
    bg_var, fg = finderPrepare(images, backgrounds, fits, camera_variance)
    if var_filter is not None:
        bg_var = var_filter.convolveBatch(bg_var)
    if fg_filter is not None:
        fg = fg_filter.convolveBatch(fg)
    return finderSNR(fg, bg_var, peak_mask)
