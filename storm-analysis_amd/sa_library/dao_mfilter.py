#!/usr/bin/env python
"""
FFT image convolution on the GPU: the reference's MatchedFilter
(storm_analysis/sa_library/matched_filter_c.py:43-88) over
libmatched_filter.so built for MI355X -- same constructor, convolve() and
cleanup(), same results to FFT rounding -- plus convolveBatch() for many
images already on the GPU (torch tensors, no host copy).

Needs the HIP library and a GPU: otherwise the calls raise
MatchedFilterException (there is no CPU path).
"""
import ctypes

import numpy
from numpy.ctypeslib import ndpointer

from . import loadclib

_lib = None


class MatchedFilterException(Exception):

    def __init__(self, message):
        Exception.__init__(self, message)


def loadMatchedFilterC():
    global _lib
    if _lib is None:
        lib = loadclib.loadCLibrary("matched_filter")
        d = ndpointer(dtype=numpy.float64)
        lib.cleanup.argtypes = [ctypes.c_void_p]
        lib.convolve.argtypes = [ctypes.c_void_p, d, d]
        lib.convolveMemo.argtypes = [ctypes.c_void_p, d, d]
        lib.initialize.argtypes = [d, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.initialize.restype = ctypes.c_void_p
        lib.daoMatchedFilterConvolveBatch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                      ctypes.c_void_p]
        lib.daoMatchedFilterConvolveBatch.restype = ctypes.c_int
        _lib = lib
    return _lib


def recenterPSF(psf):
    """Roll the PSF by half its size on each axis (recenter_psf.py:9-17)."""
    recentered = numpy.roll(psf, int(psf.shape[0] / 2), axis=0)
    return numpy.roll(recentered, int(psf.shape[1] / 2), axis=1)


class MatchedFilter(object):

    def __init__(self, psf, fftw_estimate=False, memoize=False, max_diff=0.1):
        self.memoize = memoize
        self.psf_shape = psf.shape
        self.lib = loadMatchedFilterC()
        rc_psf = numpy.ascontiguousarray(recenterPSF(psf), dtype=numpy.float64)
        self.mfilter = self.lib.initialize(rc_psf, max_diff if memoize else 0.0, rc_psf.shape[0], rc_psf.shape[1],
                                           int(fftw_estimate))
        if not self.mfilter:
            raise MatchedFilterException("matched filter initialization failed on the device")

    def cleanup(self):
        self.lib.cleanup(self.mfilter)
        self.mfilter = None

    def convolve(self, image):
        if (image.shape[0] != self.psf_shape[0]) or (image.shape[1] != self.psf_shape[1]):
            raise MatchedFilterException("Image shape must match psf shape! " + str(image.shape) + " != " +
                                         str(self.psf_shape))
        image = numpy.ascontiguousarray(image, dtype=numpy.float64)
        result = numpy.zeros(self.psf_shape, dtype=numpy.float64)
        if self.memoize:
            self.lib.convolveMemo(self.mfilter, image, result)
        else:
            self.lib.convolve(self.mfilter, image, result)
        return result

    def convolveBatch(self, images):
        """images: n x X x Y contiguous float64 torch tensor on the GPU -> the n convolutions (GPU)."""
        import torch
        if not (isinstance(images, torch.Tensor) and images.is_cuda and images.dtype == torch.float64 and
                images.is_contiguous() and images.dim() == 3 and tuple(images.shape[1:]) == tuple(self.psf_shape)):
            raise MatchedFilterException("images must be a contiguous n x %d x %d float64 GPU tensor" %
                                         tuple(self.psf_shape))
        out = torch.empty_like(images)
        torch.cuda.current_stream(images.device).synchronize()
        rc = self.lib.daoMatchedFilterConvolveBatch(self.mfilter, ctypes.c_void_p(images.data_ptr()),
                                                    images.shape[0], ctypes.c_void_p(out.data_ptr()))
        if rc != 0:
            raise MatchedFilterException("batched convolution failed (%d)" % rc)
        return out
