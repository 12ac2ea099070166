"""
Multiplane 3D-DAOSTORM: one 2D Gaussian fitter per plane, coupled
(storm_analysis/multi_plane/mp_fit_dao_c.py).
"""
import ctypes

import numpy
from numpy.ctypeslib import ndpointer

from ..sa_library import dao_fit_c as daoFitC
from ..sa_library import loadclib
from . import mp_fit_c as mpFitC

_lib = None


def loadMPFitCDao():
    """
    c_libraries/libmp_fit_dao.so (mp_fit_dao_c.py:17-36), with every entry
    point of the multiplane fitter declared.  Raises if it is absent.
    """
    global _lib
    if _lib is None:
        _lib = setMPFitDaoArgtypes(loadclib.loadCLibrary("mp_fit_dao"))
    return _lib


def setMPFitDaoArgtypes(mp_fit):
    mp_fit.mpDaoInitialize2DChannel.argtypes = [ctypes.c_void_p,
                                                ndpointer(dtype=numpy.float64),
                                                ndpointer(dtype=numpy.float64),
                                                ctypes.c_double,
                                                ctypes.c_double,
                                                ctypes.c_int,
                                                ctypes.c_int]
    mp_fit.mpDaoNewPeaks.argtypes = [ctypes.c_void_p,
                                     ndpointer(dtype=numpy.float64),
                                     ctypes.c_char_p,
                                     ctypes.c_int]
    return mpFitC.addMPFitC(mp_fit)


class MPFitDao(mpFitC.MPFit):
    """
    Multiplane fitting with the 3D-DAOSTORM fitters (mp_fit_dao_c.py:39-62).
    """
    def __init__(self, n_channels=None, roi_size=None, sigma_range=None, **kwds):
        super(MPFitDao, self).__init__(**kwds)
        self.independent_heights = 1
        self.n_channels = n_channels
        self.roi_size = roi_size
        self.sigma_range = sigma_range
        self.clib = loadMPFitCDao()

    def newPeaks(self, peaks, peaks_type):
        """Peaks in channel 0's frame; the library maps x, y to the other channels."""
        c_peaks = daoFitC.formatPeaksGaussianPSF(peaks, peaks_type)
        self.clib.mpDaoNewPeaks(self.mfit, c_peaks, ctypes.c_char_p(peaks_type.encode()), c_peaks.shape[0])


class MPFitDao2D(MPFitDao):
    """
    One 2D Gaussian fitter (free width) per plane (mp_fit_dao_c.py:65-94).
    """
    def initializeChannel(self, rqe, variance, channel):
        super(MPFitDao2D, self).initializeChannel(rqe, variance, channel)
        width_max = 1.0 / (2.0 * self.sigma_range[0] * self.sigma_range[0])
        width_min = 1.0 / (2.0 * self.sigma_range[1] * self.sigma_range[1])
        self.clib.mpDaoInitialize2DChannel(self.mfit,
                                           numpy.ascontiguousarray(rqe, dtype=numpy.float64),
                                           numpy.ascontiguousarray(variance, dtype=numpy.float64),
                                           width_min, width_max, self.roi_size, channel)

    def setWeights(self):
        """Equal weights: z is not a fit parameter here (mp_fit_dao_c.py:83-94)."""
        n = self.n_channels
        weights = {k: numpy.ones((2, n)) / float(n) for k in ("bg", "h", "x", "y", "z")}
        super(MPFitDao2D, self).setWeights(weights, 0.0, 0.0, verbose=False)
