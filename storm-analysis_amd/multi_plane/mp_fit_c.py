"""
Multiplane fitting, the Python surface of storm_analysis/multi_plane/mp_fit_c.py.

One emitter is fitted in several image planes (channels) at once: each
channel has its own fitData (a 2D Gaussian fitter of its plane), the LM
steps of the channels are combined into one x, y update through the
channel-to-channel affine transforms, and a peak that fails in one channel
fails in all.  The work runs in c_libraries/libdao_fit.so (loaded under the
reference's name libmp_fit_dao.so); the whole doFit() is one kernel launch.
"""
import ctypes

import numpy
from numpy.ctypeslib import ndpointer

from ..sa_library import dao_fit_c as daoFitC


class mpFitData(ctypes.Structure):
    """The leading fields of the handle, in the reference's layout (mp_fit_c.py:15-52)."""
    _fields_ = [('im_size_x', ctypes.c_int),
                ('im_size_y', ctypes.c_int),
                ('independent_heights', ctypes.c_int),
                ('n_channels', ctypes.c_int),
                ('n_weights', ctypes.c_int),
                ('tolerance', ctypes.c_double),
                ('w_z_offset', ctypes.c_double),
                ('w_z_scale', ctypes.c_double),
                ('zmin', ctypes.c_double),
                ('zmax', ctypes.c_double),
                ('xt_0toN', ctypes.POINTER(ctypes.c_double)),
                ('yt_0toN', ctypes.POINTER(ctypes.c_double)),
                ('xt_Nto0', ctypes.POINTER(ctypes.c_double)),
                ('yt_Nto0', ctypes.POINTER(ctypes.c_double)),
                ('w_bg', ctypes.POINTER(ctypes.c_double)),
                ('w_h', ctypes.POINTER(ctypes.c_double)),
                ('w_x', ctypes.POINTER(ctypes.c_double)),
                ('w_y', ctypes.POINTER(ctypes.c_double)),
                ('w_z', ctypes.POINTER(ctypes.c_double)),
                ('heights', ctypes.POINTER(ctypes.c_double)),
                ('jacobian', ctypes.c_void_p),
                ('w_jacobian', ctypes.c_void_p),
                ('hessian', ctypes.c_void_p),
                ('w_hessian', ctypes.c_void_p),
                ('fit_data', ctypes.POINTER(ctypes.POINTER(daoFitC.fitData))),
                ('fn_cleanup', ctypes.c_void_p),
                ('fn_newpeaks', ctypes.c_void_p),
                ('fn_peak_xi_yi', ctypes.c_void_p),
                ('fn_update', ctypes.c_void_p),
                ('fn_z_range', ctypes.c_void_p)]


_MP_ENTRY_POINTS = ("mpInitialize", "mpIterateLM", "mpDoFit", "mpSetTransforms", "mpSetWeights",
                    "mpSetWeightsIndexing", "mpDaoInitialize2DChannel", "mpDaoNewPeaks")


def addMPFitC(mp_fit):
    """
    Declare the C-ABI shared by the multiplane fitters (mp_fit_c.py:55-127):
    the per-channel mFit* read-outs and the mp* entry points.
    """
    vp = ctypes.c_void_p
    f64 = ndpointer(dtype=numpy.float64)
    i32 = ndpointer(dtype=numpy.int32)

    daoFitC.setDaoFitArgtypes(mp_fit)

    mp_fit.mpCleanup.argtypes = [vp]
    mp_fit.mpInitialize.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    mp_fit.mpInitialize.restype = ctypes.POINTER(mpFitData)
    mp_fit.mpIterateLM.argtypes = [vp]
    if hasattr(mp_fit, "mpDoFit"):
        mp_fit.mpDoFit.argtypes = [vp, ctypes.c_int]
        mp_fit.mpDoFit.restype = ctypes.c_int
    mp_fit.mpSetTransforms.argtypes = [vp, f64, f64, f64, f64]
    mp_fit.mpSetWeights.argtypes = [vp, f64, f64, f64, f64, f64, ctypes.c_int]
    mp_fit.mpSetWeightsIndexing.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    mp_fit.mFitGetPeakPropertyInt.argtypes = [vp, i32, ctypes.c_char_p]
    return daoFitC.raiseOnLibraryError(mp_fit, _MP_ENTRY_POINTS)


class MPFit(daoFitC.MultiFitter):
    """
    Base class for multiplane fitting (mp_fit_c.py:130-361).
    """
    def __init__(self, **kwds):
        super(MPFit, self).__init__(**kwds)
        self.independent_heights = None
        self.n_channels = None

    def _channel(self, i):
        return self.mfit.contents.fit_data[i]

    def cleanup(self, spacing="  ", verbose=True):
        if self.mfit is not None:
            if verbose:
                for i in range(self.n_channels):
                    print()
                    print("Channel", i)
                    daoFitC.printFittingInfo(self._channel(i), spacing=spacing)
                print()
                print(spacing, self.n_proximity, "peaks lost to proximity.")
                print(spacing, self.n_significance, "peaks lost to low significance.")
                print(spacing, "{0:0d} fitting iterations.".format(self.iterations))
                print(spacing, "{0:.1f} fitting iterations/channel.".format(
                    float(self.iterations) / float(self.n_channels)))
            self.clib.mpCleanup(self.mfit)
            self.mfit = None

    def doFit(self, max_iterations=2000):
        """
        One coupled sweep, then more while any peak of channel 0 is RUNNING,
        at most max_iterations more (MultiFitter.doFit, dao_fit_c.py:332-356,
        driving mpIterateLM).  The HIP library runs all of them in one call.
        """
        if not self.verbose and hasattr(self.clib, "mpDoFit"):
            self.clib.mpDoFit(self.mfit, max_iterations)
            self.getIterations()
            return
        i = 0
        self.iterate()
        while self.getUnconverged() and (i < max_iterations):
            if self.verbose and ((i % 20) == 0):
                print("iteration", i)
            self.iterate()
            i += 1
        if self.verbose:
            if i == max_iterations:
                print(" Failed to converge in:", i, self.getUnconverged())
            else:
                print(" Multi-fit converged in:", i, self.getUnconverged())
            print("")
        self.getIterations()

    def getFitImage(self):
        """The fit image of each channel (a list)."""
        fit_images = []
        for i in range(self.n_channels):
            fit_image = numpy.zeros(self.im_shape)
            self.clib.mFitGetFitImage(self._channel(i), fit_image)
            fit_images.append(fit_image)
        return fit_images

    def getResidual(self):
        """The residual of each channel (a list); not in the reference's multiplane surface."""
        out = []
        for i in range(self.n_channels):
            r = numpy.zeros(self.im_shape)
            self.clib.mFitGetResidual(self._channel(i), r)
            out.append(r)
        return out

    def getIterations(self):
        """Fold every channel's cycle counter into the Python one and reset them."""
        for i in range(self.n_channels):
            fit_data = self._channel(i)
            self.iterations += fit_data.contents.n_iterations
            fit_data.contents.n_iterations = 0
        return self.iterations

    def getNError(self, channel=0):
        return self.clib.mFitGetNError(self._channel(channel))

    def getNFit(self, channel=0):
        return self._channel(channel).contents.nfit

    def getNFitMax(self, channel=0):
        return self._channel(channel).contents.max_nfit

    def getPeakProperty(self, p_name, channel=0):
        """
        A property of the peaks of one channel; "significance" sums the
        channels' fg_sum and bg_sum first (mp_fit_c.py:197-236).
        """
        if p_name not in self.peak_properties:
            raise daoFitC.MultiFitterException("No such property '" + p_name + "'")
        kind = self.peak_properties[p_name]
        if kind == "compound":
            if self.getNFit() == 0:
                return numpy.zeros(0, dtype=numpy.float64)
            if p_name == "significance":
                bg_sum = numpy.zeros(self.getNFit(), dtype=numpy.float64)
                fg_sum = numpy.zeros(self.getNFit(), dtype=numpy.float64)
                for i in range(self.n_channels):
                    bg_sum += self.getPeakProperty("bg_sum", channel=i)
                    fg_sum += self.getPeakProperty("fg_sum", channel=i)
                return fg_sum / numpy.sqrt(bg_sum)
        elif kind == "float":
            if p_name == "jacobian":
                values = numpy.zeros((self.getNFit(channel), self._channel(channel).contents.jac_size))
            else:
                values = numpy.zeros(self.getNFit(channel), dtype=numpy.float64)
            self.clib.mFitGetPeakPropertyDouble(self._channel(channel), values, ctypes.c_char_p(p_name.encode()))
            return values
        elif kind == "int":
            values = numpy.zeros(self.getNFit(channel), dtype=numpy.int32)
            self.clib.mFitGetPeakPropertyInt(self._channel(channel), values, ctypes.c_char_p(p_name.encode()))
            return values

    def getUnconverged(self, channel=0):
        return self.clib.mFitGetUnconverged(self._channel(channel))

    def initializeC(self, variance):
        self.im_shape = variance.shape
        # the single-channel tolerance times the number of channels (mp_fit_c.py:262-269)
        self.mfit = self.clib.mpInitialize(self.default_tol * self.n_channels, self.n_channels,
                                           self.independent_heights, variance.shape[1], variance.shape[0])
        if not self.mfit:
            raise daoFitC.MultiFitterException("mpInitialize failed (no usable GPU, or bad sizes)")

    def initializeChannel(self, rqe, variance, channel):
        """
        Check one channel's calibration; creates the handle on the first call
        (mp_fit_c.py:271-286).
        """
        if self.mfit is None:
            self.initializeC(variance)
        if (rqe.shape[0] != self.im_shape[0]) or (rqe.shape[1] != self.im_shape[1]):
            raise daoFitC.MultiFitterException("Current RQE shape and the original variance shape are not the same.")
        if (variance.shape[0] != self.im_shape[0]) or (variance.shape[1] != self.im_shape[1]):
            raise daoFitC.MultiFitterException(
                "Current variance shape and the original variance shape are not the same.")

    def iterate(self):
        self.clib.mpIterateLM(self.mfit)

    def newBackground(self, background):
        """background: one estimate per channel."""
        if len(background) != self.n_channels:
            raise daoFitC.MultiFitterException("Number of images does not match number of channels.")
        for i in range(self.n_channels):
            if (background[i].shape[0] != self.im_shape[0]) or (background[i].shape[1] != self.im_shape[1]):
                raise daoFitC.MultiFitterException(
                    "Background image shape and the original image shape are not the same.")
            self.clib.mFitNewBackground(self._channel(i), numpy.ascontiguousarray(background[i], dtype=numpy.float64))

    def newImage(self, image):
        """image: one image per channel."""
        if len(image) != self.n_channels:
            raise daoFitC.MultiFitterException("Number of images does not match number of channels.")
        for i in range(self.n_channels):
            if (image[i].shape[0] != self.im_shape[0]) or (image[i].shape[1] != self.im_shape[1]):
                raise daoFitC.MultiFitterException("Current image shape and the original image shape are not the same.")
            self.clib.mFitNewImage(self._channel(i), numpy.ascontiguousarray(image[i], dtype=numpy.float64))

    def removeErrorPeaks(self, check=True):
        for i in range(self.n_channels):
            self.clib.mFitRemoveErrorPeaks(self._channel(i))
        if check:
            for i in range(1, self.n_channels):
                assert self.getNFit(0) == self.getNFit(i)

    def removeRunningPeaks(self, check=True):
        for i in range(self.n_channels):
            self.clib.mFitRemoveRunningPeaks(self._channel(i))
        if check:
            for i in range(1, self.n_channels):
                assert self.getNFit(0) == self.getNFit(i)

    def rescaleZ(self, z):
        return z

    def setMapping(self, xt_0toN, yt_0toN, xt_Nto0, yt_Nto0):
        """Affine transforms channel 0 -> N and N -> 0, each [n_channels, 3] (mp_fit_c.py:309-314)."""
        self.clib.mpSetTransforms(self.mfit,
                                  numpy.ascontiguousarray(xt_0toN, dtype=numpy.float64),
                                  numpy.ascontiguousarray(yt_0toN, dtype=numpy.float64),
                                  numpy.ascontiguousarray(xt_Nto0, dtype=numpy.float64),
                                  numpy.ascontiguousarray(yt_Nto0, dtype=numpy.float64))

    def setPeakStatus(self, status):
        assert status.size == self.getNFit()
        for i in range(self.n_channels):
            self.clib.mFitSetPeakStatus(self._channel(i), numpy.ascontiguousarray(status, dtype=numpy.int32))

    def setWeights(self, weights, z_offset, z_scale, verbose=True):
        """
        Per-channel weights of the parameter updates, indexed by z, then
        channel (mp_fit_c.py:323-361); all must be positive.
        """
        if weights["h"].shape[1] != self.n_channels:
            raise daoFitC.MultiFitterException("Incorrect shape for weights array.")
        w = [numpy.ascontiguousarray(weights[k], dtype=numpy.float64) for k in ("bg", "h", "x", "y", "z")]
        for v in w:
            if numpy.count_nonzero(v > 0.0) != v.size:
                raise daoFitC.MultiFitterException("Weights must be positive.")
        self.clib.mpSetWeights(self.mfit, w[0], w[1], w[2], w[3], w[4], w[0].shape[0])
        if verbose:
            print("weights z scaling - offset: {0:.3f} scale: {1:.3f}".format(z_offset, z_scale))
        self.clib.mpSetWeightsIndexing(self.mfit, z_offset, z_scale)
