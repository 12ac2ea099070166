#!/usr/bin/env python
"""
Python surface of the MI355X 3D-DAOSTORM fitter.

Drop-in for storm_analysis/sa_library/dao_fit_c.py: the same classes
(MultiFitter, MultiFitterGaussian, MultiFitter2DFixed{,ALS,LS,DWLS,FWLS,NC,
DWLSNC}, MultiFitter2D{,ALS,LS,DWLS,FWLS}, MultiFitter3D, MultiFitterZ), the same
methods and the same C-ABI symbol names (dao_fit_c.py:147-238).  Every call
goes through ctypes into c_libraries/libdao_fit.so, whose entry points run HIP
kernels on the GPU and return only when the work is finished.
"""
import ctypes
import numpy
from numpy.ctypeslib import ndpointer

from . import loadclib

# Peak status values (multi_fit.h:33-35).
RUNNING = 0
CONVERGED = 1
ERROR = 2


class fitData(ctypes.Structure):
    """
    Leading fields of the native handle; layout of the reference's ctypes
    mirror (dao_fit_c.py:22-77, multi_fit.h:80-143).  Only the prefix is read
    from Python; the native side keeps the counters here current after every
    call, and reads n_iterations back (Python resets it).
    """
    _fields_ = [('n_dposv', ctypes.c_int),
                ('n_iterations', ctypes.c_int),
                ('n_lost', ctypes.c_int),
                ('n_margin', ctypes.c_int),
                ('n_neg_fi', ctypes.c_int),
                ('n_neg_height', ctypes.c_int),
                ('n_non_converged', ctypes.c_int),
                ('n_non_decr', ctypes.c_int),

                ('fit_size_x', ctypes.c_int),
                ('fit_size_y', ctypes.c_int),
                ('image_size_x', ctypes.c_int),
                ('image_size_y', ctypes.c_int),
                ('jac_size', ctypes.c_int),
                ('max_nfit', ctypes.c_int),
                ('nfit', ctypes.c_int),
                ('roi_n_index', ctypes.c_int),

                ('min_height', ctypes.c_double),
                ('xoff', ctypes.c_double),
                ('yoff', ctypes.c_double),
                ('zoff', ctypes.c_double),
                ('tolerance', ctypes.c_double),

                ('bg_counts', ctypes.POINTER(ctypes.c_int)),
                ('roi_x_index', ctypes.POINTER(ctypes.c_int)),
                ('roi_y_index', ctypes.POINTER(ctypes.c_int)),
                ('stale', ctypes.POINTER(ctypes.c_int)),

                ('as_xi', ctypes.POINTER(ctypes.c_double)),
                ('bg_data', ctypes.POINTER(ctypes.c_double)),
                ('bg_estimate', ctypes.POINTER(ctypes.c_double)),
                ('err_i', ctypes.POINTER(ctypes.c_double)),
                ('f_data', ctypes.POINTER(ctypes.c_double)),
                ('rqe', ctypes.POINTER(ctypes.c_double)),
                ('scmos_term', ctypes.POINTER(ctypes.c_double)),
                ('t_fi', ctypes.POINTER(ctypes.c_double)),
                ('x_data', ctypes.POINTER(ctypes.c_double)),

                ('working_peak', ctypes.c_void_p),
                ('fit', ctypes.c_void_p),

                ('fit_model', ctypes.c_void_p),

                ('fn_alloc_peaks', ctypes.c_void_p),
                ('fn_calc_JH', ctypes.c_void_p),
                ('fn_calc_peak_shape', ctypes.c_void_p),
                ('fn_check', ctypes.c_void_p),
                ('fn_copy_peak', ctypes.c_void_p),
                ('fn_error_fn', ctypes.c_void_p),
                ('fn_free_peaks', ctypes.c_void_p),
                ('fn_peak_sum', ctypes.c_void_p),
                ('fn_update', ctypes.c_void_p)]


class MultiFitterException(Exception):
    pass


def formatPeaksGaussianPSF(peaks, peaks_type):
    """
    (N,4) {x,y,z,sigma} for "finder"/"testing", (N,7) {x,y,z,background,
    height,xsigma,ysigma} for "text"/"hdf5" (dao_fit_c.py:114-144).
    """
    if (peaks_type == "testing") or (peaks_type == "finder"):
        cols = ("x", "y", "z", "sigma")
    elif (peaks_type == "text") or (peaks_type == "hdf5"):
        cols = ("x", "y", "z", "background", "height", "xsigma", "ysigma")
    else:
        raise MultiFitterException("Unknown peaks type '" + peaks_type + "'")
    return numpy.ascontiguousarray(numpy.stack([peaks[c] for c in cols], axis=1),
                                   dtype=numpy.float64)


def setDaoFitArgtypes(lib):
    """
    Declare the C-ABI (dao_fit_c.py:147-238) on a loaded library.
    """
    vp = ctypes.c_void_p
    f64 = ndpointer(dtype=numpy.float64)
    i32 = ndpointer(dtype=numpy.int32)

    lib.mFitAnscombeTransformImage.argtypes = [vp]
    lib.mFitGetFitImage.argtypes = [vp, f64]
    lib.mFitGetNError.argtypes = [vp]
    lib.mFitGetNError.restype = ctypes.c_int
    lib.mFitGetPeakPropertyDouble.argtypes = [vp, f64, ctypes.c_char_p]
    lib.mFitGetPeakPropertyInt.argtypes = [vp, i32, ctypes.c_char_p]
    lib.mFitGetResidual.argtypes = [vp, f64]
    lib.mFitGetUnconverged.argtypes = [vp]
    lib.mFitGetUnconverged.restype = ctypes.c_int
    lib.mFitIterateLM.argtypes = [vp]
    if hasattr(lib, "mFitDoFit"):
        lib.mFitDoFit.argtypes = [vp, ctypes.c_int]
        lib.mFitDoFit.restype = ctypes.c_int
    lib.mFitNewBackground.argtypes = [vp, f64]
    lib.mFitNewImage.argtypes = [vp, f64]
    lib.mFitRemoveErrorPeaks.argtypes = [vp]
    lib.mFitRemoveRunningPeaks.argtypes = [vp]
    lib.mFitSetPeakStatus.argtypes = [vp, i32]

    lib.daoCleanup.argtypes = [vp]
    lib.daoInitialize.argtypes = [f64, f64, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int]
    lib.daoInitialize.restype = ctypes.POINTER(fitData)
    for name in ("daoInitialize2DFixed", "daoInitialize2DFixedALS", "daoInitialize2DFixedLS",
                 "daoInitialize2DFixedDWLS", "daoInitialize2DFixedFWLS"):
        getattr(lib, name).argtypes = [vp]
    for name in ("daoInitialize2D", "daoInitialize2DALS", "daoInitialize2DLS",
                 "daoInitialize2DDWLS", "daoInitialize2DFWLS", "daoInitialize3D"):
        getattr(lib, name).argtypes = [vp, ctypes.c_double, ctypes.c_double]
    lib.daoInitializeZ.argtypes = [vp, f64, f64, ctypes.c_double, ctypes.c_double]
    lib.daoNewPeaks.argtypes = [vp, f64, ctypes.c_char_p, ctypes.c_int]
    if hasattr(lib, "daoSetTeam"):
        lib.daoSetTeam.argtypes = [vp, ctypes.c_int]
    if hasattr(lib, "daoSetWaitLimit"):
        lib.daoSetWaitLimit.argtypes = [vp, ctypes.c_longlong]
    return raiseOnLibraryError(lib, _FIT_ENTRY_POINTS)


_FIT_ENTRY_POINTS = (
    "mFitAnscombeTransformImage", "mFitGetFitImage", "mFitGetNError", "mFitGetPeakPropertyDouble",
    "mFitGetPeakPropertyInt", "mFitGetResidual", "mFitGetUnconverged", "mFitIterateLM", "mFitDoFit",
    "mFitNewBackground", "mFitNewImage", "mFitRemoveErrorPeaks", "mFitRemoveRunningPeaks", "mFitSetPeakStatus",
    "daoInitialize", "daoNewPeaks")


def lastLibraryError(lib):
    """(code, message) of the calling thread's last failed call of the library; clears it."""
    msg = ctypes.create_string_buffer(512)
    code = lib.daoGetLastError(msg, len(msg))
    return code, msg.value.decode(errors="replace")


def raiseOnLibraryError(lib, names, exception=None):
    """
    Not in the reference: the library never aborts the process; a call
    that fails returns and leaves its error for daoGetLastError().  This
    makes each named entry point raise (MultiFitterException, or what
    exception(code, message, func, args) returns) after such a call.  A
    library without daoGetLastError (the oracle) is returned unchanged.
    """
    if not hasattr(lib, "daoGetLastError"):
        return lib
    lib.daoGetLastError.argtypes = [ctypes.c_char_p, ctypes.c_int]
    lib.daoGetLastError.restype = ctypes.c_int

    def errcheck(result, func, args):
        code, msg = lastLibraryError(lib)
        if code != 0:
            if exception is not None:
                raise exception(code, msg, func, args)
            raise MultiFitterException("%s failed (code %d): %s" % (func.__name__, code, msg))
        return result

    for name in names:
        if hasattr(lib, name):
            getattr(lib, name).errcheck = errcheck
    return lib


_lib = None


def loadDaoFitC():
    """
    Load c_libraries/libdao_fit.so (the HIP build).  Raises if it is absent.
    """
    global _lib
    if _lib is None:
        _lib = setDaoFitArgtypes(loadclib.loadCLibrary("dao_fit"))
    return _lib


def printFittingInfo(mfit, spacing="  "):
    """
    Print the fitter's counters (dao_fit_c.py:241-251).
    """
    print(spacing, mfit.contents.n_dposv, "fits reset due to Cholesky failure.")
    print(spacing, mfit.contents.n_margin, "fits reset due to image margin.")
    print(spacing, mfit.contents.n_neg_fi, "fits reset due to negative value in fit function.")
    print(spacing, mfit.contents.n_neg_height, "fits reset due to negative height.")
    print(spacing, mfit.contents.n_non_decr, "fits reset due to non-decreasing error (LM).")
    print(spacing, mfit.contents.n_non_converged, "fits did not converge.")
    print(spacing, mfit.contents.n_lost, "fits were lost.")


class MultiFitter(object):
    """
    Base class for fitting many, possibly overlapping, emitters on one frame
    (dao_fit_c.py:257-549):

     1. create one instance of a sub-class,
     2. newImage() for each frame,
     3. newBackground() with a background estimate,
     4. newPeaks() with starting peaks,
     5. doFit(),
     6. getPeakProperty() for the results,
     7. cleanup().

    Instances share no state and can be used side by side.
    """
    def __init__(self, rqe=None, scmos_cal=None, verbose=False, min_z=None, max_z=None, **kwds):
        super(MultiFitter, self).__init__(**kwds)
        self.clib = None
        self.default_tol = 1.0e-6
        self.im_shape = None
        self.iterations = 0
        self.max_z = max_z
        self.mfit = None
        self.min_z = min_z
        self.n_proximity = 0
        self.n_significance = 0

        self.peak_properties = {"background": "float",
                                "bg_sum": "float",
                                "error": "float",
                                "fg_sum": "float",
                                "height": "float",
                                "iterations": "int",
                                "jacobian": "float",
                                "significance": "compound",
                                "sum": "float",
                                "status": "int",
                                "x": "float",
                                "xsigma": "float",
                                "y": "float",
                                "ysigma": "float",
                                "z": "float"}

        self.rqe = rqe
        self.scmos_cal = scmos_cal
        self.verbose = verbose

    def cleanup(self, spacing="  ", verbose=True):
        """
        Prints the analysis statistics (the sub-class frees the handle).
        """
        if self.mfit is not None:
            if verbose:
                printFittingInfo(self.mfit, spacing=spacing)
                print(spacing, self.n_proximity, "peaks lost to proximity filter.")
                print(spacing, self.n_significance, "peaks lost to low significance.")
                print(spacing, self.iterations, "fitting iterations.")

    def doFit(self, max_iterations=2000):
        """
        One sweep, then more while any peak is RUNNING, at most max_iterations
        more (dao_fit_c.py:332-356).
        """
        if not self.verbose and hasattr(self.clib, "mFitDoFit"):
            # the same sweeps in one native call
            self.clib.mFitDoFit(self.mfit, max_iterations)
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
        fit_image = numpy.zeros(self.im_shape, dtype=numpy.float64)
        self.clib.mFitGetFitImage(self.mfit, fit_image)
        return fit_image

    def getIterations(self):
        """
        Fold the native cycle counter into the Python one and reset it
        (dao_fit_c.py:367-375).
        """
        self.iterations += self.mfit.contents.n_iterations
        self.mfit.contents.n_iterations = 0
        return self.iterations

    def getNError(self):
        return self.clib.mFitGetNError(self.mfit)

    def getNFit(self):
        return self.mfit.contents.nfit

    def getNFitMax(self):
        return self.mfit.contents.max_nfit

    def getPeakProperty(self, p_name):
        """
        Return a numpy array with the requested property (dao_fit_c.py:398-439).
        """
        if p_name not in self.peak_properties:
            raise MultiFitterException("No such property '" + p_name + "'")

        kind = self.peak_properties[p_name]
        if kind == "compound":
            if self.getNFit() == 0:
                return numpy.zeros(0, dtype=numpy.float64)
            if p_name == "significance":
                bg_sum = self.getPeakProperty("bg_sum")
                fg_sum = self.getPeakProperty("fg_sum")
                return fg_sum / numpy.sqrt(bg_sum)

        elif kind == "float":
            if p_name == "jacobian":
                values = numpy.zeros((self.getNFit(), self.mfit.contents.jac_size),
                                     dtype=numpy.float64)
            else:
                values = numpy.zeros(self.getNFit(), dtype=numpy.float64)
            self.clib.mFitGetPeakPropertyDouble(self.mfit, values,
                                                ctypes.c_char_p(p_name.encode()))
            return values

        elif kind == "int":
            values = numpy.zeros(self.getNFit(), dtype=numpy.int32)
            self.clib.mFitGetPeakPropertyInt(self.mfit, values, ctypes.c_char_p(p_name.encode()))
            return values

    def getResidual(self):
        residual = numpy.zeros(self.im_shape, dtype=numpy.float64)
        self.clib.mFitGetResidual(self.mfit, residual)
        return residual

    def getUnconverged(self):
        return self.clib.mFitGetUnconverged(self.mfit)

    def incProximityCounter(self, n_inc):
        self.n_proximity += n_inc

    def incSignificanceCounter(self, n_inc):
        self.n_significance += n_inc

    def initializeC(self, image):
        """
        Check / default the camera calibration arrays (dao_fit_c.py:461-488).
        """
        if self.scmos_cal is None:
            if self.verbose:
                print("Using zeros for sCMOS calibration data.")
            self.scmos_cal = numpy.ascontiguousarray(numpy.zeros(image.shape), dtype=numpy.float64)
        else:
            self.scmos_cal = numpy.ascontiguousarray(self.scmos_cal, dtype=numpy.float64)

        if self.rqe is None:
            if self.verbose:
                print("Using ones for relative quantum efficiency data.")
            self.rqe = numpy.ascontiguousarray(numpy.ones(image.shape), dtype=numpy.float64)
        else:
            self.rqe = numpy.ascontiguousarray(self.rqe, dtype=numpy.float64)

        if (image.shape[0] != self.scmos_cal.shape[0]) or (image.shape[1] != self.scmos_cal.shape[1]):
            raise MultiFitterException("Image shape and sCMOS calibration shape do not match.")
        if (image.shape[0] != self.rqe.shape[0]) or (image.shape[1] != self.rqe.shape[1]):
            raise MultiFitterException("Image shape and RQE shape do not match.")

        self.im_shape = self.scmos_cal.shape

    def isInitialized(self):
        return self.mfit is not None

    def iterate(self):
        self.clib.mFitIterateLM(self.mfit)

    def newBackground(self, background):
        if (background.shape[0] != self.im_shape[0]) or (background.shape[1] != self.im_shape[1]):
            raise MultiFitterException("Background image shape and the original image shape are not the same.")
        self.clib.mFitNewBackground(self.mfit, numpy.ascontiguousarray(background, dtype=numpy.float64))

    def newImage(self, image):
        if (image.shape[0] != self.im_shape[0]) or (image.shape[1] != self.im_shape[1]):
            raise MultiFitterException("Current image shape and the original image shape are not the same.")
        self.clib.mFitNewImage(self.mfit, numpy.ascontiguousarray(image, dtype=numpy.float64))

    def newPeaks(self, peaks, peaks_type):
        raise MultiFitterException("newPeaks() method not defined.")

    def removeErrorPeaks(self):
        self.clib.mFitRemoveErrorPeaks(self.mfit)

    def removeRunningPeaks(self):
        self.clib.mFitRemoveRunningPeaks(self.mfit)

    def rescaleZ(self, z):
        return z

    def setPeakStatus(self, status):
        assert status.size == self.getNFit()
        self.clib.mFitSetPeakStatus(self.mfit, numpy.ascontiguousarray(status, dtype=numpy.int32))


class MultiFitterGaussian(MultiFitter):
    """
    Gaussian-PSF fitters (dao_fit_c.py:560-603).  The default ROI is 10 px.
    """
    def __init__(self, roi_size=10, wx_params=None, wy_params=None, team=0, **kwds):
        super(MultiFitterGaussian, self).__init__(**kwds)
        self.roi_size = roi_size
        self.team = team
        self.wx_params = wx_params
        self.wy_params = wy_params
        self.clib = loadDaoFitC()

    def cleanup(self, verbose=True):
        super(MultiFitterGaussian, self).cleanup(verbose=verbose)
        if self.mfit is not None:
            self.clib.daoCleanup(self.mfit)
            self.mfit = None

    def formatPeaks(self, peaks, peaks_type):
        return formatPeaksGaussianPSF(peaks, peaks_type)

    def initializeC(self, image):
        super(MultiFitterGaussian, self).initializeC(image)
        self.mfit = self.clib.daoInitialize(self.rqe, self.scmos_cal, self.default_tol,
                                            self.scmos_cal.shape[1], self.scmos_cal.shape[0],
                                            self.roi_size)
        if not self.mfit:
            raise MultiFitterException("daoInitialize failed (no usable GPU, or bad sizes)")
        if self.team:
            self.clib.daoSetTeam(self.mfit, self.team)

    def setWaitLimit(self, polls):
        """
        Not in the reference: polls after which a scheduler wait of the
        fitter gives up (0: the default).  A wait that gives up makes
        iterate() / doFit() raise MultiFitterException and drops the
        frame's peaks.
        """
        self.clib.daoSetWaitLimit(self.mfit, int(polls))

    def setTeam(self, peaks_per_workgroup):
        """
        Not in the reference: fit the frame with several workgroups, one per
        `peaks_per_workgroup` peaks (at most 64; 0: one workgroup).  Same
        results, lower latency for frames with many peaks.
        """
        self.team = peaks_per_workgroup
        if self.mfit is not None:
            self.clib.daoSetTeam(self.mfit, self.team)

    def newPeaks(self, peaks, peaks_type):
        c_peaks = self.formatPeaks(peaks, peaks_type)
        self.clib.daoNewPeaks(self.mfit, c_peaks, ctypes.c_char_p(peaks_type.encode()),
                              c_peaks.shape[0])


def _widths(sigma_range):
    """
    sigma range -> (width_min, width_max), width = 1/(2 sigma^2)
    (dao_fit_c.py:708-712).
    """
    width_max = 1.0 / (2.0 * sigma_range[0] * sigma_range[0])
    width_min = 1.0 / (2.0 * sigma_range[1] * sigma_range[1])
    return width_min, width_max


class MultiFitter2DFixed(MultiFitterGaussian):
    """Fixed peak width (dao_fit_c.py:606-612)."""
    def initializeC(self, image):
        super(MultiFitter2DFixed, self).initializeC(image)
        self.clib.daoInitialize2DFixed(self.mfit)


class MultiFitter2DFixedALS(MultiFitterGaussian):
    """Fixed width, Anscombe least squares (dao_fit_c.py:615-626)."""
    def initializeC(self, image):
        super(MultiFitter2DFixedALS, self).initializeC(image)
        self.clib.daoInitialize2DFixedALS(self.mfit)

    def newImage(self, image):
        super(MultiFitter2DFixedALS, self).newImage(image)
        self.clib.mFitAnscombeTransformImage(self.mfit)


class MultiFitter2DFixedLS(MultiFitterGaussian):
    """Fixed width, least squares (dao_fit_c.py:629-636)."""
    def initializeC(self, image):
        super(MultiFitter2DFixedLS, self).initializeC(image)
        self.clib.daoInitialize2DFixedLS(self.mfit)


class MultiFitter2DFixedDWLS(MultiFitterGaussian):
    """Fixed width, data weighted least squares (dao_fit_c.py:639-646)."""
    def initializeC(self, image):
        super(MultiFitter2DFixedDWLS, self).initializeC(image)
        self.clib.daoInitialize2DFixedDWLS(self.mfit)


class MultiFitter2DFixedFWLS(MultiFitterGaussian):
    """Fixed width, fit weighted least squares (dao_fit_c.py:649-656)."""
    def initializeC(self, image):
        super(MultiFitter2DFixedFWLS, self).initializeC(image)
        self.clib.daoInitialize2DFixedFWLS(self.mfit)


class MultiFitter2DFixedNC(MultiFitter2DFixed):
    """Fixed width with the RQE correction disabled (dao_fit_c.py:659-673)."""
    def __init__(self, **kwds):
        super(MultiFitter2DFixedNC, self).__init__(**kwds)
        self.rqe = None


class MultiFitter2DFixedDWLSNC(MultiFitter2DFixedDWLS):
    """DWLS with the RQE correction disabled (dao_fit_c.py:676-693)."""
    def __init__(self, **kwds):
        super(MultiFitter2DFixedDWLSNC, self).__init__(**kwds)
        self.rqe = None


class _SigmaRange(MultiFitterGaussian):
    _init_name = None

    def __init__(self, sigma_range=None, **kwds):
        super(_SigmaRange, self).__init__(**kwds)
        self.sigma_range = sigma_range

    def initializeC(self, image):
        super(_SigmaRange, self).initializeC(image)
        width_min, width_max = _widths(self.sigma_range)
        getattr(self.clib, self._init_name)(self.mfit, width_min, width_max)


class MultiFitter2D(_SigmaRange):
    """Variable width, same in x and y (dao_fit_c.py:696-712)."""
    _init_name = "daoInitialize2D"


class MultiFitter2DALS(_SigmaRange):
    """Variable width, Anscombe least squares (dao_fit_c.py:715-736)."""
    _init_name = "daoInitialize2DALS"

    def newImage(self, image):
        super(MultiFitter2DALS, self).newImage(image)
        self.clib.mFitAnscombeTransformImage(self.mfit)


class MultiFitter2DLS(_SigmaRange):
    """Variable width, least squares (dao_fit_c.py:739-756)."""
    _init_name = "daoInitialize2DLS"


class MultiFitter2DDWLS(_SigmaRange):
    """Variable width, data weighted least squares (dao_fit_c.py:759-776)."""
    _init_name = "daoInitialize2DDWLS"


class MultiFitter2DFWLS(_SigmaRange):
    """Variable width, fit weighted least squares (dao_fit_c.py:779-796)."""
    _init_name = "daoInitialize2DFWLS"


class MultiFitter3D(_SigmaRange):
    """Independent x and y widths (dao_fit_c.py:799-816)."""
    _init_name = "daoInitialize3D"


class MultiFitterZ(MultiFitterGaussian):
    """Widths follow z through the wx/wy calibration (dao_fit_c.py:819-829)."""
    def initializeC(self, image):
        super(MultiFitterZ, self).initializeC(image)
        self.clib.daoInitializeZ(self.mfit,
                                 numpy.ascontiguousarray(self.wx_params, dtype=numpy.float64),
                                 numpy.ascontiguousarray(self.wy_params, dtype=numpy.float64),
                                 self.min_z, self.max_z)
