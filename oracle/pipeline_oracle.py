"""
The 3D-DAOSTORM analysis loop on the CPU -- TEST INFRASTRUCTURE ONLY.

A frame-by-frame restatement of PeakFinderFitter.analyzeImage
(storm_analysis/sa_library/fitting.py:866-902) with PeakFinderGaussian
(:193-501), PeakFitter.fitPeaks (:701-790), initFindAndFit
(daostorm_3d/find_peaks.py:15-100) and mufit_analysis.analyze
(daostorm_3d/mufit_analysis.py:14-38), built from the other oracles: the
FP64 fitter (liboracle_dao.so, driven through the package's dao_fit_c
classes), the FFT matched filter with convolveMemo and the finder arithmetic
(mfilter_oracle.py), the maxima search (finder_oracle.py), the post-fit
filters (filter_oracle.py) and the ingest (ingest_oracle.py).  As in the
reference, one background filter (and its memo) serves the whole movie,
frame after frame.  Only tests/ use it, as the checker of
storm_analysis_amd.sa_library.dao_analysis.
"""
import ctypes
import os
from unittest import mock

import numpy

import filter_oracle
import finder_oracle
import ingest_oracle
import mfilter_oracle

HERE = os.path.dirname(os.path.abspath(__file__))
UNCONVERGED_DIST = 5.0


def draw_gaussian_psf(shape, sigma):
    """
    gaussianPSF (fitting.py:38-45) through drawGaussiansXY
    (draw_gaussians.c:24-80, resolution 1): exp(-((j - px)^2 + (k - py)^2) /
    (2 sigma^2)) at (px, py) = (shape[0] / 2, shape[1] / 2) for rows j in
    [int(px) - int(5 sigma), int(px) + int(5 sigma)] and columns k likewise
    (clipped), zero elsewhere; normalised to unit sum.
    """
    px, py = 0.5 * shape[0], 0.5 * shape[1]
    aw = max(int(5.0 * sigma), 1)
    sg = 1.0 / (2.0 * sigma * sigma)
    img = numpy.zeros(shape)
    for j in range(max(int(px) - aw, 0), min(int(px) + aw + 1, shape[0])):
        for k in range(max(int(py) - aw, 0), min(int(py) + aw + 1, shape[1])):
            img[j, k] += numpy.exp(-1.0 * ((j - px) * (j - px) * sg + (k - py) * (k - py) * sg))
    return img / numpy.sum(img)


def peak_mask(shape, params, margin):
    """peakMask (fitting.py:134-178), square AOI, `:-1` slices as written."""
    mask = numpy.ones(shape)
    if "x_start" in params:
        mask[:, 0:params["x_start"] + margin] = 0.0
    if "x_stop" in params:
        mask[:, params["x_stop"] + margin:-1] = 0.0
    if "y_start" in params:
        mask[0:params["y_start"] + margin, :] = 0.0
    if "y_stop" in params:
        mask[params["y_stop"] + margin:-1, :] = 0.0
    return mask


def _fitter_class(D, model, emodel):
    table = {("2dfixed", "MLE"): D.MultiFitter2DFixed, ("2dfixed", "ALS"): D.MultiFitter2DFixedALS,
             ("2dfixed", "LS"): D.MultiFitter2DFixedLS, ("2dfixed", "DWLS"): D.MultiFitter2DFixedDWLS,
             ("2dfixed", "FWLS"): D.MultiFitter2DFixedFWLS, ("2d", "MLE"): D.MultiFitter2D,
             ("2d", "ALS"): D.MultiFitter2DALS, ("2d", "LS"): D.MultiFitter2DLS, ("2d", "DWLS"): D.MultiFitter2DDWLS,
             ("2d", "FWLS"): D.MultiFitter2DFWLS, ("3d", "MLE"): D.MultiFitter3D, ("Z", "MLE"): D.MultiFitterZ}
    return table[(model, emodel)]


class OracleFinderFitter(object):
    """initFindAndFit + PeakFinderFitter on one frame at a time."""
    def __init__(self, params, oracle_so=None):
        from storm_analysis_amd.sa_library import dao_fit_c as D
        self.p = params
        self.sigma = float(params["sigma"])
        self.threshold = float(params["threshold"])
        self.iterations = int(params["iterations"])
        self.radius = float(params["find_max_radius"])
        self.z_value = float(params.get("z_value", 0.0))
        model = params["model"]
        if params.get("roi_size", -1) != -1:
            self.roi_size = int(params["roi_size"])
        else:
            self.roi_size = int(8.0 * self.sigma)
            if model != "2dfixed":
                self.roi_size = int(1.5 * self.roi_size)
        self.margin = int(self.roi_size / 2 + 2)
        self.neighborhood = self.sigma * UNCONVERGED_DIST
        self.bg_filter = None
        self.peak_mask = None
        kw = {"roi_size": self.roi_size}
        if model in ("2d", "3d"):
            kw["sigma_range"] = params.get("sigma_range", [0.5 * self.sigma, 5.0 * self.sigma])
        elif model == "Z":
            w = []
            for ax in ("wx", "wy"):
                v = numpy.array([float(params.get(ax + s, 0.0)) for s in ("_wo", "_c", "_d", "A", "B", "C", "D")])
                v[0] /= params["pixel_size"]
                v[1] *= 0.001
                v[2] *= 0.001
                w.append(v)
            kw.update(wx_params=w[0], wy_params=w[1], min_z=params.get("min_z", -0.5), max_z=params.get("max_z", 0.5))
        lib = D.setDaoFitArgtypes(ctypes.cdll.LoadLibrary(oracle_so or os.path.join(HERE, "liboracle_dao.so")))
        with mock.patch.object(D, "loadDaoFitC", lambda: lib):
            self.mfitter = _fitter_class(D, model, params.get("fit_error_model", "MLE"))(**kw)

    def analyze_frame(self, image):
        """analyzeImage for one padded photo-electron frame; returns the peak dict (x, y unpadded)."""
        m = self.mfitter
        if self.peak_mask is None:
            self.peak_mask = peak_mask(image.shape, self.p, self.margin)
        if self.bg_filter is None:
            self.bg_filter = mfilter_oracle.MatchedFilterOracle(
                draw_gaussian_psf(image.shape, self.p["background_sigma"]), memoize=True, max_diff=1.0e-3)
        taken = [numpy.zeros(image.shape, dtype=numpy.int32)]
        cur = self.threshold + (4.0 if self.iterations > 4 else float(self.iterations))
        if not m.isInitialized():
            m.initializeC(image)
        m.newImage(image)
        fit = numpy.zeros(image.shape)
        for _ in range(self.iterations):
            background = self.bg_filter.convolve(image - fit)
            snr = mfilter_oracle.peak_finder_snr(image, background, fit, peak_mask=self.peak_mask)
            z, y, x, h = finder_oracle.find_local_maxima([snr], taken, self.margin + 1, 1, self.radius,
                                                         self.threshold, [self.z_value], snr.size + 1)
            new = {"x": x, "y": y, "z": z, "sigma": numpy.ones(x.size) * self.sigma}
            if cur > self.threshold:
                cur -= 1.0
                done = False
            else:
                done = x.size == 0
            m.newBackground(background)
            fit = self._fit_peaks(new)
            if done:
                break
        m.removeRunningPeaks()
        out = {k: m.getPeakProperty(k) for k in ("x", "y", "z", "height", "background", "xsigma", "ysigma",
                                                   "error", "iterations", "status")}
        out["x"] = out["x"] - float(self.margin)
        out["y"] = out["y"] - float(self.margin)
        return out

    def _fit_peaks(self, new):
        """PeakFitter.fitPeaks (fitting.py:701-790)."""
        m = self.mfitter
        if new["x"].size > 0:
            if m.getNFit() > 0:
                st = m.getPeakProperty("status")
                st = filter_oracle.running_if_has_neighbors(st, m.getPeakProperty("x"), m.getPeakProperty("y"),
                                                            new["x"], new["y"], self.neighborhood)
                m.setPeakStatus(st)
            m.newPeaks(new, "finder")
            m.doFit()
            m.removeErrorPeaks()
            st = m.getPeakProperty("status")
            px, py = m.getPeakProperty("x"), m.getPeakProperty("y")
            n_prox = filter_oracle.mark_dimmer_peaks(px, py, m.getPeakProperty("height"), st, self.sigma,
                                                     self.neighborhood)
            n_sig = filter_oracle.mark_low_significance_peaks(px, py, m.getPeakProperty("significance"), st,
                                                              self.threshold, self.neighborhood)
            if n_prox + n_sig > 0:
                m.setPeakStatus(st)
                m.removeErrorPeaks()
            if m.getUnconverged() > 0:
                m.doFit()
                m.removeErrorPeaks()
        return m.getFitImage()

    def cleanup(self):
        self.mfitter.cleanup(verbose=False)


def read_xml(filename):
    """The typed settings of a reference XML file (int, float, string)."""
    import xml.etree.ElementTree as ElementTree
    out = {}
    for node in ElementTree.parse(filename).getroot():
        kind = node.attrib.get("type", "string")
        text = (node.text or "").strip()
        out[node.tag] = int(text) if kind == "int" else float(text) if kind == "float" else text
    return out


def analyze_dax(dax_name, xml_name, width, height, n_frames, big_endian=True, oracle_so=None):
    """mufit_analysis.analyze on the CPU: the per-frame peak dicts of a .dax movie (standard camera)."""
    p = read_xml(xml_name)
    ff = OracleFinderFitter(p, oracle_so)
    out = []
    with open(dax_name, "rb") as fp:
        for f in range(n_frames):
            fp.seek(f * width * height * 2)
            raw = numpy.fromfile(fp, dtype=numpy.uint16, count=width * height).reshape(1, height, width)
            img = ingest_oracle.ingest(raw, p["camera_offset"], 1.0 / p["camera_gain"], 1.0, ff.margin,
                                       big_endian)[0]
            out.append(ff.analyze_frame(img))
    ff.cleanup()
    return out
