#!/usr/bin/env python
"""
Local-maxima peak finding on the GPU, as the reference's MaximaFinder
(storm_analysis/sa_library/ia_utilities_c.py:132-252) does it with
findLocalMaxima (ia_utilities.c:105-220): same constructor, same findMaxima /
resetTaken methods, same results (peak list, order, heights and the taken
counters that limit how often a pixel is returned).  The taken counters live
on the GPU; images may be numpy arrays or torch CUDA tensors (no host copy).

    mxf = MaximaFinder(margin=10, radius=2.0, threshold=6.0, z_values=[0.0])
    [x, y, z] = mxf.findMaxima([image])

Needs the HIP library and a GPU: otherwise the calls raise
MultiFitterException (there is no CPU path).
"""
import ctypes

import numpy

from . import loadclib
from .dao_fit_c import MultiFitterException

_lib = None


def loadFinderC():
    global _lib
    if _lib is None:
        lib = loadclib.loadCLibrary("dao_fit")
        vp = ctypes.c_void_p
        ci = ctypes.c_int
        cd = ctypes.c_double
        lib.daoFindLocalMaxima.argtypes = [vp, vp, ci, ci, ci, ci, ci, cd, cd, vp, ci, vp, vp, vp, vp,
                                           ctypes.POINTER(ci)]
        lib.daoFindLocalMaxima.restype = ci
        _lib = lib
    return _lib


class MaximaFinder(object):
    """
    For finding local maxima in an image or an image stack (one plane per z value).
    """
    def __init__(self, margin=None, n_duplicates=1, radius=None, threshold=None, z_range=None, z_values=None,
                 device="cuda", **kwds):
        """
        margin - Margin around the edge to avoid in the search.
        n_duplicates - How many times we can return the same location (in multiple calls) before
                       giving up.
        radius - Radius in pixels over which the current pixel needs to be a maxima.
        threshold - Minimum value to be considered as a maxima.
        z_range - Number of z planes over which the current pixel needs to be a maxima.
        z_values - Z values to fill in the z array, there should be one of these for each plane.
        """
        super(MaximaFinder, self).__init__(**kwds)
        self.n_duplicates = n_duplicates
        self.n_planes = len(z_values)
        self.taken = None
        self.device = device
        if z_range is None:
            z_range = self.n_planes
        # as the reference: one more than the margin, so that no peak lies on its edge
        self.margin = int(margin + 1)
        self.z_range = int(z_range)
        self.radius = float(radius)
        self.threshold = float(threshold)
        self.z_values = numpy.ascontiguousarray(numpy.array(z_values), dtype=numpy.float64)

    def _stack(self, images):
        import torch
        if isinstance(images, torch.Tensor):
            st = images if images.dim() == 3 else images.unsqueeze(0)
        else:
            planes = [im if isinstance(im, torch.Tensor) else torch.from_numpy(numpy.ascontiguousarray(im))
                      for im in images]
            st = torch.stack([p.to(self.device, torch.float64) for p in planes])
        st = st.to(self.device, torch.float64).contiguous()
        return st

    def findMaxima(self, images, want_height=False):
        """
        Find the (local) maxima in a list of one or more images, in the order of
        z_values.  Returns [x, y, z] (and h with want_height) as numpy arrays.
        """
        import torch
        assert (len(images) == self.n_planes), "Number of planes does not match number of Z planes."
        st = self._stack(images)
        zs, ys, xs = st.shape
        if self.taken is None:
            self.taken = torch.full((zs, ys, xs), 1 - self.n_duplicates, dtype=torch.int32, device=self.device)
        assert (ys == self.taken.shape[1]), "Unexpected image x size!"
        assert (xs == self.taken.shape[2]), "Unexpected image y size!"
        # as the reference: room for every candidate pixel, so the search never cuts
        max_npeaks = zs * ys * xs + 1
        out = torch.empty((4, max(max_npeaks, 1)), dtype=torch.float64, device=self.device)
        n = ctypes.c_int(0)
        torch.cuda.current_stream(st.device).synchronize()
        rc = loadFinderC().daoFindLocalMaxima(
            ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(self.taken.data_ptr()), zs, ys, xs, self.margin,
            self.z_range, self.radius, self.threshold, self.z_values.ctypes.data_as(ctypes.c_void_p), max_npeaks,
            ctypes.c_void_p(out[0].data_ptr()), ctypes.c_void_p(out[1].data_ptr()),
            ctypes.c_void_p(out[2].data_ptr()), ctypes.c_void_p(out[3].data_ptr()), ctypes.byref(n))
        if rc != 0:
            raise MultiFitterException("local maxima search failed (%d)" % rc)
        res = out[:, :n.value].cpu().numpy()
        z, y, x, h = (numpy.ascontiguousarray(r) for r in res)
        if want_height:
            return [x, y, z, h]
        return [x, y, z]

    def resetTaken(self):
        """Restore the taken counters to their original values."""
        if self.taken is None:
            return
        self.taken.fill_(1 - self.n_duplicates)
