#!/usr/bin/env python
"""
Frame ingest on the GPU: raw camera frames to padded photo-electron frames in
device memory, ready for BatchFitter.newImages (no host copy of the result).

The reference does this per frame on the host (storm_analysis/sa_library/fitting.py:934-937):
DaxReader.loadAFrame (datareader.py:234-243) reads uint16 counts (byte-swapped
for big-endian movies), FrameReaderStd / FrameReaderSCMOS.loadAFrame
(analysis_io.py:243-295) convert them to photo-electrons, and padArray
(fitting.py:112-132) mirror-pads the frame by the finder's margin.  The
classes here keep the readers' names, calibration conventions and results
(bit-identical), for a whole batch of frames per call:

    reader = FrameReaderStd(camera_gain=2.0, camera_offset=100.0)
    images = reader.loadFrames(raw, pad=10)       # torch float64 on the GPU
    bf.newImages(images)                          # BatchFitter, no host copy

Needs torch with a ROCm device; without the HIP library or a GPU the calls
raise MultiFitterException (there is no CPU path).
"""
import ctypes

import numpy

from . import loadclib
from .dao_fit_c import MultiFitterException

_lib = None


def loadIngestC():
    global _lib
    if _lib is None:
        lib = loadclib.loadCLibrary("dao_fit")
        vp = ctypes.c_void_p
        ci = ctypes.c_int
        lib.daoIngestFrames.argtypes = [vp, ci, ci, ci, ci, ci, vp, vp, vp, ci, ci, ci, vp]
        lib.daoIngestFrames.restype = ci
        lib.daoPadFrames.argtypes = [vp, ci, ci, ci, ci, ci, vp]
        lib.daoPadFrames.restype = ci
        _lib = lib
    return _lib


def _device_out(n, height, width, pad, device):
    import torch
    return torch.empty((n, height + 2 * pad, width + 2 * pad), dtype=torch.float64, device=device)


def _array(a, dtype):
    """(pointer, on_device, keep-alive) of a numpy array or torch tensor of `dtype`."""
    if hasattr(a, "data_ptr") and hasattr(a, "is_cuda"):
        import torch
        if a.is_cuda:
            want = torch.uint16 if dtype == numpy.uint16 else torch.float64
            if a.dtype != want or not a.is_contiguous():
                raise MultiFitterException("device arrays must be contiguous %s" % (want,))
            torch.cuda.current_stream(a.device).synchronize()
            return ctypes.c_void_p(a.data_ptr()), 1, a
        a = a.numpy()
    a = numpy.ascontiguousarray(a, dtype=dtype)
    return a.ctypes.data_as(ctypes.c_void_p), 0, a


def _check(rc):
    if rc == -1:
        raise MultiFitterException("bad frame sizes (pad must not exceed the frame's width or height)")
    if rc != 0:
        raise MultiFitterException("frame ingest failed on the device")


class FrameReader(object):
    """
    Conversion of raw counts to photo-electrons with the attributes of the
    reference's FrameReader (analysis_io.py:216-260): offset, gain (1 / camera
    gain) and rqe (1 / relative QE), each a number or a height x width map.
    """
    def __init__(self, offset, gain, rqe=1.0, device="cuda"):
        self.offset = offset
        self.gain = gain
        self.rqe = rqe
        self.device = device

    def _calibration(self, height, width):
        maps = [numpy.ndim(v) > 0 for v in (self.offset, self.gain, self.rqe)]
        if any(maps):
            arr = [numpy.broadcast_to(numpy.asarray(v, dtype=numpy.float64), (height, width))
                   for v in (self.offset, self.gain, self.rqe)]
            return [numpy.ascontiguousarray(a) for a in arr], 1
        return [numpy.array([float(v)]) for v in (self.offset, self.gain, self.rqe)], 0

    def loadFrames(self, raw, pad=0, big_endian=False):
        """
        raw: n x height x width uint16 counts (numpy, or a torch tensor on the
        GPU), as the movie stores them.  Returns the n frames as FrameReader.loadAFrame
        and padArray(frame, pad) would, as a torch float64 tensor on the GPU.
        """
        shape = tuple(raw.shape)
        if len(shape) == 2:
            shape = (1,) + shape
        n, height, width = shape
        lib = loadIngestC()
        out = _device_out(n, height, width, pad, self.device)
        rp, ron, keep = _array(raw, numpy.uint16)
        cal, per_pixel = self._calibration(height, width)
        ptrs = [c.ctypes.data_as(ctypes.c_void_p) for c in cal]
        _check(lib.daoIngestFrames(rp, ron, n, width, height, int(bool(big_endian)), ptrs[0], ptrs[1], ptrs[2],
                                   per_pixel, 0, pad, ctypes.c_void_p(out.data_ptr())))
        return out


class FrameReaderStd(FrameReader):
    """A standard camera (analysis_io.py:263-278): camera_gain in ADU per photo-electron."""
    def __init__(self, camera_gain, camera_offset, device="cuda"):
        super(FrameReaderStd, self).__init__(camera_offset, 1.0 / camera_gain, 1.0, device)


class FrameReaderSCMOS(FrameReader):
    """
    An sCMOS camera (analysis_io.py:281-295): per-pixel offset, gain (ADU per
    photo-electron) and relative QE, as loadCMOSCalibration returns them.
    """
    def __init__(self, offset, gain, rqe, device="cuda"):
        super(FrameReaderSCMOS, self).__init__(offset, 1.0 / numpy.asarray(gain, dtype=numpy.float64),
                                               1.0 / numpy.asarray(rqe, dtype=numpy.float64), device)


def padArrays(frames, pad, device="cuda"):
    """padArray (fitting.py:112-132) of n float64 frames (backgrounds, variance maps) on the GPU."""
    shape = tuple(frames.shape)
    if len(shape) == 2:
        shape = (1,) + shape
    n, height, width = shape
    lib = loadIngestC()
    out = _device_out(n, height, width, pad, device)
    p, on, keep = _array(frames, numpy.float64)
    _check(lib.daoPadFrames(p, on, n, width, height, pad, ctypes.c_void_p(out.data_ptr())))
    return out


def readDaxFrames(filename, width, height, frames, big_endian=False):
    """
    Raw uint16 counts of `frames` (frame numbers) of a .dax movie, as
    DaxReader.loadAFrame reads them (datareader.py:234-243) but without the
    byte swap: pass big_endian to FrameReader.loadFrames, which swaps on the
    device.
    """
    out = numpy.empty((len(frames), height, width), dtype=numpy.uint16)
    with open(filename, "rb") as fp:
        for k, f in enumerate(frames):
            fp.seek(f * height * width * 2)
            out[k] = numpy.fromfile(fp, dtype=numpy.uint16, count=height * width).reshape(height, width)
    return out
