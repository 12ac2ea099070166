"""
GPU local-maxima finder (findLocalMaxima / calcMaxPeaks drop-ins,
daoFindLocalMaxima, and storm_analysis_amd.sa_library.dao_finder.MaximaFinder)
against the oracle (oracle/finder_oracle.py): the same peaks in the same
order, the same heights and the same taken counters.
"""
import ctypes
import os
import sys

import numpy
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import finder_oracle as FD  # noqa: E402
from test_finder_oracle import SETTINGS, check_reference_cases, flm_call, random_stack, run  # noqa: E402

from storm_analysis_amd.sa_library import dao_finder, loadclib  # noqa: E402


@pytest.mark.gpu
def test_reference_cases():
    check_reference_cases(lambda **kw: dao_finder.MaximaFinder(**kw))


@pytest.mark.gpu
@pytest.mark.parametrize("s", SETTINGS)
def test_drop_in_matches_oracle(s):
    """findLocalMaxima / calcMaxPeaks through the reference's flmData interface."""
    lib = loadclib.loadCLibrary("dao_fit")
    images = random_stack(s[0], s[1], s[2], s[3])
    t_gpu = [numpy.zeros(im.shape, dtype=numpy.int32) for im in images]
    t_or = [t.copy() for t in t_gpu]
    for _ in range(2):
        n_cand = FD.calc_max_peaks(images, t_or, s[4], s[7])
        got, n_gpu_cand = flm_call(lib, images, t_gpu, s[4], s[5], s[6], s[7], [0.5 * i for i in range(s[1])], s[8])
        want = run(FD.find_local_maxima, images, t_or, s)
        assert n_gpu_cand == n_cand
        for u, v in zip(got, want):
            numpy.testing.assert_array_equal(u, v)
        for u, v in zip(t_gpu, t_or):
            numpy.testing.assert_array_equal(u, v)


def blob_frame(seed, size=512, n=1000, sigma=1.5):
    """Gaussian spots on a noisy background (a filtered camera frame)."""
    rng = numpy.random.default_rng(seed)
    yy, xx = numpy.mgrid[0:size, 0:size]
    img = rng.normal(0.0, 1.0, size=(size, size))
    for y, x, h in zip(rng.uniform(0, size, n), rng.uniform(0, size, n), rng.uniform(5, 50, n)):
        y0, y1, x0, x1 = int(max(0, y - 6)), int(min(size, y + 7)), int(max(0, x - 6)), int(min(size, x + 7))
        img[y0:y1, x0:x1] += h * numpy.exp(-((yy[y0:y1, x0:x1] - y) ** 2 + (xx[y0:y1, x0:x1] - x) ** 2) /
                                           (2 * sigma * sigma))
    return numpy.round(img, 1)  # coarse values: ties between neighbours


@pytest.mark.gpu
def test_full_frame_and_device_input_match_oracle():
    import torch
    img = blob_frame(71)
    mxf = dao_finder.MaximaFinder(margin=10, radius=3.0, threshold=6.0, z_values=[0.0], n_duplicates=2)
    t_or = [numpy.full(img.shape, -1, dtype=numpy.int32)]
    for k in range(3):
        images = [torch.from_numpy(img).cuda()] if k == 1 else [img]
        x, y, z, h = mxf.findMaxima(images, want_height=True)
        n_max = FD.calc_max_peaks([img], t_or, 11, 6.0) + 1
        wz, wy, wx, wh = FD.find_local_maxima([img], t_or, 11, 1, 3.0, 6.0, [0.0], n_max)
        for u, v in ((x, wx), (y, wy), (z, wz), (h, wh)):
            numpy.testing.assert_array_equal(u, v)
        numpy.testing.assert_array_equal(mxf.taken[0].cpu().numpy(), t_or[0])
        assert (x.size > 200) if k < 2 else (x.size == 0)


@pytest.mark.gpu
def test_frames_as_planes_with_a_cut():
    """A batch of frames as planes with z_range 0 (each frame searched on its
    own), cut after 1000 peaks as the reference stops at n_peaks."""
    import torch
    frames = numpy.stack([blob_frame(80 + i, size=256, n=400) for i in range(6)])
    taken = torch.zeros(frames.shape, dtype=torch.int32, device="cuda")
    out = torch.empty((4, 1000), dtype=torch.float64, device="cuda")
    n = ctypes.c_int(0)
    zv = numpy.arange(6, dtype=numpy.float64)
    lib = dao_finder.loadFinderC()
    st = torch.from_numpy(frames).cuda()
    torch.cuda.synchronize()
    rc = lib.daoFindLocalMaxima(ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(taken.data_ptr()), 6, 256, 256, 5,
                                0, 2.5, 5.0, zv.ctypes.data_as(ctypes.c_void_p), 1000,
                                *[ctypes.c_void_p(out[i].data_ptr()) for i in range(4)], ctypes.byref(n))
    assert rc == 0 and n.value == 1000
    t_or = [numpy.zeros((256, 256), dtype=numpy.int32) for _ in range(6)]
    want = FD.find_local_maxima(list(frames), t_or, 5, 0, 2.5, 5.0, zv, 1000)
    got = out.cpu().numpy()
    for i in range(4):
        numpy.testing.assert_array_equal(got[i], want[i])
    numpy.testing.assert_array_equal(taken.cpu().numpy(), numpy.stack(t_or))


@pytest.mark.gpu
def test_bad_arguments_rejected():
    lib = dao_finder.loadFinderC()
    n = ctypes.c_int(0)
    zv = numpy.zeros(1)
    nul = ctypes.c_void_p(0)
    assert lib.daoFindLocalMaxima(nul, nul, 1, 8, 8, 1, 0, 2.0, 1.0, zv.ctypes.data_as(ctypes.c_void_p), 0,
                                  nul, nul, nul, nul, ctypes.byref(n)) == -1
    import torch
    a = torch.zeros((1, 8, 8), dtype=torch.float64, device="cuda")
    t = torch.zeros((1, 8, 8), dtype=torch.int32, device="cuda")
    assert lib.daoFindLocalMaxima(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(t.data_ptr()), 1, 8, 8, 1, 0,
                                  float("nan"), 1.0, zv.ctypes.data_as(ctypes.c_void_p), 0, nul, nul, nul, nul,
                                  ctypes.byref(n)) == -1
