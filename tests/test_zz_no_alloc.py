"""
The allocation-free local-maxima search (daoFindLocalMaximaNoAlloc) against
the allocating one (daoFindLocalMaxima), both through the C-ABI on the same
device arrays: identical peaks, heights, taken counters and counts, on two
calls in a row -- the second with a smaller input, which finds the scratch
still holding the first call's mask and row counts -- and on the caller's
stream.  Inputs beyond the fixed scratch are refused before any device work.
"""
import ctypes

import numpy
import pytest

from storm_analysis_amd.sa_library import dao_finder


def _lib():
    lib = dao_finder.loadFinderC()
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.daoFindLocalMaximaNoAlloc.argtypes = [vp, vp, ci, ci, ci, ci, ci, cd, cd, vp, ci, vp, vp, vp, vp, vp, vp]
    lib.daoFindLocalMaximaNoAlloc.restype = ci
    return lib


def _stack(seed, planes, ys, xs, n):
    """Gaussian spots on noise, with many exact ties (values rounded)."""
    rng = numpy.random.default_rng(seed)
    yy, xx = numpy.mgrid[0:ys, 0:xs]
    out = numpy.empty((planes, ys, xs))
    for p in range(planes):
        img = rng.normal(0.0, 1.0, size=(ys, xs))
        for y, x, h in zip(rng.uniform(0, ys, n), rng.uniform(0, xs, n), rng.uniform(5, 50, n)):
            img += h * numpy.exp(-((yy - y) ** 2 + (xx - x) ** 2) / (2 * 1.5 ** 2))
        out[p] = numpy.round(img, 1)
    return out


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _run_both(lib, images, max_peaks, stream):
    import torch
    zs, ys, xs = images.shape
    dev = torch.from_numpy(images).cuda()
    z_host = numpy.arange(zs, dtype=numpy.float64) * 0.5
    z_dev = torch.from_numpy(z_host).cuda()
    taken_a = torch.zeros((zs, ys, xs), dtype=torch.int32, device="cuda")
    taken_b = taken_a.clone()
    m = max(max_peaks, 1)
    out_a = [torch.full((m,), -1.0, dtype=torch.float64, device="cuda") for _ in range(4)]
    out_b = [torch.full((m,), -1.0, dtype=torch.float64, device="cuda") for _ in range(4)]
    counts = torch.full((2,), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    n_found = ctypes.c_int(0)
    rc = lib.daoFindLocalMaxima(_ptr(dev), _ptr(taken_a), zs, ys, xs, 6, 0, 3.0, 6.0,
                                z_host.ctypes.data_as(ctypes.c_void_p), max_peaks,
                                *[_ptr(t) for t in out_a], ctypes.byref(n_found))
    assert rc == 0
    with torch.cuda.stream(stream):
        rc = lib.daoFindLocalMaximaNoAlloc(_ptr(dev), _ptr(taken_b), zs, ys, xs, 6, 0, 3.0, 6.0, _ptr(z_dev),
                                           max_peaks, *[_ptr(t) for t in out_b], _ptr(counts),
                                           ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0
    stream.synchronize()
    c = counts.cpu().numpy()
    # n_found as the allocating search reports it; the maxima count before
    # the cut is at least that, and equal to it when nothing was cut
    assert c[0] == n_found.value, (c, n_found.value)
    assert c[1] >= c[0] and (c[0] == max_peaks or c[1] == c[0]), c
    total = int(c[1])
    for a, b in zip(out_a, out_b):
        numpy.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    numpy.testing.assert_array_equal(taken_a.cpu().numpy(), taken_b.cpu().numpy())
    return n_found.value, total


@pytest.mark.gpu
def test_no_alloc_finder_matches_the_allocating_one():
    import torch
    lib = _lib()
    stream = torch.cuda.Stream()
    # first call: 6 planes of 512 x 384 (1.2 M pixels), cut at 3000 peaks
    n1, t1 = _run_both(lib, _stack(1, 6, 512, 384, 900), 3000, stream)
    assert n1 == 3000 and t1 > 3000
    # second call, smaller: 2 planes of 96 x 130, no cut; the scratch still
    # holds the first call's mask and counts
    n2, t2 = _run_both(lib, _stack(2, 2, 96, 130, 40), 10000, stream)
    assert 0 < n2 == t2 < 10000
    # and once more on the null stream
    _run_both(lib, _stack(3, 1, 64, 64, 10), 5, torch.cuda.default_stream())


@pytest.mark.gpu
def test_no_alloc_finder_refuses_inputs_beyond_its_scratch():
    import torch
    lib = _lib()
    t = torch.zeros(16, dtype=torch.float64, device="cuda")
    ti = torch.zeros(16, dtype=torch.int32, device="cuda")
    p = _ptr(t)
    for zs, ys, xs in ((1, 4097, 4096), (5000, 4, 4), (2, 40000, 8)):   # pixels, planes, rows
        rc = lib.daoFindLocalMaximaNoAlloc(p, _ptr(ti), zs, ys, xs, 6, 0, 3.0, 6.0, p, 4, p, p, p, p, _ptr(ti),
                                           None)
        assert rc == -1, (zs, ys, xs)
