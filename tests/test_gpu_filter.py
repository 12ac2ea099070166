"""
GPU post-fit filters (markDimmerPeaks, markLowSignificancePeaks,
runningIfHasNeighbors and their batched forms, through
storm_analysis_amd.sa_library.dao_filter) against the oracle
(oracle/filter_oracle.py): statuses and removed counts identical.
"""
import os
import sys

import numpy
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import filter_oracle as FO  # noqa: E402
from test_filter_oracle import (LOWSIG_CASES, RUNNING_CASES, check_dimmer_case, dimmer_case,  # noqa: E402
                                random_peaks, running_case)

from storm_analysis_amd.sa_library import dao_filter as DF  # noqa: E402

R, C, E = FO.RUNNING, FO.CONVERGED, FO.ERROR


# ---- the reference's test cases through the drop-ins ----

@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(RUNNING_CASES)))
def test_running_if_has_neighbors_reference_cases(case):
    (c_x, c_y, n_x, n_y), status, radius, want = running_case(RUNNING_CASES[case])
    assert list(DF.runningIfHasNeighbors(status, c_x, c_y, n_x, n_y, radius)) == want


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_mark_dimmer_peaks_reference_case(seed):
    x, y, h, status = dimmer_case(seed)
    check_dimmer_case(x, status, DF.markDimmerPeaks(x, y, h, status, 2.0, 5.0))


@pytest.mark.gpu
@pytest.mark.parametrize("case", LOWSIG_CASES)
def test_mark_low_significance_reference_cases(case):
    min_sig, r_nb, n_want, want = case
    px = numpy.array([30.0, 35.0, 40.0])
    py = numpy.array([30.0, 30.0, 30.0])
    sig = numpy.array([1.0, 2.0, 3.0])
    status = numpy.full(3, C, dtype=numpy.int32)
    assert DF.markLowSignificancePeaks(px, py, sig, status, min_sig, r_nb) == n_want
    assert list(status) == want


# ---- seeded single lists against the oracle ----

@pytest.mark.gpu
@pytest.mark.parametrize("seed,n,extent,ra,rb", [
    (21, 0, 10.0, 1.5, 7.5),
    (22, 1, 10.0, 1.5, 7.5),
    (23, 300, 30.0, 1.5, 7.5),      # crowded
    (24, 1000, 512.0, 1.5, 7.5),    # one frame of the 512 x 512, 1000-emitter workload
    (25, 4000, 2048.0, 1.0, 5.0),   # wider than the 64 x 64 grid cap at this radius
    (26, 500, 50.0, float("inf"), 3.0),   # infinite radius: brute-force search
    (27, 500, 50.0, 0.0, 3.0),      # zero removal radius: coincident peaks only
])
def test_single_list_matches_oracle(seed, n, extent, ra, rb):
    x, y, h, sig, status = random_peaks(seed, max(n, 1), extent)
    x, y, h, sig, status = x[:n].copy(), y[:n].copy(), h[:n].copy(), sig[:n].copy(), status[:n].copy()
    if n >= 10:
        x[3], y[3] = x[4], y[4]  # a coincident pair
    for gpu_fn, or_fn, v in ((DF.markDimmerPeaks, FO.mark_dimmer_peaks, h),
                             (DF.markLowSignificancePeaks, FO.mark_low_significance_peaks, sig)):
        thr = ra if gpu_fn is DF.markDimmerPeaks else 4.0
        s_g, s_o = status.copy(), status.copy()
        n_g = gpu_fn(x, y, v, s_g, thr, rb)
        n_o = or_fn(x, y, v, s_o, thr, rb)
        assert n_g == n_o, gpu_fn.__name__
        numpy.testing.assert_array_equal(s_g, s_o, err_msg=gpu_fn.__name__)
    k = n // 3
    nx, ny = (x[:k] + 0.8).copy(), (y[:k] - 0.5).copy()
    s_g, s_o = status.copy(), status.copy()
    DF.runningIfHasNeighbors(s_g, x, y, nx, ny, rb)
    FO.running_if_has_neighbors(s_o, x, y, nx, ny, rb)
    numpy.testing.assert_array_equal(s_g, s_o)


@pytest.mark.gpu
def test_exact_distance_is_inclusive_on_an_integer_lattice():
    """Peaks on a unit lattice with radii 1 and sqrt(2)-free integers: every
    distance test lands on an exact tie somewhere, and the device keeps the
    kd-tree's inclusive test."""
    gx, gy = numpy.meshgrid(numpy.arange(12.0), numpy.arange(9.0))
    x, y = gx.ravel().copy(), gy.ravel().copy()
    rng = numpy.random.default_rng(31)
    h = rng.integers(1, 4, size=x.size).astype(numpy.float64)
    status = rng.choice([R, C, C, E], size=x.size).astype(numpy.int32)
    for ra, rb in ((1.0, 1.0), (1.0, 2.0), (2.0, 3.0)):
        s_g, s_o = status.copy(), status.copy()
        assert DF.markDimmerPeaks(x, y, h, s_g, ra, rb) == FO.mark_dimmer_peaks(x, y, h, s_o, ra, rb)
        numpy.testing.assert_array_equal(s_g, s_o)


# ---- batches of frames ----

def ragged_batch(seed, nframes, mean, extent):
    rng = numpy.random.default_rng(seed)
    counts = rng.integers(0, 2 * mean + 1, size=nframes)
    counts[:3] = [0, 1, 0]  # empty and single-peak frames
    off = numpy.concatenate([[0], numpy.cumsum(counts)]).astype(numpy.int32)
    x, y, h, sig, status = random_peaks(seed + 1, int(off[-1]), extent)
    return off, x, y, h, sig, status


def cuda(a):
    import torch
    return torch.from_numpy(numpy.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("seed,nframes,mean,extent", [(41, 64, 200, 128.0), (42, 300, 1000, 512.0)])
def test_batched_marks_match_oracle_per_frame(seed, nframes, mean, extent):
    off, x, y, h, sig, status = ragged_batch(seed, nframes, mean, extent)
    for batch_fn, or_fn, v, ra in ((DF.markDimmerPeaksBatch, FO.mark_dimmer_peaks, h, 1.5),
                                   (DF.markLowSignificancePeaksBatch, FO.mark_low_significance_peaks, sig, 3.0)):
        st = cuda(status)
        removed = batch_fn(cuda(x), cuda(y), cuda(v), st, off, ra, 7.5).cpu().numpy()
        st = st.cpu().numpy()
        for f in range(nframes):
            a, b = off[f], off[f + 1]
            s_o = status[a:b].copy()
            n_o = or_fn(x[a:b], y[a:b], v[a:b], s_o, ra, 7.5)
            assert removed[f] == n_o, (batch_fn.__name__, f)
            numpy.testing.assert_array_equal(st[a:b], s_o, err_msg="%s frame %d" % (batch_fn.__name__, f))


@pytest.mark.gpu
def test_batched_running_if_has_neighbors_matches_oracle_per_frame():
    off, x, y, h, sig, status = ragged_batch(51, 100, 300, 256.0)
    rng = numpy.random.default_rng(52)
    ncounts = rng.integers(0, 40, size=100)
    ncounts[5] = 0
    noff = numpy.concatenate([[0], numpy.cumsum(ncounts)]).astype(numpy.int32)
    nx = rng.uniform(0, 256.0, size=noff[-1])
    ny = rng.uniform(0, 256.0, size=noff[-1])
    st = cuda(status)
    DF.runningIfHasNeighborsBatch(st, cuda(x), cuda(y), off, cuda(nx), cuda(ny), noff, 7.5)
    st = st.cpu().numpy()
    for f in range(100):
        a, b = off[f], off[f + 1]
        s_o = status[a:b].copy()
        FO.running_if_has_neighbors(s_o, x[a:b], y[a:b], nx[noff[f]:noff[f + 1]], ny[noff[f]:noff[f + 1]], 7.5)
        numpy.testing.assert_array_equal(st[a:b], s_o, err_msg="frame %d" % f)


@pytest.mark.gpu
def test_batched_filters_are_idempotent_at_full_size():
    """500 frames x 1000 peaks in 512 x 512 (the fitting workload's peak
    lists): a second pass removes nothing more and changes no status, and the
    removed counts add up to the new ERROR peaks."""
    import torch
    rng = numpy.random.default_rng(61)
    nframes, per = 500, 1000
    off = (numpy.arange(nframes + 1) * per).astype(numpy.int32)
    x = cuda(rng.uniform(0, 512.0, size=nframes * per))
    y = cuda(rng.uniform(0, 512.0, size=nframes * per))
    h = cuda(rng.uniform(50.0, 500.0, size=nframes * per))
    st = torch.full((nframes * per,), C, dtype=torch.int32, device="cuda")
    removed = DF.markDimmerPeaksBatch(x, y, h, st, off, 1.5, 7.5)
    n_err = int((st == E).sum())
    assert int(removed.sum()) == n_err > 0
    before = st.clone()
    again = DF.markDimmerPeaksBatch(x, y, h, st, off, 1.5, 7.5)
    assert int(again.sum()) == 0
    assert torch.equal(st, before)
    # spot-check three frames against the oracle
    xs, ys, hs, ss = (t.cpu().numpy() for t in (x, y, h, before))
    for f in (0, 257, 499):
        a, b = off[f], off[f + 1]
        s_o = numpy.full(per, C, dtype=numpy.int32)
        FO.mark_dimmer_peaks(xs[a:b], ys[a:b], hs[a:b], s_o, 1.5, 7.5)
        numpy.testing.assert_array_equal(ss[a:b], s_o)


@pytest.mark.gpu
def test_bad_offsets_raise():
    from storm_analysis_amd.sa_library.dao_fit_c import MultiFitterException
    x = cuda(numpy.zeros(4))
    st = cuda(numpy.ones(4, dtype=numpy.int32))
    with pytest.raises(MultiFitterException):
        DF.markDimmerPeaksBatch(x, x, x, st, [0, 3], 1.0, 1.0)
    with pytest.raises(MultiFitterException):
        DF.markDimmerPeaksBatch(x, x, x, st, [0, 3, 2, 4], 1.0, 1.0)
