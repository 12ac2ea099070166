"""
Pins of the post-fit filter oracle (oracle/filter_oracle.py), CPU only:

- the reference's own expectations for these functions
  (storm_analysis/test/test_ia_utilities.py: test_ia_util_9, 10, 13), restated
  as fixtures with the same inputs and expected statuses;
- the sequential restatement against the two-pass one on seeded inputs
  (the device kernels follow the two-pass form);
- both against the reference's compiled C (ia_utilities.c + dbscan/kdtree.c,
  built by oracle/Makefile into oracle/_ref/) where that library is present.
"""
import ctypes
import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import filter_oracle as FO  # noqa: E402

REF_SO = os.path.join(ROOT, "oracle", "_ref", "libia_utilities.so")
R, C, E = FO.RUNNING, FO.CONVERGED, FO.ERROR

SEQ = {"dimmer": FO.mark_dimmer_peaks_seq, "lowsig": FO.mark_low_significance_peaks_seq,
       "running": FO.running_if_has_neighbors_seq}
VEC = {"dimmer": FO.mark_dimmer_peaks, "lowsig": FO.mark_low_significance_peaks,
       "running": FO.running_if_has_neighbors}


# ---- the reference's test cases (test_ia_utilities.py:301-480) ----

RUNNING_CASES = [
    # c_x, c_y, n_x, n_y, status, radius, expected
    ([1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5, [1.1, 2.1, 3.1, 4.1], [1.1] * 4, [0, 1, 2, 1, 1], 0.5, [0, 0, 2, 0, 1]),
    ([1.0, 2.0, 3.0, 4.0, 5.0], [1.0] * 5, [1.9, 2.1], [1.1, 1.1], [0, 1, 2, 1, 1], 0.5, [0, 0, 2, 1, 1]),
    ([2.0, 3.0], [2.0, 2.0], [1.0], [2.0], [1, 1], 1.5, [0, 1]),
]

LOWSIG_CASES = [
    # min_sig, r_neighbors, expected removed, expected status; px 30/35/40, sig 1/2/3
    (0.0, 7.0, 0, [C, C, C]),
    (1.1, 3.0, 1, [E, C, C]),
    (1.1, 7.0, 1, [E, R, C]),
    (1.1, 11.0, 1, [E, R, R]),
    (2.1, 7.0, 2, [E, E, R]),
]


def running_case(case):
    c_x, c_y, n_x, n_y, status, radius, want = case
    a = [numpy.array(v, dtype=numpy.float64) for v in (c_x, c_y, n_x, n_y)]
    return a, numpy.array(status, dtype=numpy.int32), radius, want


def dimmer_case(seed):
    """test_ia_util_10: 25 peaks in the unit square, the first the tallest,
    the last moved to x = 4 (outside r_removal 2, inside r_neighbors 5), the
    second to last to x = 40 (away from everything)."""
    rng = numpy.random.default_rng(seed)
    n = 25
    x = rng.uniform(size=n)
    y = rng.uniform(size=n)
    h = rng.uniform(size=n) + 1.0
    h[0] = 4.0
    x[-1] = 4.0
    x[-2] = 40.0
    status = numpy.full(n, C, dtype=numpy.int32)
    return x, y, h, status


def check_dimmer_case(x, status, removed):
    n = x.size
    assert removed == n - 3
    assert all(status[i] == E for i in range(1, n - 2))
    assert status[0] == R and status[-1] == R and status[-2] == C


@pytest.mark.parametrize("impl", ["seq", "vec"])
@pytest.mark.parametrize("case", range(len(RUNNING_CASES)))
def test_running_if_has_neighbors_reference_cases(impl, case):
    (c_x, c_y, n_x, n_y), status, radius, want = running_case(RUNNING_CASES[case])
    fn = SEQ["running"] if impl == "seq" else VEC["running"]
    assert list(fn(status, c_x, c_y, n_x, n_y, radius)) == want


@pytest.mark.parametrize("impl", ["seq", "vec"])
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_mark_dimmer_peaks_reference_case(impl, seed):
    x, y, h, status = dimmer_case(seed)
    fn = SEQ["dimmer"] if impl == "seq" else VEC["dimmer"]
    check_dimmer_case(x, status, fn(x, y, h, status, 2.0, 5.0))


@pytest.mark.parametrize("impl", ["seq", "vec"])
@pytest.mark.parametrize("case", LOWSIG_CASES)
def test_mark_low_significance_reference_cases(impl, case):
    min_sig, r_nb, n_want, want = case
    px = numpy.array([30.0, 35.0, 40.0])
    py = numpy.array([30.0, 30.0, 30.0])
    sig = numpy.array([1.0, 2.0, 3.0])
    status = numpy.full(3, C, dtype=numpy.int32)
    fn = SEQ["lowsig"] if impl == "seq" else VEC["lowsig"]
    assert fn(px, py, sig, status, min_sig, r_nb) == n_want
    assert list(status) == want


# ---- closed forms that a plausible mistake breaks ----

def test_distance_is_inclusive_and_error_peaks_still_outshine():
    """Peaks on a line 1 apart with r_removal 1: the test is <= (an exclusive
    test removes nothing); an ERROR peak is skipped but still counts as the
    brighter neighbour; equal heights do not remove each other."""
    x = numpy.array([0.0, 1.0, 2.0, 10.0, 11.0])
    y = numpy.zeros(5)
    h = numpy.array([5.0, 4.0, 4.0, 3.0, 3.0])
    for impl in ("seq", "vec"):
        status = numpy.array([E, C, C, C, C], dtype=numpy.int32)
        n = (SEQ if impl == "seq" else VEC)["dimmer"](x, y, h, status, 1.0, 1.0)
        # peak 1 has the (ERROR) peak 0 as brighter neighbour; peak 2's only
        # neighbour within 1 is peak 1, of equal height; peaks 3, 4 tie.
        assert n == 1
        assert list(status) == [E, E, R, C, C]


def test_running_peaks_stay_running_and_error_peaks_do_not_mark():
    x = numpy.array([0.0, 0.5, 1.0, 5.0])
    y = numpy.zeros(4)
    sig = numpy.array([0.0, 9.0, 9.0, 0.0])
    for impl in ("seq", "vec"):
        status = numpy.array([C, R, C, E], dtype=numpy.int32)
        n = (SEQ if impl == "seq" else VEC)["lowsig"](x, y, sig, status, 1.0, 1.0)
        # peak 0 removed (sig 0 <= 1) and marks peak 2 (distance 1, inclusive);
        # peak 3 is already ERROR: not counted, marks nobody.
        assert n == 1
        assert list(status) == [E, R, R, E]


# ---- sequential == two-pass on seeded inputs ----

def random_peaks(seed, n, extent, n_status=True):
    rng = numpy.random.default_rng(seed)
    x = rng.uniform(0, extent, size=n)
    y = rng.uniform(0, extent, size=n)
    h = rng.uniform(1.0, 2.0, size=n)
    h[rng.integers(0, n, size=n // 10)] = 1.5  # some equal heights
    sig = rng.uniform(0.0, 10.0, size=n)
    status = rng.choice([R, C, C, C, E], size=n).astype(numpy.int32) if n_status else \
        numpy.full(n, C, dtype=numpy.int32)
    return x, y, h, sig, status


@pytest.mark.parametrize("seed,n,extent", [(1, 60, 10.0), (2, 150, 20.0), (3, 200, 8.0), (4, 1, 1.0)])
def test_sequential_equals_two_pass(seed, n, extent):
    x, y, h, sig, status = random_peaks(seed, n, extent)
    for name, va, ra, rb in (("dimmer", h, 1.5, 4.0), ("lowsig", sig, 3.0, 2.5)):
        s1, s2 = status.copy(), status.copy()
        assert SEQ[name](x, y, va, s1, ra, rb) == VEC[name](x, y, va, s2, ra, rb)
        numpy.testing.assert_array_equal(s1, s2)
    nx, ny = x[: n // 3] + 0.7, y[: n // 3] - 0.4
    s1, s2 = status.copy(), status.copy()
    SEQ["running"](s1, x, y, nx, ny, 1.2)
    VEC["running"](s2, x, y, nx, ny, 1.2)
    numpy.testing.assert_array_equal(s1, s2)


# ---- against the reference's own C where it was built ----

def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libia_utilities.so not built (needs the reference's sources)")
    lib = ctypes.cdll.LoadLibrary(REF_SO)
    P = numpy.ctypeslib.ndpointer
    d, i32 = P(dtype=numpy.float64), P(dtype=numpy.int32)
    for name in ("markDimmerPeaks", "markLowSignificancePeaks"):
        getattr(lib, name).argtypes = [d, d, d, i32, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        getattr(lib, name).restype = ctypes.c_int
    lib.runningIfHasNeighbors.argtypes = [d, d, d, d, i32, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    lib.runningIfHasNeighbors.restype = None
    return lib


@pytest.mark.parametrize("seed,n,extent", [(11, 40, 6.0), (12, 500, 40.0), (13, 2000, 60.0), (14, 3000, 512.0)])
def test_oracle_matches_reference_c(seed, n, extent):
    lib = ref_lib()
    x, y, h, sig, status = random_peaks(seed, n, extent)
    for name, cname, va, ra, rb in (("dimmer", "markDimmerPeaks", h, 1.5, 7.5),
                                    ("lowsig", "markLowSignificancePeaks", sig, 2.0, 7.5)):
        s_ref, s_or = status.copy(), status.copy()
        n_ref = getattr(lib, cname)(x, y, va, s_ref, ra, rb, n)
        n_or = VEC[name](x, y, va, s_or, ra, rb)
        assert n_ref == n_or
        numpy.testing.assert_array_equal(s_ref, s_or)
    nx, ny = x[: n // 4] + 0.9, y[: n // 4] + 0.3
    s_ref, s_or = status.copy(), status.copy()
    lib.runningIfHasNeighbors(x, y, nx, ny, s_ref, 1.5, n, nx.size)
    VEC["running"](s_or, x, y, nx, ny, 1.5)
    numpy.testing.assert_array_equal(s_ref, s_or)


def test_reference_c_reproduces_its_test_cases():
    """The built reference agrees with the expectations the fixtures above hold."""
    lib = ref_lib()
    for case in RUNNING_CASES:
        (c_x, c_y, n_x, n_y), status, radius, want = running_case(case)
        lib.runningIfHasNeighbors(c_x, c_y, n_x, n_y, status, radius, c_x.size, n_x.size)
        assert list(status) == want
    x, y, h, status = dimmer_case(5)
    check_dimmer_case(x, status, lib.markDimmerPeaks(x, y, h, status, 2.0, 5.0, x.size))
