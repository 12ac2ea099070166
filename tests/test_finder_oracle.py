"""
Pins of the local-maxima finder oracle (oracle/finder_oracle.py), CPU only:

- the reference's own expectations for MaximaFinder
  (storm_analysis/test/test_ia_utilities.py: test_ia_util_1, 2, 4-8),
  restated as fixtures with the same images and settings;
- the sequential restatement against the vectorised one on seeded images
  with ties, several planes, z ranges, radii and a peak-count cut;
- both against the reference's compiled C (ia_utilities.c, built by
  oracle/Makefile into oracle/_ref/) where that library is present.
"""
import ctypes
import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import finder_oracle as FD  # noqa: E402
from finder_oracle import flm_call  # noqa: E402,F401  (also used by the GPU tests)

REF_SO = os.path.join(ROOT, "oracle", "_ref", "libia_utilities.so")


class OracleFinder(object):
    """MaximaFinder's bookkeeping (ia_utilities_c.py:136-252) around an oracle search."""
    def __init__(self, search, margin, radius, threshold, z_values, n_duplicates=1, z_range=None):
        self.search = search
        self.margin = int(margin + 1)
        self.radius = float(radius)
        self.threshold = float(threshold)
        self.z_values = list(z_values)
        self.z_range = len(z_values) if z_range is None else int(z_range)
        self.n_duplicates = n_duplicates
        self.taken = None

    def findMaxima(self, images, want_height=False):
        if self.taken is None:
            self.taken = [numpy.ones(images[0].shape, dtype=numpy.int32) - self.n_duplicates for _ in images]
        n_max = FD.calc_max_peaks(images, self.taken, self.margin, self.threshold) + 1
        z, y, x, h = (numpy.asarray(v, dtype=numpy.float64) for v in
                      self.search(images, self.taken, self.margin, self.z_range, self.radius, self.threshold,
                                  self.z_values, n_max))
        return [x, y, z, h] if want_height else [x, y, z]

    def resetTaken(self):
        if self.taken is not None:
            for t in self.taken:
                t[:, :] = 1 - self.n_duplicates


def blank(n=1):
    return [numpy.zeros((100, 80), dtype=numpy.float64) for _ in range(n)]


def case_images(name):
    ims = blank(3 if name == "4" else 1)
    if name == "2":
        ims[0][10, 11], ims[0][10, 12] = 1.1, 1.5
        ims[0][15, 8], ims[0][15, 9] = 1.5, 1.5
        ims[0][20, 11] = 0.5
    elif name == "4":
        ims[0][20, 10], ims[1][20, 10], ims[2][20, 10] = 1.3, 1.2, 1.5
    elif name in ("5", "8"):
        ims[0][10, 21], ims[0][10, 20] = 1.1, 1.2
    elif name == "6":
        ims[0][10, 23], ims[0][10, 22], ims[0][10, 21], ims[0][10, 20] = 1.2, 1.1, 1.1, 1.2
    elif name == "7":
        ims[0][10, 5], ims[0][10, 6], ims[0][10, 7] = 1.2, 1.1, 1.1
        ims[0][20, 10], ims[0][20, 11] = 1.1, 1.2
    return ims


def check_reference_cases(make):
    """The reference's test_ia_util_1, 2, 4-8 with finders from make(**kwargs)."""
    z1 = [0.1]
    f = make(margin=1, radius=2, threshold=1, z_values=z1)
    assert f.findMaxima(case_images("1"))[0].size == 0
    x, y, z, h = make(margin=1, radius=2, threshold=1, z_values=z1).findMaxima(case_images("2"), want_height=True)
    assert x.size == 2 and numpy.allclose(z, 0.1) and numpy.allclose(h, 1.5)
    zv = [1.0, 2.0, 3.0]
    x, y, z = make(margin=1, radius=2, threshold=1, z_values=zv).findMaxima(case_images("4"))
    assert x.size == 1 and abs(z[0] - 3.0) < 1e-6
    x, y, z = make(margin=1, radius=2, threshold=1, z_range=1, z_values=zv).findMaxima(case_images("4"))
    assert x.size == 2 and abs(z[0] - 1.0) < 1e-6 and abs(z[1] - 3.0) < 1e-6
    x, y, z = make(margin=1, radius=2, threshold=1, z_range=0, z_values=zv).findMaxima(case_images("4"))
    assert x.size == 3 and numpy.allclose(z, zv)
    f = make(margin=1, radius=2, threshold=1, z_values=z1)
    assert [f.findMaxima(case_images("5"))[0].size for _ in range(2)] == [1, 0]
    f.resetTaken()
    assert f.findMaxima(case_images("5"))[0].size == 1
    assert make(margin=1, radius=3, threshold=1, z_values=z1).findMaxima(case_images("6"))[0].size == 1
    assert make(margin=1, radius=2, threshold=1, z_values=z1).findMaxima(case_images("6"))[0].size == 2
    assert make(margin=6, radius=3, threshold=1, z_values=z1).findMaxima(case_images("7"))[0].size == 1
    assert make(margin=4, radius=3, threshold=1, z_values=z1).findMaxima(case_images("7"))[0].size == 2
    f = make(margin=1, n_duplicates=2, radius=2, threshold=1, z_values=z1)
    assert [f.findMaxima(case_images("8"))[0].size for _ in range(3)] == [1, 1, 0]
    f.resetTaken()
    assert [f.findMaxima(case_images("8"))[0].size for _ in range(3)] == [1, 1, 0]


@pytest.mark.parametrize("search", [FD.find_local_maxima_seq, FD.find_local_maxima])
def test_reference_cases(search):
    check_reference_cases(lambda **kw: OracleFinder(search, **kw))


def random_stack(seed, zs, ys, xs, levels=6):
    """Quantised noise plus blobs: many equal-height neighbours (ties)."""
    rng = numpy.random.default_rng(seed)
    st = rng.integers(0, levels, size=(zs, ys, xs)).astype(numpy.float64)
    for _ in range(ys * xs // 60):
        z, y, x = rng.integers(0, zs), rng.integers(0, ys), rng.integers(0, xs)
        st[z, max(0, y - 1):y + 2, max(0, x - 1):x + 2] += rng.integers(1, 4)
    return [st[i].copy() for i in range(zs)]


SETTINGS = [
    # seed, zs, ys, xs, margin, z_range, radius, threshold, n_peaks
    (1, 1, 24, 31, 2, 1, 2.0, 2.5, 10 ** 6),
    (2, 1, 24, 31, 1, 1, 2.5, 1.5, 10 ** 6),
    (3, 3, 20, 22, 3, 1, 1.5, 2.5, 10 ** 6),
    (4, 3, 20, 22, 2, 0, 3.0, 2.0, 10 ** 6),
    (5, 2, 26, 18, 2, 5, 0.5, 3.5, 10 ** 6),
    (6, 1, 30, 30, 2, 1, 2.0, 1.5, 7),      # cut after 7 peaks
    (7, 2, 16, 16, 1, -1, -1.5, 2.5, 10 ** 6),  # negative radius and z range: empty windows
]


def run(search, images, taken, s):
    seed, zs, ys, xs, margin, z_range, radius, thr, n_peaks = s
    return search(images, taken, margin, z_range, radius, thr, [0.5 * i for i in range(zs)], n_peaks)


@pytest.mark.parametrize("s", SETTINGS)
def test_sequential_equals_vectorised(s):
    images = random_stack(s[0], s[1], s[2], s[3])
    t1 = [numpy.zeros(im.shape, dtype=numpy.int32) for im in images]
    t2 = [t.copy() for t in t1]
    t1[0][5, 5] = t2[0][5, 5] = 1  # one pixel already taken
    found = []
    for _ in range(2):  # the second call sees the first's taken counters
        a = run(FD.find_local_maxima_seq, images, t1, s)
        b = run(FD.find_local_maxima, images, t2, s)
        for u, v in zip(a, b):
            numpy.testing.assert_array_equal(numpy.asarray(u, dtype=numpy.float64), v)
        for u, v in zip(t1, t2):
            numpy.testing.assert_array_equal(u, v)
        found.append(len(a[0]))
    assert found[0] > 0


# ---- against the reference's own C where it was built ----

def ref_lib():
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libia_utilities.so not built (needs the reference's sources)")
    return ctypes.cdll.LoadLibrary(REF_SO)


@pytest.mark.parametrize("s", SETTINGS + [(8, 1, 128, 128, 4, 1, 2.0, 3.5, 10 ** 6)])
def test_oracle_matches_reference_c(s):
    lib = ref_lib()
    images = random_stack(s[0], s[1], s[2], s[3])
    t_ref = [numpy.zeros(im.shape, dtype=numpy.int32) for im in images]
    t_or = [t.copy() for t in t_ref]
    for _ in range(2):
        n_cand = FD.calc_max_peaks(images, t_or, s[4], s[7])
        got, n_ref_cand = flm_call(lib, images, t_ref, s[4], s[5], s[6], s[7], [0.5 * i for i in range(s[1])],
                                   s[8])
        want = run(FD.find_local_maxima, images, t_or, s)
        assert n_ref_cand == n_cand
        for u, v in zip(got, want):
            numpy.testing.assert_array_equal(u, v)
        for u, v in zip(t_ref, t_or):
            numpy.testing.assert_array_equal(u, v)
