"""
The device-resident analysis loop (sa_library/dao_analysis.py: ingest,
background filter, search image, maxima, fit, post-fit filters and fit
image, all frames of a batch at once on the GPU) against:

- the reference's own numbers on its test movie: test/data/test.dax with
  test_3d_{2d_fixed,2d,3d,Z}.xml gives 1998 / 1970 / 1956 / 1955
  localizations, within 5% (storm_analysis/test/test_3ddao.py:12-141);
- the CPU oracle of the same loop (oracle/pipeline_oracle.py), frame by
  frame: the same number of localizations in every frame, and positions,
  heights and backgrounds of the matched peaks within 1e-3 px / 1e-4.
"""
import os
import sys

import numpy
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import conftest  # noqa: E402,F401
import pipeline_oracle  # noqa: E402
from storm_analysis_amd.sa_library import dao_analysis  # noqa: E402

DAX = os.path.join(HERE, "golden", "dax", "test.dax")
EXPECTED = {"2d_fixed": 1998, "2d": 1970, "3d": 1956, "Z": 1955}


def match(a, b, tol=0.5):
    """Pairs (i, j) of peaks of a and b within tol px, nearest first."""
    d = numpy.hypot(a["x"][:, None] - b["x"][None, :], a["y"][:, None] - b["y"][None, :])
    pairs = []
    if d.size:
        j = numpy.argmin(d, axis=1)
        for i in range(d.shape[0]):
            if d[i, j[i]] < tol:
                pairs.append((i, j[i]))
    return pairs


@pytest.mark.gpu
@pytest.mark.parametrize("model", sorted(EXPECTED))
def test_device_analysis_of_test_dax(model):
    xml = os.path.join(HERE, "golden", "dax", "test_3d_%s.xml" % model)
    got = dao_analysis.analyzeMovie(DAX, xml)
    want = pipeline_oracle.analyze_dax(DAX, xml, 256, 256, 10)
    n = sum(p["x"].size for p in got)
    print("%s: %d localizations on the GPU, %d by the oracle, reference %d" %
          (model, n, sum(p["x"].size for p in want), EXPECTED[model]))
    assert abs(n - EXPECTED[model]) < 0.05 * EXPECTED[model]
    for f, (g, r) in enumerate(zip(got, want)):
        assert g["x"].size == r["x"].size, (f, g["x"].size, r["x"].size)
        pairs = match(g, r)
        assert len(pairs) == r["x"].size, (f, len(pairs), r["x"].size)
        gi = numpy.array([p[0] for p in pairs])
        ri = numpy.array([p[1] for p in pairs])
        for k in ("x", "y"):
            assert numpy.max(numpy.abs(g[k][gi] - r[k][ri])) < 1e-3, (f, k)
        for k in ("height", "background"):
            assert numpy.max(numpy.abs(g[k][gi] - r[k][ri]) / numpy.abs(r[k][ri])) < 1e-4, (f, k)
