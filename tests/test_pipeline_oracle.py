"""
The oracle of the analysis loop (oracle/pipeline_oracle.py: find, fit,
filter, subtract, frame by frame, built from the other oracles) against the
reference's own end-to-end numbers: test/data/test.dax analysed with
test_3d_{2d_fixed,2d,3d,Z}.xml finds 1998 / 1970 / 1956 / 1955
localizations, within 5% (storm_analysis/test/test_3ddao.py:12-141 with
verifications.py:26-41).  The movie and the settings are the reference's
test data (tests/golden/dax/).  This pins the oracle's fitter, finder,
filters, matched filter and ingest together, against numbers that none of
them produced.
"""
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import conftest  # noqa: E402,F401  (imports the package)
import pipeline_oracle  # noqa: E402

DAX = os.path.join(HERE, "golden", "dax", "test.dax")
EXPECTED = {"2d_fixed": 1998, "2d": 1970, "3d": 1956, "Z": 1955}


def close_enough(n, want, margin=0.05):
    """verifyIsCloseEnough (verifications.py:26-31)."""
    return abs(n - want) < want * margin


@pytest.mark.parametrize("model", sorted(EXPECTED))
def test_oracle_analysis_of_test_dax_finds_the_reference_count(model):
    xml = os.path.join(HERE, "golden", "dax", "test_3d_%s.xml" % model)
    peaks = pipeline_oracle.analyze_dax(DAX, xml, 256, 256, 10)
    n = sum(p["x"].size for p in peaks)
    print("%s: %d localizations (reference %d)" % (model, n, EXPECTED[model]))
    assert close_enough(n, EXPECTED[model])
