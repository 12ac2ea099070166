"""
The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls: those need a GPU).
"""
import os
import re

import pytest

import conftest  # noqa: F401  (imports the package)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "storm-analysis_amd", "c_libraries", "libdao_fit.so")
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\**([A-Za-z_][A-Za-z0-9_]*)\s*\(", re.M)


def declared():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
        text = re.sub(r"typedef\s+struct[^;]*?\{.*?\}\s*\w+\s*;", "", text, flags=re.S)
        for m in DECL.finditer(text):
            if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
                names.append((fn, m.group(1)))
    return names


def test_headers_declare_the_reference_surface():
    names = {n for _, n in declared()}
    # the reference's ctypes binding (storm_analysis/sa_library/dao_fit_c.py:147-238)
    for n in ("daoInitialize", "daoInitialize2DFixed", "daoInitialize2DFixedALS", "daoInitialize2DFixedLS",
              "daoInitialize2DFixedDWLS", "daoInitialize2DFixedFWLS", "daoInitialize2D", "daoInitialize2DALS",
              "daoInitialize2DLS", "daoInitialize2DDWLS", "daoInitialize2DFWLS", "daoInitialize3D",
              "daoInitializeZ", "daoNewPeaks", "daoCleanup", "mFitNewImage", "mFitNewBackground",
              "mFitAnscombeTransformImage", "mFitIterateLM", "mFitGetUnconverged", "mFitGetNError",
              "mFitGetFitImage", "mFitGetResidual", "mFitGetPeakPropertyDouble", "mFitGetPeakPropertyInt",
              "mFitRemoveErrorPeaks", "mFitRemoveRunningPeaks", "mFitSetPeakStatus"):
        assert n in names, n
    # the multiplane binding (storm_analysis/multi_plane/mp_fit_c.py:55-127, mp_fit_dao_c.py:17-36)
    for n in ("mpCleanup", "mpInitialize", "mpIterateLM", "mpSetTransforms", "mpSetWeights",
              "mpSetWeightsIndexing", "mpDaoInitialize2DChannel", "mpDaoNewPeaks"):
        assert n in names, n


# headers whose functions live in a library of their own (as in the reference)
HEADER_LIBRARY = {"dao_mfilter_amd.h": "matched_filter", "mp_fit_amd.h": "mp_fit_dao"}


@pytest.mark.parametrize("header,name", declared())
def test_library_exports(header, name):
    libname = HEADER_LIBRARY.get(header, "dao_fit")
    path = os.path.join(ROOT, "storm-analysis_amd", "c_libraries", "lib%s.so" % libname)
    if not os.path.exists(path):
        pytest.fail("%s is not built (run __graft_entry__.build())" % path)
    from storm_analysis_amd.sa_library import loadclib
    lib = loadclib.loadCLibrary(libname)
    assert hasattr(lib, name), "%s declared in include/%s is not exported by lib%s.so" % (name, header, libname)
