"""
Shared test setup: registers the `gpu` marker, imports the package from
storm-analysis_amd/, and provides the two back ends of the same ctypes
surface -- the HIP library (the product) and the oracle (the checker).
"""
import ctypes
import os
import sys
from unittest import mock

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

__graft_entry__.load_package()
from storm_analysis_amd.sa_library import dao_fit_c  # noqa: E402

# DAO_ORACLE_SO points the oracle back end at another build of the oracle
# (test_oracle_mutations.py runs the pins against deliberately broken copies)
ORACLE_SO = os.environ.get("DAO_ORACLE_SO", os.path.join(ROOT, "oracle", "liboracle_dao.so"))
HIP_SO = os.path.join(ROOT, "storm-analysis_amd", "c_libraries", "libdao_fit.so")
# the multiplane oracle (DAO_MP_ORACLE_SO: another build of it, as above)
MP_ORACLE_SO = os.environ.get("DAO_MP_ORACLE_SO", os.path.join(ROOT, "oracle", "liboracle_mp.so"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: make -C oracle")
        lib = dao_fit_c.setDaoFitArgtypes(ctypes.cdll.LoadLibrary(ORACLE_SO))
        lib.oraPeakErrorJH.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                       numpy.ctypeslib.ndpointer(dtype=numpy.float64),
                                       numpy.ctypeslib.ndpointer(dtype=numpy.float64),
                                       numpy.ctypeslib.ndpointer(dtype=numpy.float64)]
        lib.oraPeakErrorJH.restype = ctypes.c_int
        lib.oraSolve.argtypes = [numpy.ctypeslib.ndpointer(dtype=numpy.float64),
                                 numpy.ctypeslib.ndpointer(dtype=numpy.float64), ctypes.c_int]
        lib.oraSolve.restype = ctypes.c_int
        _oracle = lib
    return _oracle


def make_fitter(cls, backend, **kwds):
    """An instance of a dao_fit_c class bound to 'oracle' or 'hip'."""
    if backend == "oracle":
        lib = oracle_lib()
        with mock.patch.object(dao_fit_c, "loadDaoFitC", lambda: lib):
            return cls(**kwds)
    return cls(**kwds)


_mp_oracle = None


def mp_oracle_lib():
    global _mp_oracle
    if _mp_oracle is None:
        if not os.path.exists(MP_ORACLE_SO):
            raise RuntimeError("multiplane oracle not built: make -C oracle")
        from storm_analysis_amd.multi_plane import mp_fit_dao_c
        _mp_oracle = mp_fit_dao_c.setMPFitDaoArgtypes(ctypes.cdll.LoadLibrary(MP_ORACLE_SO))
    return _mp_oracle


def make_mp_fitter(cls, backend, **kwds):
    """An instance of a multi_plane.mp_fit_dao_c class bound to 'oracle' or 'hip'."""
    from storm_analysis_amd.multi_plane import mp_fit_dao_c
    if backend == "oracle":
        lib = mp_oracle_lib()
        with mock.patch.object(mp_fit_dao_c, "loadMPFitCDao", lambda: lib):
            return cls(**kwds)
    return cls(**kwds)


@pytest.fixture
def oracle():
    return oracle_lib()
