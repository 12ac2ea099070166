"""
Loads the package's native library from <package>/c_libraries/lib<name>.so,
the same convention as storm_analysis/sa_library/loadclib.py:15-70.

There is no fallback: a missing library is an error.
"""
import ctypes
import os

_here = os.path.dirname(os.path.abspath(__file__))
c_lib_path = os.path.join(os.path.dirname(_here), "c_libraries")


class LoadCLibraryException(Exception):
    pass


def libraryPath(library_filename):
    return os.path.join(c_lib_path, "lib" + library_filename + ".so")


def _share_runtime_with_torch():
    """
    A process can hold only one HSA runtime.  PyTorch ships its own copy,
    which its libraries request under an unversioned file name; if the
    fitter's library (which requests libhsa-runtime64.so.1) loads first,
    importing torch later maps a second copy that finds no device.  With
    torch imported first, the fitter's library binds to torch's copy (same
    soname) and device pointers from torch tensors are valid in both.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def loadCLibrary(library_filename):
    path = libraryPath(library_filename)
    if not os.path.exists(path):
        raise LoadCLibraryException(
            "native library " + path + " is not built; run __graft_entry__.build()")
    _share_runtime_with_torch()
    return ctypes.cdll.LoadLibrary(path)
