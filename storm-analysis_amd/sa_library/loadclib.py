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


def loadCLibrary(library_filename):
    path = libraryPath(library_filename)
    if not os.path.exists(path):
        raise LoadCLibraryException(
            "native library " + path + " is not built; run __graft_entry__.build()")
    return ctypes.cdll.LoadLibrary(path)
