"""
storm_analysis_amd -- an MI355X-native 3D-DAOSTORM multi-emitter fitter.

The Python surface mirrors storm_analysis.sa_library.dao_fit_c; the work runs
in HIP kernels inside c_libraries/libdao_fit.so, loaded with ctypes the way
the reference loads its own libdao_fit (storm_analysis/sa_library/loadclib.py).
The directory name contains a dash, so import it through
__graft_entry__.load_package() (or register it under the name
"storm_analysis_amd" with importlib).
"""
