// libmp_fit_dao.so: the library name under which the reference's Python
// loads the multiplane fitter (storm_analysis/multi_plane/mp_fit_dao_c.py:17-18).
// It holds no code of its own: libdao_fit.so, its dependency, carries the
// multiplane entry points (include/mp_fit_amd.h) and the per-channel mFit*
// entry points, and symbol look-ups through this library resolve there.
