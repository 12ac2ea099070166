"""
Local-maxima finder oracle -- test infrastructure only.

Used by tests/ (and bench.py's CPU leg) to check the device finder
(findLocalMaxima / calcMaxPeaks / daoFindLocalMaxima in libdao_fit.so) peak
for peak.  Two restatements of storm_analysis/sa_library/ia_utilities.c:52-214:

- find_local_maxima_seq: the reference's loops as written (index order, the
  search window assigned from double to int as C does, the same tie rule).
  Pure Python, for small images;
- find_local_maxima: the same decision for all pixels at once in numpy (one
  shifted comparison per window offset), then the maxima in (plane, row,
  column) order, cut at n_peaks.
Both update `taken` (one int32 array per plane) as the reference does.
"""
import ctypes
import math

import numpy


def _trunc(v):
    """C's double -> int assignment."""
    return int(math.trunc(v))


def calc_max_peaks(images, taken, margin, threshold):
    """calcMaxPeaks (ia_utilities.c:52-75)."""
    n = 0
    for img, tk in zip(images, taken):
        ys, xs = img.shape
        sub = (slice(margin, ys - margin), slice(margin, xs - margin))
        n += int(((img[sub] > threshold) & (tk[sub] < 1)).sum())
    return n


def find_local_maxima_seq(images, taken, margin, z_range, radius, threshold, z_values, n_peaks):
    """findLocalMaxima (ia_utilities.c:105-160) + isLocalMaxima (:175-214).
    Returns (z, y, x, h) lists and updates taken."""
    zs = len(images)
    ys, xs = images[0].shape
    rr = _trunc(radius * radius)
    out = ([], [], [], [])
    for zi in range(zs):
        sz = max(zi - z_range, 0)
        ez = min(zi + z_range, zs - 1)
        for yi in range(margin, ys - margin):
            sy = _trunc(yi - radius)
            sy = 0 if sy < 0 else sy
            ey = _trunc(yi + radius)
            ey = ys - 1 if ey >= ys else ey
            for xi in range(margin, xs - margin):
                cur = images[zi][yi, xi]
                if not (cur > threshold) or taken[zi][yi, xi] >= 1:
                    continue
                sx = _trunc(xi - math.ceil(radius))
                sx = 0 if sx < 0 else sx
                ex = _trunc(xi + math.ceil(radius))
                ex = xs - 1 if ex >= xs else ex
                is_max = True
                for zz in range(sz, ez + 1):
                    for yy in range(sy, ey + 1):
                        dy = (yy - yi) * (yy - yi)
                        for xx in range(sx, ex + 1):
                            if (xx - xi) * (xx - xi) + dy <= rr:
                                v = images[zz][yy, xx]
                                if yy <= yi and xx <= xi:
                                    if v > cur:
                                        is_max = False
                                elif v >= cur:
                                    is_max = False
                            if not is_max:
                                break
                        if not is_max:
                            break
                    if not is_max:
                        break
                if is_max:
                    taken[zi][yi, xi] += 1
                    for lst, v in zip(out, (z_values[zi], float(yi), float(xi), cur)):
                        lst.append(v)
                if len(out[0]) >= n_peaks:
                    return out
    return out


def find_local_maxima(images, taken, margin, z_range, radius, threshold, z_values, n_peaks):
    """The same result for all pixels at once; returns (z, y, x, h) arrays."""
    stack = numpy.asarray(images, dtype=numpy.float64)
    zs, ys, xs = stack.shape
    tk = numpy.asarray(taken)
    rr = _trunc(radius * radius)
    cr = math.ceil(radius)
    cand = numpy.zeros(stack.shape, dtype=bool)
    cand[:, margin:ys - margin, margin:xs - margin] = True
    cand &= (stack > threshold) & (tk < 1)
    is_max = cand.copy()
    # window offsets (dy, dx) the reference's loops can visit: y in
    # [trunc(y - r), trunc(y + r)], x in [trunc(x - ceil r), trunc(x + ceil r)],
    # with dx^2 + dy^2 <= rr; the truncation toward zero matters only where the
    # window is clipped at the top row, so it is applied per row below.
    if radius >= 0 and z_range >= 0:
        span = int(cr) + 1
        lo = numpy.array([max(_trunc(y - radius), 0) for y in range(ys)])
        hi = numpy.array([min(_trunc(y + radius), ys - 1) for y in range(ys)])
        for dz in range(-min(z_range, zs), min(z_range, zs) + 1):
            for dy in range(-span, span + 1):
                for dx in range(-span, span + 1):
                    if dx * dx + dy * dy > rr or abs(dx) > cr:
                        continue
                    # rows whose window [max(trunc(y - r), 0), min(trunc(y + r), ys - 1)] holds y + dy
                    ny = numpy.arange(ys) + dy
                    row_ok = (ny >= lo) & (ny <= hi)
                    for zi in range(zs):
                        zz = zi + dz
                        if zz < 0 or zz >= zs:
                            continue
                        src = numpy.full((ys, xs), -numpy.inf)
                        y0, y1 = max(0, -dy), min(ys, ys - dy)
                        x0, x1 = max(0, -dx), min(xs, xs - dx)
                        src[y0:y1, x0:x1] = stack[zz, y0 + dy:y1 + dy, x0 + dx:x1 + dx]
                        cur = stack[zi]
                        veto = (src > cur) if (dy <= 0 and dx <= 0) else (src >= cur)
                        veto &= row_ok[:, None]
                        is_max[zi] &= ~veto
    zi, yi, xi = numpy.nonzero(is_max)
    n = min(zi.size, max(n_peaks, 0))
    zi, yi, xi = zi[:n], yi[:n], xi[:n]
    for a, b, c in zip(zi, yi, xi):
        taken[a][b, c] += 1
    z_values = numpy.asarray(z_values, dtype=numpy.float64)
    return z_values[zi], yi.astype(numpy.float64), xi.astype(numpy.float64), stack[zi, yi, xi]


class flmData(ctypes.Structure):
    """The reference's finder state (ia_utilities.c:19-38)."""
    _fields_ = [('margin', ctypes.c_int), ('npeaks', ctypes.c_int), ('zrange', ctypes.c_int),
                ('xsize', ctypes.c_int), ('ysize', ctypes.c_int), ('zsize', ctypes.c_int),
                ('radius', ctypes.c_double), ('threshold', ctypes.c_double),
                ('zvalues', ctypes.c_void_p), ('taken', ctypes.c_void_p), ('images', ctypes.c_void_p)]


def flm_call(lib, images, taken, margin, z_range, radius, threshold, z_values, n_peaks):
    """findLocalMaxima of a library exporting the reference's C interface."""
    zs = len(images)
    ys, xs = images[0].shape
    P = ctypes.c_void_p * zs
    c_im, c_tk = P(), P()
    for i in range(zs):
        c_im[i] = images[i].ctypes.data
        c_tk[i] = taken[i].ctypes.data
    zv = numpy.ascontiguousarray(z_values, dtype=numpy.float64)
    f = flmData(margin=margin, npeaks=n_peaks, zrange=z_range, xsize=xs, ysize=ys, zsize=zs, radius=radius,
                threshold=threshold, zvalues=zv.ctypes.data, taken=ctypes.addressof(c_tk),
                images=ctypes.addressof(c_im))
    out = [numpy.zeros(max(n_peaks, 1)) for _ in range(4)]
    lib.findLocalMaxima.argtypes = [ctypes.POINTER(flmData)] + [numpy.ctypeslib.ndpointer(dtype=numpy.float64)] * 4
    lib.findLocalMaxima.restype = None
    lib.calcMaxPeaks.argtypes = [ctypes.POINTER(flmData)]
    lib.calcMaxPeaks.restype = ctypes.c_int
    n_cand = lib.calcMaxPeaks(ctypes.byref(f))
    lib.findLocalMaxima(ctypes.byref(f), *out)
    return [o[:f.npeaks] for o in out], n_cand
