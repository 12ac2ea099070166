"""
Post-fit filter oracle -- test infrastructure only.

Used by tests/ (and bench.py's CPU leg) to check the device filters
(markDimmerPeaks, markLowSignificancePeaks, runningIfHasNeighbors and their
batched forms in libdao_fit.so) status for status.  Two restatements of
storm_analysis/sa_library/ia_utilities.c:221-430:

- *_seq: the reference's loops as written, in index order, with the kd-tree's
  range query (storm_analysis/dbscan/kdtree.c:262-293) replaced by a scan of
  every peak with the same inclusive test, dist_sq <= range * range, where
  dist_sq = (x_k - x_i)^2 + (y_k - y_i)^2 (kdtree.c:269-273).  Pure Python,
  for small lists;
- the unsuffixed functions: the order-independent two-pass form (flag the
  removed peaks, then update their neighbours) in numpy, for larger lists.

Positions are finite and radii positive (for other inputs the reference's
kd-tree pruning, kdtree.c:283, makes its answer depend on the tree's shape).
"""
import numpy

RUNNING = 0
CONVERGED = 1
ERROR = 2


def _near(x, y, i, r):
    """Indices k with the kd-tree's dist_sq of peak k from peak i <= r * r."""
    r2 = r * r
    return [k for k in range(x.size) if (x[k] - x[i]) ** 2 + (y[k] - y[i]) ** 2 <= r2]


def mark_dimmer_peaks_seq(x, y, h, status, r_removal, r_neighbors):
    """markDimmerPeaks (ia_utilities.c:221-284); updates status, returns the removed count."""
    removed = 0
    for i in range(x.size):
        if status[i] == ERROR:
            continue
        set_r = _near(x, y, i, r_removal)
        if len(set_r) < 2:
            continue
        if any(h[k] > h[i] for k in set_r):
            removed += 1
            status[i] = ERROR
            for k in _near(x, y, i, r_neighbors):
                if status[k] == CONVERGED:
                    status[k] = RUNNING
    return removed


def mark_low_significance_peaks_seq(x, y, sig, status, min_sig, r_neighbors):
    """markLowSignificancePeaks (ia_utilities.c:287-339)."""
    removed = 0
    for i in range(x.size):
        if status[i] == ERROR:
            continue
        if sig[i] > min_sig:
            continue
        status[i] = ERROR
        removed += 1
        for k in _near(x, y, i, r_neighbors):
            if status[k] == CONVERGED:
                status[k] = RUNNING
    return removed


def running_if_has_neighbors_seq(status, c_x, c_y, n_x, n_y, radius):
    """runningIfHasNeighbors (ia_utilities.c:394-427); returns status (updated in place)."""
    r2 = radius * radius
    for i in range(c_x.size):
        if status[i] == RUNNING or status[i] == ERROR:
            continue
        if any((n_x[k] - c_x[i]) ** 2 + (n_y[k] - c_y[i]) ** 2 <= r2 for k in range(n_x.size)):
            status[i] = RUNNING
    return status


def _pairs_within(px, py, qx, qy, r, chunk=2048):
    """Boolean matrix [query, point]: the kd-tree's test of point from query."""
    r2 = r * r
    out = numpy.zeros((qx.size, px.size), dtype=bool)
    for s in range(0, qx.size, chunk):
        dx = px[None, :] - qx[s:s + chunk, None]
        dy = py[None, :] - qy[s:s + chunk, None]
        out[s:s + chunk] = (dx * dx + dy * dy) <= r2
    return out


def _update_neighbours(x, y, status, flag, r_neighbors):
    near = _pairs_within(x, y, x, y, r_neighbors)
    hit = (near & flag[None, :]).any(axis=1)
    status[(~flag) & (status == CONVERGED) & hit] = RUNNING
    status[flag] = ERROR
    return int(flag.sum())


def mark_dimmer_peaks(x, y, h, status, r_removal, r_neighbors):
    """markDimmerPeaks, two-pass form; same result as mark_dimmer_peaks_seq."""
    near = _pairs_within(x, y, x, y, r_removal)
    brighter = (near & (h[None, :] > h[:, None])).any(axis=1)
    flag = (status != ERROR) & brighter
    return _update_neighbours(x, y, status, flag, r_neighbors)


def mark_low_significance_peaks(x, y, sig, status, min_sig, r_neighbors):
    """markLowSignificancePeaks, two-pass form."""
    flag = (status != ERROR) & ~(sig > min_sig)
    return _update_neighbours(x, y, status, flag, r_neighbors)


def running_if_has_neighbors(status, c_x, c_y, n_x, n_y, radius):
    """runningIfHasNeighbors, vectorised."""
    if n_x.size == 0:
        return status
    hit = _pairs_within(n_x, n_y, c_x, c_y, radius).any(axis=1)
    status[(status != RUNNING) & (status != ERROR) & hit] = RUNNING
    return status
