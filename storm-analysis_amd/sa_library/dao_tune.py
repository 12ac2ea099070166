"""
Autotuner of the batched fitter: times fits of the caller's own frames for
each candidate and keeps the fastest.  No counterpart in the
reference, whose fitter has no launch shape to choose.

The knobs:
- waves: the peaks a frame's workgroup fits at once (one wavefront per peak
  turn); fewer waves leave room for more frames per CU;
- tile: the side of the reserved box each peak turn holds, the ROI plus the
  margin on each side (for ROI 18: 18^2 with no margin up to 24^2 with
  three pixels); a larger tile lets a peak's ROI move without a rollback
  but serializes more neighbouring turns.  A tile smaller than the ROI
  cannot hold a turn and is not a candidate;
- team: peaks per workgroup when several workgroups share a frame (0: one
  workgroup per frame), which lets a large frame use many CUs.

Results do not depend on the choice (a turn that leaves its tile is rolled
back exactly), so tuning on the workload itself is safe.  Trials are timed
by the fit kernel's HIP events, in two stages when `trial_sweeps` is given:
waves and tile are chosen on fits of `trial_sweeps` sweeps (the rollbacks a
small tile causes and the occupancy of the waves show in the first sweeps),
then the team size on whole fits at that choice: teams change the cost of
every sweep of a large frame, and the long tail of sweeps with a few
hundred peaks still running can hold most of the time (on an uneven sCMOS
batch with a 2048^2 frame, 10-sweep trials ranked one workgroup per frame
first, whose whole fit then took 1.7x as long as with teams).  Without
`trial_sweeps` every candidate is a whole fit.  The pick is kept in a JSON
cache keyed by the workload's signature -- model, error model, ROI, precision,
each frame's size and its peak count rounded to a power of two, and the
library build -- so a restarted server reuses it without trials.
"""
import json
import os
import time

from . import loadclib


def workload_key(bf):
    """Signature of the workload held by BatchFitter bf (see the module docstring)."""
    sizes = {}
    for f, (h, w) in enumerate(bf.shapes):
        n = max(1, int(bf.getNFit(f)))
        k = "%dx%d:%d" % (h, w, 1 << (n - 1).bit_length())
        sizes[k] = sizes.get(k, 0) + 1
    lib = loadclib.libraryPath("dao_fit")
    try:
        st = os.stat(lib)
        build = "%d-%d" % (st.st_size, int(st.st_mtime))
    except OSError:
        build = "unknown"
    return "|".join([bf.model, bf.error_model, "roi%d" % bf.roi_size, bf.precision, build] +
                    ["%s*%d" % (k, v) for k, v in sorted(sizes.items())])


def margins_for(tiles, roi_size):
    """Reserved-box margins of the tile sides that can hold an ROI of roi_size."""
    out = []
    for t in tiles:
        if t >= roi_size and (t - roi_size) % 2 == 0 and (t - roi_size) // 2 not in out:
            out.append((t - roi_size) // 2)
    return tuple(out)


def _load_cache(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_cache(path, key, best):
    data = _load_cache(path)
    data[key] = {"waves": best[0], "margin": best[1], "team": best[2]}
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def autotune(bf, prepare, waves=(2, 4, 8), margins=None, teams=(0,), tiles=None, max_iterations=2000,
             trial_sweeps=None, cache=None, log=None):
    """
    bf: a BatchFitter; prepare(): reloads the same images and start peaks into
    it.  Candidates: waves x (margins, or the tile sides `tiles`, default
    margins 1, 2, 3) x teams.  With trial_sweeps, waves and tile come from
    fits of that many sweeps (with the first team size), then the team from
    whole fits (max_iterations); without, every candidate is a whole fit.
    With `cache` (a JSON file path) a pick stored for the same workload
    signature is used without trials, and a new pick is stored.  Returns
    (best, trials): best = (waves, margin, team), left set on bf; trials
    lists {"waves", "margin", "tile", "team", "kernel_ms", "wall_ms",
    "sweeps", "stage"} (empty when the cache answered).
    """
    key = workload_key(bf) if cache else None
    if cache:
        hit = _load_cache(cache).get(key)
        if hit is not None:
            best = (int(hit["waves"]), int(hit["margin"]), int(hit["team"]))
            bf.setTuning(*best)
            if log is not None:
                log("autotune: cached pick for this workload: waves %d margin %d team %d" % best)
            return best, []
    if margins is None:
        margins = margins_for(tiles, bf.roi_size) if tiles is not None else (1, 2, 3)
    if not margins:
        raise ValueError("no tile can hold the ROI")
    trials = []
    bf.setTiming(True)

    def trial(w, m, t, sweeps, stage):
        bf.setTuning(w, m, t)
        prepare()
        s0 = bf.stats()
        t0 = time.perf_counter()
        bf.doFit(sweeps)
        wall = (time.perf_counter() - t0) * 1e3
        ms = bf.stats()["lm_ms"] - s0["lm_ms"]
        trials.append({"waves": w, "margin": m, "tile": bf.roi_size + 2 * m, "team": t, "kernel_ms": ms,
                       "wall_ms": wall, "sweeps": sweeps, "stage": stage})
        if log is not None:
            log("autotune: waves %d tile %d^2 team %d: fit kernel %.1f ms (%d sweeps)" %
                (w, bf.roi_size + 2 * m, t, ms, sweeps))
        return ms

    if trial_sweeps is None:
        for t in teams:
            for w in waves:
                for m in margins:
                    trial(w, m, t, max_iterations, 0)
        final = trials
    else:
        # stage 1: waves and tile on short fits with the first team size
        _, w1, m1 = min((trial(w, m, teams[0], trial_sweeps, 1), w, m) for w in waves for m in margins)
        # stage 2: the team size on whole fits
        for t in teams:
            trial(w1, m1, t, max_iterations, 2)
        final = [r for r in trials if r["stage"] == 2]
    b = min(final, key=lambda r: r["kernel_ms"])
    best = (b["waves"], b["margin"], b["team"])
    bf.setTuning(*best)
    if cache:
        _store_cache(cache, key, best)
    return best, trials
