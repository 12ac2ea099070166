"""
Autotuner of the batched fitter: times doFit on the caller's own frames for
each candidate (waves per workgroup, reserved-box margin, team size) and
keeps the fastest.

Waves per workgroup is the number of peaks a frame's workgroup fits at once
(fewer waves leave room for more frames per CU); the margin is the slack of
each peak's reserved box around its ROI (a larger box lets the ROI move
without a rollback but serializes more neighbours); the team size is the
number of peaks per workgroup when several workgroups share a frame (0: one
workgroup per frame), which lets a large frame use many CUs.  Results do not
depend on the choice (to the parity bar), so tuning on the workload itself
is safe.
"""
import time


def autotune(bf, prepare, waves=(2, 4, 8), margins=(1, 2, 3), teams=(0,), max_iterations=2000, log=None):
    """
    bf: a BatchFitter; prepare(): reloads the same images and start peaks into
    it.  Returns (best, trials) where best = (waves, margin, team) is left set
    on bf and trials lists {"waves", "margin", "team", "kernel_ms", "wall_ms"}.
    """
    trials = []
    bf.setTiming(True)
    for t in teams:
        for w in waves:
            for m in margins:
                bf.setTuning(w, m, t)
                prepare()
                s0 = bf.stats()
                t0 = time.perf_counter()
                bf.doFit(max_iterations)
                wall = (time.perf_counter() - t0) * 1e3
                ms = bf.stats()["lm_ms"] - s0["lm_ms"]
                trials.append({"waves": w, "margin": m, "team": t, "kernel_ms": ms, "wall_ms": wall})
                if log is not None:
                    log("autotune: waves %d margin %d team %d: fit kernel %.1f ms" % (w, m, t, ms))
    best = min(trials, key=lambda r: r["kernel_ms"])
    bf.setTuning(best["waves"], best["margin"], best["team"])
    return (best["waves"], best["margin"], best["team"]), trials
