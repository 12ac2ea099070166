"""
Autotuner of the batched fitter: times doFit on the caller's own frames for
each candidate (waves per workgroup, reserved-box margin) and keeps the
fastest.

Waves per workgroup is the number of peaks a frame's workgroup fits at once
(fewer waves leave room for more frames per CU); the margin is the slack of
each peak's reserved box around its ROI (a larger box lets the ROI move
without a rollback but serializes more neighbours).  Results do not depend on
the choice (to the parity bar), so tuning on the workload itself is safe.
"""
import time


def autotune(bf, prepare, waves=(2, 4, 8), margins=(1, 2, 3), max_iterations=2000, log=None):
    """
    bf: a BatchFitter; prepare(): reloads the same images and start peaks into
    it.  Returns (best, trials) where best = (waves, margin) is left set on bf
    and trials lists {"waves", "margin", "kernel_ms", "wall_ms"}.
    """
    trials = []
    bf.setTiming(True)
    for w in waves:
        for m in margins:
            bf.setTuning(w, m)
            prepare()
            s0 = bf.stats()
            t0 = time.perf_counter()
            bf.doFit(max_iterations)
            wall = (time.perf_counter() - t0) * 1e3
            ms = bf.stats()["lm_ms"] - s0["lm_ms"]
            trials.append({"waves": w, "margin": m, "kernel_ms": ms, "wall_ms": wall})
            if log is not None:
                log("autotune: waves %d margin %d: fit kernel %.1f ms" % (w, m, ms))
    best = min(trials, key=lambda t: t["kernel_ms"])
    bf.setTuning(best["waves"], best["margin"])
    return (best["waves"], best["margin"]), trials
