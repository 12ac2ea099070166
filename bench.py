#!/usr/bin/env python
"""
Benchmark: converged emitter fits/s of the MI355X 3D-DAOSTORM fitter, batched
vs plain, with frame latency, on the 3D-astigmatism workload: 500 frames of
512 x 512, 1000 emitters per frame, "Z" model (wx = [2.0 px, +0.15 um, 0.4 um,
0, 0], wy = [2.0, -0.15, 0.4, 0, 0], z in [-0.5, 0.5] um), ROI 18 px, MLE,
tolerance 1e-6.

One step = newImage + newBackground + newPeaks ("hdf5" start peaks) + doFit
(until no peak is RUNNING, at most 2000 sweeps) + all peak properties back on
the host, for every frame of the batch; images and start peaks are resident
in HBM when the timed region starts.

`value` is the batched converged fits/s in FP64, the reference's arithmetic
(its results match the oracle to the parity bar).  Beside it: the same in
mixed precision (FP32 deviance terms, Jacobian and Hessian; FP64 peak shape,
fit-image accumulators, deviance sums, Cholesky solve and accept/reject) with
its accuracy against FP64 and the oracle; for each precision two
host-to-host rates (timed from the upload of host frames to the results on
the host: float64 frames, and raw uint16 camera frames converted on the GPU
by daoIngestFrames); the plain path (MultiFitterZ, one frame at a time
through the single-frame C-ABI, FP64, host arrays in and out) with its
p50/p99 frame latency; the fit kernel's HBM roofline (algorithmic bytes = LM
cycles x N_roi x 36 B over the kernel's event-timed duration); and the
oracle timed on the host's cores.
"""
import argparse
import concurrent.futures as cf
import ctypes
import json
import os
import sys
import time
from unittest import mock

import numpy

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__  # noqa: E402

HBM_PEAK_GBS = 8000.0
ROI = 18
N_ROI_18 = 256
ALG_BYTES_PER_PIXEL = 36
# HBM traffic of one batched k_fit_frames launch on the default workload (500
# frames of 512 x 512, 1000 emitters), per precision, from two rocprofv3 --pmc
# runs of `bench.py --steps 1 --warmup 0 --plain-frames 1 --no-cpu
# --precision P`: FETCH_SIZE and WRITE_SIZE in KiB per dispatch, scaled by
# what the same counters read for a 1 GiB device copy (`bench.py
# --copy-calibration 1024`: FETCH_SIZE reads half the bytes moved, WRITE_SIZE
# all of them).
MEASURED_TRAFFIC = {
    "mixed": {"frames": 500, "size": 512, "emitters": 1000, "margin": 2, "waves": 0, "commit": "79033c9",
              "fetch_kib": 94242076.2, "write_kib": 105784091.6, "fetch_scale": 2.0, "write_scale": 1.0},
    "fp64": {"frames": 500, "size": 512, "emitters": 1000, "margin": 2, "waves": 0, "commit": "79033c9",
             "fetch_kib": 94054058.6, "write_kib": 105365188.1, "fetch_scale": 2.0, "write_scale": 1.0},
}
# accuracy bar of mixed precision against FP64 for x and y
TOL_MIXED_XY_PX = 6.25e-3  # 1 nm at 160 nm pixels


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--frames", type=int, default=500)
    p.add_argument("--emitters", type=int, default=1000)
    p.add_argument("--size", type=int, default=512)
    p.add_argument("--plain-frames", type=int, default=20)
    p.add_argument("--cpu-frames", type=int, default=500)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-host", action="store_true", help="skip the host-to-host rates")
    p.add_argument("--margin", type=int, default=2)
    p.add_argument("--no-ingest", action="store_true", help="skip the frame-ingest leg")
    p.add_argument("--no-filter", action="store_true", help="skip the post-fit filter leg")
    p.add_argument("--no-finder", action="store_true", help="skip the local-maxima finder leg")
    p.add_argument("--no-mfilter", action="store_true", help="skip the FFT convolution leg")
    p.add_argument("--waves", type=int, default=0,
                   help="waves per workgroup of the batched fitter (0: the kernel's default)")
    p.add_argument("--serve-seconds", type=float, default=3.0,
                   help="seconds of simulated camera traffic per offered load for the serving leg (0: skip)")
    p.add_argument("--serve-loads", default="0.05,0.1,0.2",
                   help="offered loads of the serving leg, as fractions of 16 cameras at U(50, 400) frames/s")
    p.add_argument("--latency-frames", type=int, default=100,
                   help="frames fitted one at a time for the single-frame latency leg (0: skip)")
    p.add_argument("--team", type=int, default=64,
                   help="peaks per workgroup of the frame-per-team fitter in the latency and serving legs")
    p.add_argument("--serve-workers", type=int, default=2,
                   help="workers of the serving leg's latency set-up")
    p.add_argument("--serve-slots", type=int, default=32,
                   help="frame slots per worker in the serving leg's latency set-up")
    p.add_argument("--serve-setups", default="resident,latency",
                   help="serving set-ups to run: resident (frames admitted into a running kernel), "
                        "throughput, latency (a launch per batch)")
    p.add_argument("--resident-slots", type=int, default=0,
                   help="frame slots of the resident server (0: one per resident workgroup)")
    p.add_argument("--tune-sweeps", type=int, default=10,
                   help="sweeps of each autotuner trial on the uneven batch (its first sweeps hold most of the work)")
    p.add_argument("--resident-team", type=int, default=64,
                   help="peaks per workgroup of the resident server's team set-up (a 512^2 frame of 1000 "
                        "emitters gets 16 workgroups)")
    p.add_argument("--serve-team", type=int, default=250,
                   help="peaks per workgroup in the serving leg's latency set-up")
    p.add_argument("--no-dense", action="store_true",
                   help="skip the high-density leg (one 2048^2 frame, 20000 emitters, 2d model)")
    p.add_argument("--no-uneven", action="store_true",
                   help="skip the batched-uneven-frames leg (sCMOS 3D, 128^2..2048^2, autotuned)")
    p.add_argument("--precision", choices=("both", "mixed", "fp64"), default="both",
                   help="batched precisions to time (value = mixed when timed)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a plain device-to-device copy of MIB MiB three times and exit "
                        "(the known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def copy_calibration(mib):
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copied %d MiB three times (%d bytes read and %d written per copy)" % (mib, 4 * n, 4 * n))


def zkw():
    from storm_analysis_amd.diagnostics import synth
    return {"wx_params": synth.WX_Z, "wy_params": synth.WY_Z, "min_z": -0.5, "max_z": 0.5}


def frame_rows(rows, counts, f):
    c0 = int(counts[:f].sum())
    return rows[c0:c0 + int(counts[f])]


def peaks_dict(r):
    return {"x": r[:, 0], "y": r[:, 1], "z": r[:, 2], "background": r[:, 3], "height": r[:, 4],
            "xsigma": r[:, 5], "ysigma": r[:, 6]}


def fit_plain(cls_factory, image, background, rows, lib=None):
    m = cls_factory()
    if lib is not None and m.clib is not lib:
        raise RuntimeError("fitter is not bound to the expected library")
    m.initializeC(image)
    m.newImage(image)
    m.newBackground(background)
    m.newPeaks(peaks_dict(rows), "hdf5")
    m.doFit()
    out = {k: m.getPeakProperty(k) for k in ("x", "y", "z", "height", "background", "status")}
    m.cleanup(verbose=False)
    return out


def cpu_baseline(images, backgrounds, rows, counts, nframes, passes=2, single_frames=16):
    """
    The oracle (FP64 C restatement) on the host's cores, one fitter per frame:
    `passes` passes over the first nframes frames on up to 16 threads (the
    reported value), then `single_frames` frames on one thread.
    """
    from storm_analysis_amd.sa_library import dao_fit_c as D
    lib = D.setDaoFitArgtypes(ctypes.cdll.LoadLibrary(os.path.join(ROOT, "oracle", "liboracle_dao.so")))
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    # a GPU host allots 16 CPUs to each GPU; use them all, but no more
    cores = max(1, min(visible, 16))

    fac = lambda: D.MultiFitterZ(roi_size=ROI, **zkw())  # noqa: E731

    def one(f):
        out = fit_plain(fac, images[f], backgrounds[f], frame_rows(rows, counts, f), lib=lib)
        return int((out["status"] == D.CONVERGED).sum()), out

    # Every fitter built while the patch is active binds the oracle library;
    # the patch stays in place until all threads have finished.
    with mock.patch.object(D, "loadDaoFitC", lambda: lib):
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(cores) as ex:
            for _ in range(passes):
                res = list(ex.map(one, range(nframes)))
        dt = time.perf_counter() - t0
        t1 = time.perf_counter()
        conv1 = sum(one(f)[0] for f in range(min(single_frames, nframes)))
        dt1 = time.perf_counter() - t1
    conv = passes * sum(r[0] for r in res)
    return {"value": conv / dt, "unit": "converged fits/s", "cores": cores, "kind": "port",
            "host_cpus_visible": visible, "os_cpu_count": os.cpu_count(),
            "sample": "%d passes over %d of the workload's frames (%d converged fits) in %.1f s on %d threads" %
                      (passes, nframes, conv, dt, cores),
            "one_thread": {"value": conv1 / dt1, "frames": min(single_frames, nframes), "seconds": dt1}}, \
        [r[1] for r in res]


def host_steps(bf, precision, args, host, counts_np):
    """
    Host-to-host rates of the batch: per step the frames, backgrounds and
    start peaks are uploaded from (pinned) host memory, fitted, and the peak
    properties copied back to the host.  Two forms: float64 frames; raw uint16
    camera frames (ADU = photo-electrons + 100 offset, gain 1, as generated)
    converted to photo-electrons on the GPU by daoIngestFrames before the fit.
    """
    import torch
    from storm_analysis_amd.sa_library import dao_fit_c as D
    from storm_analysis_amd.sa_library.dao_ingest import FrameReaderStd
    reader = FrameReaderStd(camera_gain=1.0, camera_offset=100.0, device="cuda:0")
    out = {}
    total = int(counts_np.sum())
    pin_rows = torch.empty((total, 9), dtype=torch.float64, pin_memory=True).numpy()
    pin_status = torch.empty(total, dtype=torch.int32, pin_memory=True).numpy()
    for form in ("float64", "uint16"):
        def step():
            if form == "float64":
                bf.newFrames(host["images"], host["backgrounds"])
            else:
                bf.newImages(reader.loadFrames(host["raw"]))
                bf.newBackgrounds(host["backgrounds"])
            bf.newPeaks(host["rows"], "hdf5", counts=counts_np)
            bf.doFit()
            return bf.getAllPeaks(pin_rows, pin_status)
        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        conv = 0
        for _ in range(args.steps):
            _, status, _ = step()
            conv += int((status == D.CONVERGED).sum())
        dt = time.perf_counter() - t0
        out[form] = {"fits_per_s": conv / dt, "ms_per_step": dt * 1e3 / args.steps,
                     "upload_bytes_per_step": int(host["bytes"][form])}
        log("%s host-to-host (%s frames): %.1f ms/step, %.0f converged fits/s" %
            (precision, form, dt * 1e3 / args.steps, conv / dt))
    return out


def run_batched(precision, args, images, backgrounds, rows, counts_np, host=None):
    """Time args.steps batched steps in one precision; returns the measurements and results."""
    import torch
    from storm_analysis_amd.sa_library import dao_batch, dao_fit_c as D
    F, S, N = args.frames, args.size, args.emitters
    bf = dao_batch.BatchFitter([(S, S)] * F, [N] * F, model="Z", roi_size=ROI, margin=args.margin,
                               precision=precision, **zkw())
    if args.waves:
        bf.setTuning(args.waves, -1, -1)

    # the peak rows come back into pinned host memory, which the device
    # writes directly (no staging through pageable memory)
    total = int(counts_np.sum())
    pin_rows = torch.empty((total, 9), dtype=torch.float64, pin_memory=True).numpy()
    pin_status = torch.empty(total, dtype=torch.int32, pin_memory=True).numpy()

    def step():
        bf.newFrames(images, backgrounds)  # newImage + newBackground, one pass
        bf.newPeaks(rows, "hdf5", counts=counts_np)
        bf.doFit()
        return bf.getAllPeaks(pin_rows, pin_status)

    log("%s: batch of %d frames created; warmup %d step(s)" % (precision, F, args.warmup))
    for _ in range(args.warmup):
        step()
    for f in range(F):
        bf.getCounters(f)
    bf.setTiming(True)
    s0 = bf.stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    conv = 0
    for _ in range(args.steps):
        out, status, per_frame = step()
        conv += int((status == D.CONVERGED).sum())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out, status = out.copy(), status.copy()
    s1 = bf.stats()
    cycles = sum(bf.getCounters(f)["n_iterations"] for f in range(F))
    lm_ms = s1["lm_ms"] - s0["lm_ms"]
    launches = s1["lm_launches"] - s0["lm_launches"]
    alg_bytes = float(cycles) * N_ROI_18 * ALG_BYTES_PER_PIXEL
    achieved = alg_bytes / (lm_ms * 1e-3) / 1e9 if lm_ms > 0 else 0.0
    traffic = None
    mt = MEASURED_TRAFFIC.get(precision)
    if mt and (F, S, N, args.margin, args.waves) == (mt["frames"], mt["size"], mt["emitters"], mt["margin"],
                                                      mt["waves"]):
        traffic = 1024.0 * (mt["fetch_kib"] * mt["fetch_scale"] + mt["write_kib"] * mt["write_scale"])
    st = bf.stats()
    bf.setTiming(False)
    log("%s: %.1f ms/step, %d converged per step, fit kernel %.1f ms per launch" %
        (precision, dt * 1e3 / args.steps, conv // max(1, args.steps), lm_ms / max(1, launches)))
    host_to_host = host_steps(bf, precision, args, host, counts_np) if host is not None else None
    bf.cleanup()
    return {
        "fits_per_s": conv / dt,
        "ms_per_step": dt * 1e3 / args.steps,
        "converged_per_step": conv // max(1, args.steps),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 --pmc, copy-scaled)",
                     "traffic_commit": mt.get("commit") if traffic is not None else None,
                     "alg_bytes_per_launch": alg_bytes / max(1, launches),
                     "kernel": "k_fit_frames (persistent frame-per-workgroup doFit, %s)" % precision,
                     "lm_ms": lm_ms, "lm_launches": launches,
                     "avg_launch_us": (lm_ms * 1e3 / launches) if launches else None,
                     "lm_cycles": cycles, "bytes_per_cycle": N_ROI_18 * ALG_BYTES_PER_PIXEL},
        "scheduler": {"sweeps": st["sweeps"], "escape_reruns": st["escape_reruns"]},
        "host_to_host": host_to_host,
    }, (out, status, per_frame)


def agreement(a, b, D):
    """Flag agreement and, over peaks CONVERGED in both, the |dx|,|dy| and |dz| distributions."""
    out_a, st_a = a
    out_b, st_b = b
    agree = float(numpy.mean(st_a == st_b)) if st_a.size else 1.0
    both = (st_a == D.CONVERGED) & (st_b == D.CONVERGED)
    dxy = numpy.maximum(numpy.abs(out_a[both, 0] - out_b[both, 0]), numpy.abs(out_a[both, 1] - out_b[both, 1]))
    dz = numpy.abs(out_a[both, 2] - out_b[both, 2])
    q = lambda v, t: float(numpy.quantile(v, t)) if v.size else 0.0  # noqa: E731
    return {"flag_agreement": agree, "converged_a": int((st_a == D.CONVERGED).sum()),
            "converged_b": int((st_b == D.CONVERGED).sum()),
            "frac_dxy_within_1nm": float(numpy.mean(dxy < TOL_MIXED_XY_PX)) if dxy.size else 1.0,
            "dxy_px_p50": q(dxy, 0.5), "dxy_px_p99": q(dxy, 0.99), "dxy_px_p999": q(dxy, 0.999),
            "max_dxy_px": float(dxy.max(initial=0.0)),
            "dz_um_p99": q(dz, 0.99), "max_dz_um": float(dz.max(initial=0.0))}


def result_rows(outs):
    rows = numpy.concatenate([numpy.stack([o["x"], o["y"], o["z"]], axis=1) for o in outs])
    st = numpy.concatenate([o["status"] for o in outs])
    return rows, st


def uneven_leg(log, tune_sweeps=10, seed=20261209):
    """
    Batched uneven frames in one launch with sCMOS weighting, '3d' model (free
    wx, wy), sigma U(1.2, 2.0), ROI 18: 8 frames of 128^2, 8 of 256^2, 4 of
    512^2 at densities U(0.1, 3) emitters/um^2 (160 nm pixels) and one 2048^2
    frame at 0.2 /um^2; per-pixel offset N(100, 5), gain N(2, 0.1), read
    variance lognormal(ln 4, 0.8).  The autotuner picks the reserved tile
    (ROI 18 plus 0-3 pixels each side: 18^2 to 24^2) on fits of
    tune_sweeps sweeps, then the team size (peaks per workgroup; 0 = one
    workgroup per frame, which leaves the 2048^2 frame on one CU) on whole
    fits; FP64 and mixed precision are timed with the pick.
    """
    import numpy as np
    from storm_analysis_amd.sa_library import dao_batch, dao_fit_c as D, dao_tune
    from storm_analysis_amd.diagnostics import synth
    rng = np.random.default_rng(seed)
    spec = [(128, d) for d in rng.uniform(0.1, 3.0, 8)] + [(256, d) for d in rng.uniform(0.1, 3.0, 8)] + \
           [(512, d) for d in rng.uniform(0.1, 3.0, 4)] + [(2048, 0.2)]
    t0 = time.perf_counter()
    frames = [synth.frame_3d_scmos(rng, size=s, density=d) for s, d in spec]
    log("uneven: %d frames, %d emitters generated in %.1f s" %
        (len(frames), sum(len(f.peaks["x"]) for f in frames), time.perf_counter() - t0))
    kw = {"roi_size": ROI, "sigma_range": [0.6, 10.0]}
    bgs = [np.full(f.image.shape, 100.0) for f in frames]
    out = {"frames": [[s, round(float(d), 3), int(len(f.peaks["x"]))] for (s, d), f in zip(spec, frames)],
           "model": "3d", "calibration": "sCMOS"}
    for prec in ("mixed", "fp64"):
        bf = dao_batch.BatchFitter([f.image.shape for f in frames], [len(f.peaks["x"]) for f in frames],
                                   model="3d", precision=prec, **kw)

        def prepare():
            for k, f in enumerate(frames):
                bf.setCalibration(k, rqe=f.rqe, scmos_cal=f.scmos_cal)
            bf.newImages([f.image for f in frames])
            bf.newBackgrounds(bgs)
            bf.newPeaks([f.peaks for f in frames], "hdf5")

        if prec == "mixed":
            t_tune = time.perf_counter()
            best, trials = dao_tune.autotune(bf, prepare, waves=(8,), tiles=(ROI, ROI + 2, ROI + 4, ROI + 6),
                                             teams=(0, 500, 250, 125), trial_sweeps=tune_sweeps, log=log)
            out["autotune"] = {"best": {"waves": best[0], "margin": best[1], "tile": ROI + 2 * best[1],
                                        "team": best[2]},
                               "trials": trials, "trial_sweeps": tune_sweeps,
                               "seconds": time.perf_counter() - t_tune}
        else:
            b = out["autotune"]["best"]
            bf.setTuning(b["waves"], b["margin"], b["team"])
        prepare()
        # counters since the last read: drop what the autotuner's trial fits
        # left, so lm_cycles counts the timed fit only
        for k in range(len(frames)):
            bf.getCounters(k)
        s0 = bf.stats()
        t0 = time.perf_counter()
        bf.doFit()
        rows, status, counts = bf.getAllPeaks()
        dt = time.perf_counter() - t0
        conv = int((status == D.CONVERGED).sum())
        cycles = sum(bf.getCounters(k)["n_iterations"] for k in range(len(frames)))
        sweeps = bf.stats()["sweeps"] - s0["sweeps"]
        out[prec] = {"fit_and_readback_s": dt, "converged": conv, "fits_per_s": conv / dt,
                     "running_at_end": int((status == D.RUNNING).sum()), "lm_cycles": int(cycles),
                     "sweeps_of_slowest_frame": int(sweeps)}
        log("uneven %s: %d converged in %.3f s (%.0f fits/s); %d LM cycles, %d sweeps, %d still running" %
            (prec, conv, dt, conv / dt, cycles, sweeps, out[prec]["running_at_end"]))
        bf.cleanup()
    return out


def latency_leg(args, images, backgrounds, rows, counts_np, log):
    """
    Single-frame latency: frames of the main workload fitted one at a time,
    inputs already in HBM, time = load + doFit + all peaks back to the host.
    Mixed precision; one workgroup per frame, then the frame-per-team fitter
    (one workgroup per args.team peaks, reserved-box margin 1).
    """
    import numpy as np
    import torch
    from storm_analysis_amd.sa_library import dao_batch, dao_fit_c as D
    S, N = images.shape[1], int(counts_np[0])
    nfr = min(args.latency_frames, images.shape[0])
    starts = np.concatenate([[0], np.cumsum(counts_np)])
    out = {"frames": nfr, "frame_size": S, "emitters": N, "precision": "mixed", "runs": []}
    for team, margin in ((0, 2), (args.team, 1)):
        bf = dao_batch.BatchFitter([(S, S)], [N], model="Z", precision="mixed", roi_size=ROI, **zkw())
        bf.setTuning(0, margin, team)
        ts, conv = [], 0
        for k in range(nfr + 1):
            r = rows[int(starts[k]):int(starts[k + 1])]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bf.loadFrames([0], images[k:k + 1], backgrounds[k:k + 1], [r], "hdf5")
            bf.doFit()
            _, status, _ = bf.getAllPeaks()
            dt = (time.perf_counter() - t0) * 1e3
            if k > 0:  # the first frame warms the kernels
                ts.append(dt)
                conv += int((status == D.CONVERGED).sum())
        bf.cleanup()
        ts = np.array(ts)
        run = {"team": team, "margin": margin, "p50_ms": float(np.percentile(ts, 50)),
               "p99_ms": float(np.percentile(ts, 99)), "max_ms": float(ts.max()), "converged": conv}
        out["runs"].append(run)
        log("latency team %d margin %d: p50 %.1f ms, p99 %.1f ms per %dx%d frame" %
            (team, margin, run["p50_ms"], run["p99_ms"], S, S))
    return out


def ingest_leg(log, frames=500, size=512, pad=10, reps=10, cpu_frames=20, seed=20261211):
    """
    Frame ingest (daoIngestFrames): 500 big-endian uint16 frames of 512x512,
    resident on the GPU, to photo-electrons (standard camera: gain 2 ADU/e-,
    offset 100) mirror-padded by 10 px, as FrameReader.loadAFrame + padArray
    prepare each frame on the host in the reference.  HBM bound: 2 B read per
    frame pixel, 8 B written per padded pixel.
    """
    import torch
    from storm_analysis_amd.sa_library import dao_ingest
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
    import ingest_oracle
    g = torch.Generator(device="cuda").manual_seed(seed)
    raw = torch.randint(90, 4000, (frames, size, size), dtype=torch.int32, device="cuda",
                        generator=g).to(torch.uint16)
    reader = dao_ingest.FrameReaderStd(camera_gain=2.0, camera_offset=100.0)
    out = reader.loadFrames(raw, pad=pad, big_endian=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = reader.loadFrames(raw, pad=pad, big_endian=True)
    dt = (time.perf_counter() - t0) / reps
    bytes_ = frames * size * size * 2 + frames * (size + 2 * pad) ** 2 * 8
    host_raw = raw[:cpu_frames].cpu().numpy()
    t1 = time.perf_counter()
    want = ingest_oracle.ingest(host_raw, reader.offset, reader.gain, reader.rqe, pad, True)
    dt_cpu = time.perf_counter() - t1
    exact = bool(numpy.array_equal(out[:cpu_frames].cpu().numpy(), want))
    gbs = bytes_ / dt / 1e9
    log("ingest: %d frames %dx%d -> padded by %d: %.3f ms per call (%.0f GB/s, %.1f%% of HBM), "
        "bit-identical to the oracle: %s; CPU (numpy, 1 thread) %.1f frames/s" %
        (frames, size, size, pad, dt * 1e3, gbs, 100 * gbs / HBM_PEAK_GBS, exact, cpu_frames / dt_cpu))
    return {"frames": frames, "frame_size": size, "pad": pad, "big_endian": True, "ms_per_call": dt * 1e3,
            "frames_per_s": frames / dt, "bytes_per_call": bytes_,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS},
            "bit_identical_to_oracle": exact,
            "cpu_baseline": {"value": cpu_frames / dt_cpu, "unit": "frames/s", "cores": 1, "kind": "port",
                             "sample": "%d frames through the numpy restatement of loadAFrame + padArray" %
                                       cpu_frames}}


def filter_leg(log, frames=500, per=1000, new_per=100, size=512, sigma=1.5, reps=20, cpu_frames=100,
               seed=20261212):
    """
    Post-fit filters over a batch of peak lists resident on the GPU: 500
    frames x 1000 peaks uniform in 512 x 512, as fitting.py:718-782 runs them
    after each fit pass: runningIfHasNeighbors against 100 new peaks per frame,
    then markDimmerPeaks (r_removal = sigma 1.5, r_neighbors = 5 sigma) and
    markLowSignificancePeaks (minimum significance 1.5).  One pass = the three
    batched calls (each synchronous, so launch latency is included) plus the
    copy that restores the statuses.  Bytes per pass: x, y, status in and out
    for every peak in each of the three calls, heights or significances for
    two, and x, y of the new peaks.
    """
    import ctypes
    import torch
    from storm_analysis_amd.sa_library import dao_filter as DF
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
    import filter_oracle as FO
    rng = numpy.random.default_rng(seed)
    n, nn = frames * per, frames * new_per
    host = {"x": rng.uniform(0, size, n), "y": rng.uniform(0, size, n), "h": rng.uniform(50.0, 500.0, n),
            "sig": rng.uniform(0.0, 10.0, n), "nx": rng.uniform(0, size, nn), "ny": rng.uniform(0, size, nn),
            "status": rng.choice([FO.RUNNING, FO.CONVERGED, FO.CONVERGED, FO.CONVERGED], n).astype(numpy.int32)}
    dev = {k: torch.from_numpy(v).cuda() for k, v in host.items()}
    off = torch.arange(0, n + 1, per, dtype=torch.int32, device="cuda")
    noff = torch.arange(0, nn + 1, new_per, dtype=torch.int32, device="cuda")
    removed = torch.zeros(frames, dtype=torch.int32, device="cuda")
    st = dev["status"].clone()
    lib = DF.loadFilterC()
    p = {k: ctypes.c_void_p(v.data_ptr()) for k, v in dev.items()}
    ps, po, pn, pr = (ctypes.c_void_p(t.data_ptr()) for t in (st, off, noff, removed))
    r_nb = 5.0 * sigma

    def one_pass():
        st.copy_(dev["status"])
        torch.cuda.synchronize()
        rc = lib.daoBatchRunningIfHasNeighbors(p["x"], p["y"], po, p["nx"], p["ny"], pn, ps, frames, r_nb)
        rc |= lib.daoBatchMarkDimmerPeaks(p["x"], p["y"], p["h"], ps, po, frames, sigma, r_nb, pr)
        n_dim = int(removed.sum())
        rc |= lib.daoBatchMarkLowSignificancePeaks(p["x"], p["y"], p["sig"], ps, po, frames, 1.5, r_nb, pr)
        if rc != 0:
            raise RuntimeError("filter call failed")
        return n_dim
    one_pass()
    t0 = time.perf_counter()
    for _ in range(reps):
        one_pass()
    dt = (time.perf_counter() - t0) / reps
    got = st.cpu().numpy()

    def cpu_frame(f, impl):
        a, b = f * per, (f + 1) * per
        s = host["status"][a:b].copy()
        x, y = host["x"][a:b].copy(), host["y"][a:b].copy()
        nx, ny = host["nx"][f * new_per:(f + 1) * new_per].copy(), host["ny"][f * new_per:(f + 1) * new_per].copy()
        h, sig = host["h"][a:b].copy(), host["sig"][a:b].copy()
        if impl is None:
            FO.running_if_has_neighbors(s, x, y, nx, ny, r_nb)
            FO.mark_dimmer_peaks(x, y, h, s, sigma, r_nb)
            FO.mark_low_significance_peaks(x, y, sig, s, 1.5, r_nb)
        else:
            impl.runningIfHasNeighbors(x, y, nx, ny, s, r_nb, per, new_per)
            impl.markDimmerPeaks(x, y, h, s, sigma, r_nb, per)
            impl.markLowSignificancePeaks(x, y, sig, s, 1.5, r_nb, per)
        return s
    exact = all(numpy.array_equal(got[f * per:(f + 1) * per], cpu_frame(f, None)) for f in (0, frames // 2, frames - 1))
    ref_so = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "_ref", "libia_utilities.so")
    impl, kind, what = None, "port", "the numpy restatement (oracle/filter_oracle.py)"
    if os.path.exists(ref_so):
        impl = ctypes.cdll.LoadLibrary(ref_so)
        P = numpy.ctypeslib.ndpointer
        d, i32 = P(dtype=numpy.float64), P(dtype=numpy.int32)
        for name in ("markDimmerPeaks", "markLowSignificancePeaks"):
            getattr(impl, name).argtypes = [d, d, d, i32, ctypes.c_double, ctypes.c_double, ctypes.c_int]
            getattr(impl, name).restype = ctypes.c_int
        impl.runningIfHasNeighbors.argtypes = [d, d, d, d, i32, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        impl.runningIfHasNeighbors.restype = None
        kind, what = "reference", "the reference's ia_utilities.c + kdtree.c built from source (oracle/_ref)"
    t1 = time.perf_counter()
    for f in range(cpu_frames):
        s = cpu_frame(f, impl)
        if impl is not None and f in (0, frames // 2):
            exact = exact and numpy.array_equal(got[f * per:(f + 1) * per], s)
    dt_cpu = (time.perf_counter() - t1) / cpu_frames
    bytes_ = n * (3 * 16 + 3 * 8 + 2 * 8) + nn * 16
    gbs = bytes_ / dt / 1e9
    log("filters: %d frames x %d peaks: %.3f ms per pass (runningIfHasNeighbors + markDimmerPeaks + "
        "markLowSignificancePeaks, %.1f M peaks/s, %.2f%% of HBM), statuses identical to the oracle%s: %s; "
        "CPU (%s, 1 thread) %.2f ms per frame" %
        (frames, per, dt * 1e3, n / dt / 1e6, 100 * gbs / HBM_PEAK_GBS,
         " and the reference C" if impl is not None else "", exact, kind, dt_cpu * 1e3))
    return {"frames": frames, "peaks_per_frame": per, "new_peaks_per_frame": new_per, "r_removal": sigma,
            "r_neighbors": r_nb, "min_significance": 1.5, "ms_per_pass": dt * 1e3, "peaks_per_s": n / dt,
            "bytes_per_pass": bytes_,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS},
            "identical_to_oracle": bool(exact),
            "cpu_baseline": {"value": per / dt_cpu, "unit": "peaks/s", "cores": 1, "kind": kind,
                             "sample": "%d frames of %d peaks through %s" % (cpu_frames, per, what)}}


def finder_leg(log, frames=500, size=512, n=1000, reps=10, cpu_frames=20, seed=20261213):
    """
    Local-maxima peak finding (daoFindLocalMaxima) over 500 frames of 512 x 512
    resident on the GPU, searched as 500 planes with z_range 0 (each frame on
    its own, as MaximaFinder.findMaxima does per frame): 1000 Gaussian spots per
    frame (sigma 1.5, heights U(5, 50)) on unit noise, margin 11, radius 3,
    threshold 6, values rounded to 0.1 (ties).  The taken counters are reset
    before each call.  Bytes per pixel: the image (8 B) and taken (4 B) read,
    the flag mask written and read back (2 B).
    """
    import ctypes
    import torch
    from storm_analysis_amd.sa_library import dao_finder
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
    import finder_oracle as FD
    rng = numpy.random.default_rng(seed)
    yy, xx = numpy.mgrid[0:size, 0:size]
    base = numpy.zeros((size, size))
    for y, x, h in zip(rng.uniform(0, size, n), rng.uniform(0, size, n), rng.uniform(5, 50, n)):
        y0, y1, x0, x1 = int(max(0, y - 6)), int(min(size, y + 7)), int(max(0, x - 6)), int(min(size, x + 7))
        base[y0:y1, x0:x1] += h * numpy.exp(-((yy[y0:y1, x0:x1] - y) ** 2 + (xx[y0:y1, x0:x1] - x) ** 2) / 4.5)
    g = torch.Generator(device="cuda").manual_seed(seed)
    st = torch.round((torch.from_numpy(base).cuda()[None] +
                      torch.randn((frames, size, size), dtype=torch.float64, device="cuda", generator=g)) * 10) / 10
    st = st.contiguous()
    taken = torch.zeros(st.shape, dtype=torch.int32, device="cuda")
    cap = frames * size * size // 8
    out = torch.empty((4, cap), dtype=torch.float64, device="cuda")
    zv = numpy.arange(frames, dtype=numpy.float64)
    lib = dao_finder.loadFinderC()
    nf = ctypes.c_int(0)
    ptrs = [ctypes.c_void_p(out[i].data_ptr()) for i in range(4)]

    def call():
        taken.zero_()
        torch.cuda.synchronize()
        rc = lib.daoFindLocalMaxima(ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(taken.data_ptr()), frames, size,
                                    size, 11, 0, 3.0, 6.0, zv.ctypes.data_as(ctypes.c_void_p), cap, *ptrs,
                                    ctypes.byref(nf))
        if rc != 0:
            raise RuntimeError("daoFindLocalMaxima failed (%d)" % rc)
        return nf.value
    found = call()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    dt = (time.perf_counter() - t0) / reps
    host = st[:cpu_frames].cpu().numpy()
    res = out[:, :found].cpu().numpy()
    ref_so = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "_ref", "libia_utilities.so")
    exact = True
    t1 = time.perf_counter()
    if os.path.exists(ref_so):
        rlib = ctypes.cdll.LoadLibrary(ref_so)
        kind, what = "reference", "the reference's findLocalMaxima (ia_utilities.c built from source, oracle/_ref)"
        outs = []
        for f in range(cpu_frames):
            tk = [numpy.zeros((size, size), dtype=numpy.int32)]
            (z, y, x, h), _ = FD.flm_call(rlib, [host[f]], tk, 11, 0, 3.0, 6.0, [float(f)], size * size)
            outs.append((z, y, x, h))
    else:
        kind, what = "port", "the numpy restatement (oracle/finder_oracle.py)"
        outs = []
        for f in range(cpu_frames):
            tk = [numpy.zeros((size, size), dtype=numpy.int32)]
            outs.append(FD.find_local_maxima([host[f]], tk, 11, 0, 3.0, 6.0, [float(f)], size * size))
    dt_cpu = (time.perf_counter() - t1) / cpu_frames
    want = numpy.concatenate([numpy.stack(o) for o in outs], axis=1)
    exact = bool(numpy.array_equal(res[:, :want.shape[1]], want))
    px = frames * size * size
    bytes_ = px * 14
    gbs = bytes_ / dt / 1e9
    log("finder: %d frames %dx%d, %d maxima: %.3f ms per call (%.0f GB/s, %.1f%% of HBM), first %d frames "
        "identical to the %s: %s; CPU (1 thread) %.2f ms per frame" %
        (frames, size, size, found, dt * 1e3, gbs, 100 * gbs / HBM_PEAK_GBS, cpu_frames, kind, exact, dt_cpu * 1e3))
    return {"frames": frames, "frame_size": size, "spots_per_frame": n, "maxima": found, "ms_per_call": dt * 1e3,
            "frames_per_s": frames / dt, "bytes_per_call": bytes_,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS},
            "identical_to_cpu": exact,
            "cpu_baseline": {"value": 1.0 / dt_cpu, "unit": "frames/s", "cores": 1, "kind": kind,
                             "sample": "%d frames through %s" % (cpu_frames, what)}}


def mfilter_leg(log, frames=500, size=512, reps=10, cpu_frames=20, seed=20261214):
    """
    FFT convolution (libmatched_filter.so, daoMatchedFilterConvolveBatch) of
    500 frames of 512 x 512 resident on the GPU with a unit-sum Gaussian PSF
    (sigma 8, the background filter's scale): one batched forward transform,
    the spectrum product and one batched inverse transform.  Bytes per pixel
    at the boundary: the frame read and the result written (16 B).
    """
    import torch
    from storm_analysis_amd.sa_library import dao_mfilter
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle"))
    import mfilter_oracle as MO
    psf = MO.gaussian_psf((size, size), 8.0)
    flt = dao_mfilter.MatchedFilter(psf)
    g = torch.Generator(device="cuda").manual_seed(seed)
    images = torch.rand((frames, size, size), dtype=torch.float64, device="cuda", generator=g) * 100.0
    out = flt.convolveBatch(images)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = flt.convolveBatch(images)
    dt = (time.perf_counter() - t0) / reps
    host = images[:cpu_frames].cpu().numpy()
    got = out[:cpu_frames].cpu().numpy()
    orc = MO.MatchedFilterOracle(psf)
    t1 = time.perf_counter()
    want = [orc.convolve(host[f]) for f in range(cpu_frames)]
    dt_cpu = (time.perf_counter() - t1) / cpu_frames
    err = max(float(numpy.max(numpy.abs(got[f] - want[f])) / numpy.max(numpy.abs(want[f])))
              for f in range(cpu_frames))
    flt.cleanup()
    bytes_ = frames * size * size * 16
    gbs = bytes_ / dt / 1e9
    log("mfilter: %d frames %dx%d: %.3f ms per batch (%.0f GB/s at the boundary, %.1f%% of HBM), max relative "
        "error vs the oracle %.1e; CPU (numpy FFT, 1 thread) %.2f ms per frame" %
        (frames, size, size, dt * 1e3, gbs, 100 * gbs / HBM_PEAK_GBS, err, dt_cpu * 1e3))
    return {"frames": frames, "frame_size": size, "psf_sigma": 8.0, "ms_per_batch": dt * 1e3,
            "frames_per_s": frames / dt, "bytes_per_batch": bytes_,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS},
            "max_rel_error_vs_oracle": err,
            "cpu_baseline": {"value": 1.0 / dt_cpu, "unit": "frames/s", "cores": 1, "kind": "port",
                             "sample": "%d frames through the numpy FFT restatement (oracle/mfilter_oracle.py); "
                                       "the reference's C needs FFTW, not in its sources" % cpu_frames}}


def dense_leg(log, seed=20261210):
    """
    High-density overlap stress: one 2048^2 frame with 20000 emitters
    (~4.6 /um^2 at 160 nm pixels), '2d' model (free width, sigma U(1.2, 2.0)),
    heights U(200, 5000), background 100, Poisson noise.  FP64 and mixed,
    one workgroup for the frame and the team fitter (500 peaks per
    workgroup); time = doFit; the team result must equal the one-workgroup
    result.
    """
    import numpy as np
    from storm_analysis_amd.sa_library import dao_batch, dao_fit_c as D
    from storm_analysis_amd.diagnostics import synth
    rng = np.random.default_rng(seed)
    fr = synth.frame_2d_dense(rng, size=2048, n=20000)
    kw = {"roi_size": ROI, "sigma_range": [0.96, 8.0]}
    out = {"frame_size": 2048, "emitters": int(len(fr.peaks["x"])), "model": "2d", "runs": []}
    res = {}
    for prec in ("mixed", "fp64"):
        for team in (0, 500):
            bf = dao_batch.BatchFitter([fr.image.shape], [len(fr.peaks["x"])], model="2d", precision=prec, **kw)
            bf.setTuning(0, 2, team)
            bf.setTiming(True)
            bf.newImages([fr.image])
            bf.newBackgrounds([np.full(fr.image.shape, 100.0)])
            bf.newPeaks([fr.peaks], "hdf5")
            t0 = time.perf_counter()
            bf.doFit()
            dt = time.perf_counter() - t0
            p = bf.getPeaks(0)
            c = bf.getCounters(0)
            sweeps = bf.stats()["sweeps"]
            bf.cleanup()
            res[(prec, team)] = p
            run = {"precision": prec, "team": team, "fit_s": dt, "converged": int((p["status"] == D.CONVERGED).sum()),
                   "running_at_end": int((p["status"] == D.RUNNING).sum()), "lm_cycles": int(c["n_iterations"]),
                   "sweeps": int(sweeps)}
            out["runs"].append(run)
            log("dense 2048^2 %s team %d: %.2f s, %d converged, %d LM cycles, %d sweeps" %
                (prec, team, dt, run["converged"], run["lm_cycles"], run["sweeps"]))
    for prec in ("mixed", "fp64"):
        a, b = res[(prec, 0)], res[(prec, 500)]
        out["%s_team_vs_one" % prec] = {"status_equal": bool(np.array_equal(a["status"], b["status"])),
                                        "max_dxy_px": float(np.hypot(a["x"] - b["x"], a["y"] - b["y"]).max())}
    a, b = res[("mixed", 0)], res[("fp64", 0)]
    d = np.hypot(a["x"] - b["x"], a["y"] - b["y"])
    ok = (a["status"] == D.CONVERGED) & (b["status"] == D.CONVERGED)
    out["mixed_vs_fp64"] = {"flag_agreement": float((a["status"] == b["status"]).mean()),
                            "frac_dxy_within_1nm": float((d[ok] < TOL_MIXED_XY_PX).mean()),
                            "dxy_px_p50": float(np.percentile(d[ok], 50)), "dxy_px_p99": float(np.percentile(d[ok], 99))}
    return out


def serve_leg(args, precision, log):
    """
    Continuous-batching server (slots of 1024^2 / 4000 peaks) under 16
    simulated cameras with Poisson arrivals, sensors of 256^2, 512^2 or
    1024^2 and ~1000 emitters per 512^2, Z model; p50/p99 latency and fits/s
    per offered load.  Two set-ups: throughput (8 workers x 32 slots, one
    workgroup per frame) and latency (args.serve_workers workers x
    args.serve_slots slots, team fitter with one workgroup per
    args.serve_team peaks: a 1024^2 frame gets 16 CUs, and few workers keep
    their launches on separate hardware queues).
    """
    from storm_analysis_amd.sa_library import dao_serve
    from storm_analysis_amd.diagnostics import serve_sim
    pools = serve_sim.make_pools((256, 512, 1024))
    slots = args.resident_slots or None  # None: one per resident workgroup
    setups = {"resident": dict(slots=slots, margin=2),
              "resident_team": dict(slots=slots, margin=1, team=args.resident_team),
              "throughput": dict(workers=8, slots=32, team=0, margin=2),
              "latency": dict(workers=args.serve_workers, slots=args.serve_slots, team=args.serve_team, margin=1)}
    setups = {k: v for k, v in setups.items() if k in args.serve_setups.split(",")}
    base = dict(max_shape=(1024, 1024), max_peaks=4000, model="Z", roi_size=ROI, precision=precision, **zkw())
    out = {"server": {"max_shape": [1024, 1024], "max_peaks": 4000, "precision": precision, "setups": setups},
           "cameras": 16, "rates_fps": [50, 400], "sizes": [256, 512, 1024], "runs": []}

    def server(name, kw):
        # resident: frames admitted into one running kernel (ResidentServer);
        # the others: a launch per batch of queued frames (FitServer)
        return dao_serve.ResidentServer(**kw) if name.startswith("resident") else dao_serve.FitServer(**kw)

    for name, su in setups.items():
        srv_kw = dict(base, **su)
        srv = server(name, srv_kw)  # warm the kernels once
        serve_sim.run(srv, duration_s=0.2, load=0.05, pools=pools)
        srv.close()
        for load in [float(x) for x in args.serve_loads.split(",") if x]:
            srv = server(name, srv_kw)
            try:
                summ = serve_sim.run(srv, duration_s=args.serve_seconds, load=load, pools=pools)
            finally:
                srv.close()
            summ["setup"] = name
            log("serving %s load %.2f: %d frames, p50 %.1f ms, p99 %.1f ms, %.0f fits/s (offered %.0f emitters/s)%s"
                % (name, load, summ["requests"], summ["p50_ms"], summ["p99_ms"], summ["fits_per_s"],
                   summ["offered_emitters_per_s"],
                   "".join("; %s p50 %.1f p99 %.1f ms" % (k, v["p50_ms"], v["p99_ms"])
                           for k, v in summ.get("by_size", {}).items())))
            out["runs"].append(summ)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the fitter has no multi-GPU path (frames would be "
              "sharded across GPUs with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    import torch
    __graft_entry__.load_package()
    from storm_analysis_amd.sa_library import dao_fit_c as D
    from storm_analysis_amd.diagnostics import synth_gpu

    dev = "cuda:0"
    F, S, N = args.frames, args.size, args.emitters
    t_gen = time.perf_counter()
    log("generating %d frames of %dx%d with %d emitters on %s" % (F, S, S, N, dev))
    images, backgrounds, rows, counts, truth = synth_gpu.frames_z(F, size=S, n=N, device=dev)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t_gen
    counts_np = counts.numpy()

    # host copies of the inputs for the host-to-host rates (pinned, as a
    # camera pipeline's staging buffers would be)
    host = None
    if not args.no_host:
        raw = (images + 100.0).to(torch.uint16)
        host = {"images": images.cpu().pin_memory(), "backgrounds": backgrounds.cpu().pin_memory(),
                "rows": rows.cpu().pin_memory(), "raw": raw.cpu().pin_memory()}
        extra = host["backgrounds"].numel() * 8 + host["rows"].numel() * 8
        host["bytes"] = {"float64": host["images"].numel() * 8 + extra, "uint16": host["raw"].numel() * 2 + extra}
        del raw
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    meas, res = {}, {}
    for p in precisions:
        meas[p], res[p] = run_batched(p, args, images, backgrounds, rows, counts_np, host)
    head = precisions[0]

    # plain path: one frame at a time through the single-frame C-ABI (FP64)
    host_images = images[:args.plain_frames].cpu().numpy()
    host_bgs = backgrounds[:args.plain_frames].cpu().numpy()
    host_rows = rows.cpu().numpy()
    fac = lambda: D.MultiFitterZ(roi_size=ROI, **zkw())  # noqa: E731
    lat, plain_out, plain_conv = [], [], 0
    if args.plain_frames > 0:
        fit_plain(fac, host_images[0], host_bgs[0], frame_rows(host_rows, counts_np, 0))  # warm
    t1 = time.perf_counter()
    for f in range(args.plain_frames):
        ta = time.perf_counter()
        o = fit_plain(fac, host_images[f], host_bgs[f], frame_rows(host_rows, counts_np, f))
        lat.append(time.perf_counter() - ta)
        plain_conv += int((o["status"] == D.CONVERGED).sum())
        plain_out.append(o)
    plain_dt = time.perf_counter() - t1
    plain_fps = plain_conv / plain_dt if plain_dt > 0 else 0.0
    lat_ms = numpy.array(lat) * 1e3 if lat else numpy.zeros(1)
    log("plain: %d frames, p99 %.2f ms" % (args.plain_frames, numpy.percentile(lat_ms, 99)))

    # accuracy: batched FP64 against plain (same arithmetic), mixed against FP64
    accuracy = {}
    if "fp64" in res:
        out64, st64, per_frame = res["fp64"]
        n_plain = int(numpy.sum(per_frame[:args.plain_frames]))
        if n_plain:
            prow, pst = result_rows(plain_out)
            accuracy["fp64_batched_vs_plain"] = agreement((out64[:n_plain], st64[:n_plain]), (prow, pst), D)
        if "mixed" in res:
            outm, stm, _ = res["mixed"]
            accuracy["mixed_vs_fp64_all_frames"] = agreement((outm, stm), (out64, st64), D)

    cpu = None
    if not args.no_cpu:
        nfr = min(args.cpu_frames, F)
        log("cpu baseline (oracle) on %d frames" % nfr)
        cpu, cpu_out = cpu_baseline(images[:nfr].cpu().numpy(), backgrounds[:nfr].cpu().numpy(), host_rows,
                                    counts_np, nfr)
        orow, ost = result_rows(cpu_out)
        for p in precisions:
            out_p, st_p, per_frame = res[p]
            n = int(numpy.sum(per_frame[:nfr]))
            accuracy["%s_vs_oracle" % p] = agreement((out_p[:n], st_p[:n]), (orow, ost), D)

    latency = None
    if args.latency_frames > 0:
        latency = latency_leg(args, images, backgrounds, rows, counts_np, log)

    uneven = None
    if not args.no_uneven:
        uneven = uneven_leg(log, tune_sweeps=args.tune_sweeps)

    dense = None
    if not args.no_dense:
        dense = dense_leg(log)

    serving = None
    if args.serve_seconds > 0:
        serving = serve_leg(args, "mixed" if args.precision == "both" else head, log)

    ingest = None if args.no_ingest else ingest_leg(log)
    filters = None if args.no_filter else filter_leg(log)
    finder = None if args.no_finder else finder_leg(log)
    mfilter = None if args.no_mfilter else mfilter_leg(log)

    hm = meas[head]
    line = {
        "metric": "converged emitter fits/sec + p99 frame latency; batched vs plain, 1 MI355X",
        "value": hm["fits_per_s"],
        "unit": "converged fits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "mixed: fp32 deviance/J/H terms, fp64 shape/accumulators/solve/chi2",
        "data": "synthetic (seeded GPU generator: Poisson frames, offset 100 ADU, gain 1)",
        "config": {"workload": "3D astigmatic Z model, %d frames of %dx%d, %d emitters/frame, ROI %d, MLE" %
                               (F, S, S, N, ROI),
                   "frames": F, "frame_size": S, "emitters_per_frame": N, "model": "Z", "roi_size": ROI,
                   "precision": head},
        "roofline": hm["roofline"],
        "cpu_baseline": cpu,
        "batched": meas,
        "host_to_host": {p: meas[p].get("host_to_host") for p in meas},
        "plain": {"fits_per_s": plain_fps, "frames": args.plain_frames,
                  "p50_ms": float(numpy.percentile(lat_ms, 50)), "p99_ms": float(numpy.percentile(lat_ms, 99)),
                  "path": "MultiFitterZ, one frame per call, FP64"},
        "batched_vs_plain": hm["fits_per_s"] / plain_fps if plain_fps > 0 else None,
        "batched_vs_plain_by_precision": {p: meas[p]["fits_per_s"] / plain_fps if plain_fps > 0 else None
                                          for p in meas},
        "frame_latency_p99_ms": float(numpy.percentile(lat_ms, 99)),
        "single_frame_latency": latency,
        "accuracy": accuracy,
        "mixed_xy_bar_px": TOL_MIXED_XY_PX,
        "uneven_batch": uneven,
        "dense_2048": dense,
        "serving": serving,
        "ingest": ingest,
        "filters": filters,
        "finder": finder,
        "mfilter": mfilter,
        "converged_per_step": hm["converged_per_step"],
        "generate_s": t_gen,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
