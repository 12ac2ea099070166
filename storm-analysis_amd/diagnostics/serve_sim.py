"""
Serving workload: simulated cameras, each a Poisson arrival process of
frames, replayed in real time against a FitServer.

Each camera has a sensor size drawn from `sizes` and a frame rate drawn from
U(rate_min, rate_max) x load; frames carry the 3D-astigmatism ("Z") model's
emitters at ~1000 per 512 x 512 (250 per 256^2, 4000 per 1024^2).  Frames are
drawn from pools generated on the GPU beforehand (synth_gpu.frames_z), so the
replay feeds device-resident images.
"""
import math
import threading
import time

import numpy

from . import synth_gpu


def make_pools(sizes, per_size=6, seed=20261207, device="cuda:0"):
    pools = {}
    for k, s in enumerate(sizes):
        n = int(round(1000 * (s / 512.0) ** 2))
        images, bgs, rows, counts, _ = synth_gpu.frames_z(per_size, size=s, n=n, seed=seed + k, device=device)
        pools[s] = (images, bgs, rows, n)
    return pools


def schedule(cameras, duration_s, sizes, rate_min, rate_max, load, seed):
    """Arrival times (s from start), camera, size; sorted by time."""
    rng = numpy.random.default_rng(seed)
    # the cameras (sensor size, base rate) do not depend on the load
    cam_size = [int(rng.choice(sizes)) for _ in range(cameras)]
    cam_rate = [rng.uniform(rate_min, rate_max) for _ in range(cameras)]
    events = []
    for c in range(cameras):
        size = cam_size[c]
        rate = cam_rate[c] * load
        t = rng.exponential(1.0 / rate)
        while t < duration_s:
            events.append((t, c, size))
            t += rng.exponential(1.0 / rate)
    events.sort()
    return events


def run(server, cameras=16, duration_s=4.0, sizes=(256, 512, 1024), rate_min=50.0, rate_max=400.0, load=1.0,
        seed=20261208, pools=None, log=None):
    """
    Replays the schedule against `server` (a FitServer) in real time; returns
    the server's summary plus the offered load and how late the feeder ran.
    """
    if pools is None:
        pools = make_pools(sizes)
    events = schedule(cameras, duration_s, sizes, rate_min, rate_max, load, seed)
    reqs = []
    lag = 0.0
    t_begin = time.perf_counter()
    for i, (t, c, s) in enumerate(events):
        now = time.perf_counter() - t_begin
        if t > now:
            time.sleep(t - now)
        else:
            lag = max(lag, now - t)
        images, bgs, rows, n = pools[s]
        j = i % images.shape[0]
        reqs.append(server.submit(images[j], bgs[j], rows[j * n:(j + 1) * n], "hdf5", camera=c,
                                  t_arrival=t_begin + t))
        if log is not None and i % 2000 == 0:
            log("serving: %d of %d frames submitted" % (i, len(events)))
    for r in reqs:
        r.wait(timeout=600)
    summ = server.summary(t_begin=t_begin)
    emitters = sum(int(round(1000 * (s / 512.0) ** 2)) for _, _, s in events)
    summ.update({"offered_frames_per_s": len(events) / duration_s, "offered_emitters_per_s": emitters / duration_s,
                 "duration_s": duration_s, "cameras": cameras, "load": load, "feeder_max_lag_ms": lag * 1e3})
    return summ
