"""
Serving: frames from many acquisitions fitted as they arrive.

The reference fits one frame per MultiFitter, frame after frame
(dao_fit_c.py:257-549, driven by fitting.PeakFitter, fitting.py:701-824).
FitServer keeps a fixed pool of frame slots resident on the GPU -- `workers`
batch fitters of `slots` slots each, every slot sized for the largest frame
it accepts -- and runs continuous batching over it:

- admission: the slot pool is sized from an HBM budget; a request waits in
  the queue until a worker has free slots, and frames larger than a slot are
  refused at submit;
- each worker, on its own host thread and HIP stream, takes whatever
  requests are queued (up to its slot count), loads them into its slots,
  fits them in one persistent launch (MultiFitter.doFit for every slot,
  dao_fit_c.py:332-356) and hands back the results; while one worker fits,
  the others load and fit the next requests, so work from different
  requests overlaps on the device;
- latency = result ready - arrival, per request;
- failures stay with their request: a frame with non-finite or non-positive
  pixels (the MLE deviance takes their logarithm), a non-finite background
  or start peak is refused at submit (that request is done with an error),
  and a frame the fitter reports as failed (FrameFitError) fails only its
  own request; the rest of the launch returns its results.

Frames share no state in the fitter, so a frame's result does not depend on
which other frames share its launch (to the parity bar; bit-identical unless
a reserved box overflowed and a rollback ran).
"""
import collections
import ctypes
import queue
import threading
import time

import numpy
from numpy.ctypeslib import ndpointer

from . import dao_batch
from .dao_fit_c import MultiFitterException, formatPeaksGaussianPSF, lastLibraryError, raiseOnLibraryError

# device bytes per slot pixel (two 32-byte pixel records + the Anscombe
# image) and per slot peak (two 96-byte peak records + scheduling scratch),
# besides the peak's LM turn journal (journal_bytes)
BYTES_PER_PIXEL = 72
BYTES_PER_PEAK = 840
JOURNAL_SNAPSHOTS = 4


def roi_pixels(roi_size):
    """Pixels of the round ROI of a roi_size fitting area (the fitter's ROI table)."""
    c = 0.5 * roi_size
    d = numpy.arange(roi_size) - c + 0.5
    return int(((d[:, None] ** 2 + d[None, :] ** 2) <= c * c).sum())


def journal_bytes(roi_size):
    """Device bytes of one peak's LM turn journal (JOURNAL_SNAPSHOTS ROI snapshots)."""
    per_lane = -(-roi_pixels(roi_size) // 64)
    ppl = next(p for p in (2, 4, 6, 8, 12, 16) if per_lane <= p or p == 16)
    return JOURNAL_SNAPSHOTS * (3 * ppl + 1) * 64 * 8


def slot_bytes(max_shape, max_peaks, roi_size=18):
    """Device bytes one slot of max_shape pixels and max_peaks peaks holds."""
    return max_shape[0] * max_shape[1] * BYTES_PER_PIXEL + max_peaks * (BYTES_PER_PEAK + journal_bytes(roi_size))


def pool_size(workers, slots, max_shape, max_peaks, hbm_budget_gb=None, roi_size=18):
    """(workers, slots) after admission by the HBM budget (fewer slots first)."""
    if hbm_budget_gb is None:
        return workers, slots
    fit = int(hbm_budget_gb * 1e9 // slot_bytes(max_shape, max_peaks, roi_size))
    if fit < 1:
        raise MultiFitterException("HBM budget holds no slot of %s" % (tuple(max_shape),))
    while workers * slots > fit and slots > 1:
        slots -= 1
    return max(1, min(workers, fit // slots)), slots


def _all_finite(a, positive=False):
    if hasattr(a, "is_cuda"):
        import torch
        ok = torch.isfinite(a).all()
        if positive:
            ok = ok & (a > 0).all()
        return bool(ok)
    a = numpy.asarray(a)
    ok = numpy.isfinite(a).all()
    return bool(ok and (not positive or (a > 0).all()))


def _invalid_input(image, background, rows):
    """Why a request cannot be fitted, or None."""
    if not _all_finite(image, positive=True):
        return "the frame holds non-finite or non-positive pixels"
    if not _all_finite(background):
        return "the background holds non-finite values"
    if rows is not None and rows.shape[0] and not _all_finite(rows):
        return "a start peak holds non-finite values"
    return None


def latency_by_size(done):
    """p50 / p99 latency (ms) of the completed requests per frame size."""
    by_size = {}
    for r in done:
        by_size.setdefault("%dx%d" % (r.image.shape[0], r.image.shape[1]), []).append(r.latency() * 1e3)
    return {k: {"frames": len(v), "p50_ms": float(numpy.percentile(v, 50)), "p99_ms": float(numpy.percentile(v, 99))}
            for k, v in sorted(by_size.items())}


class FitRequest(object):
    """One frame to fit; `result` (peaks dict, as BatchFitter.getPeaks) once `done` is set."""
    def __init__(self, image, background, rows, peaks_type, camera, t_arrival):
        self.image = image
        self.background = background
        self.rows = rows
        self.peaks_type = peaks_type
        self.camera = camera
        self.t_arrival = t_arrival
        self.t_start = None
        self.t_done = None
        self.worker = None
        self.batch_size = None
        self.result = None
        self.error = None
        self.done = threading.Event()

    def latency(self):
        return self.t_done - self.t_arrival

    def wait(self, timeout=None):
        if not self.done.wait(timeout):
            raise TimeoutError("fit request not done")
        if self.error is not None:
            raise self.error
        return self.result


class FitServer(object):
    """
    workers x slots frame slots of max_shape (height, width) and max_peaks
    peaks each.  hbm_budget_gb, when given, caps workers x slots to what the
    budget holds.  team > 0 fits each frame with one workgroup per `team`
    peaks (lower latency per frame), margin >= 0 sets the reserved-box
    margin.  The remaining keywords are BatchFitter's (model, precision,
    roi_size, wx_params, ...).
    """
    def __init__(self, workers=4, slots=32, max_shape=(1024, 1024), max_peaks=4000, hbm_budget_gb=None,
                 max_iterations=2000, team=0, margin=-1, wait_limit=0, **fitter_kw):
        self.max_shape = tuple(max_shape)
        self.max_peaks = max_peaks
        self.max_iterations = max_iterations
        roi_size = fitter_kw.get("roi_size", 18)
        workers, slots = pool_size(workers, slots, max_shape, max_peaks, hbm_budget_gb, roi_size)
        self.workers = workers
        self.slots = slots
        self.slot_bytes = slot_bytes(max_shape, max_peaks, roi_size)
        self.fitters = [dao_batch.BatchFitter([self.max_shape] * slots, [max_peaks] * slots, **fitter_kw)
                        for _ in range(workers)]
        for bf in self.fitters:
            bf.setTuning(0, margin, team)
            bf.setWaitLimit(wait_limit)
        self.queue = queue.Queue()
        self.completed = []
        self._lock = threading.Lock()
        self.launches = 0
        self.threads = [threading.Thread(target=self._work, args=(w,), daemon=True) for w in range(workers)]
        for t in self.threads:
            t.start()

    def submit(self, image, background, rows, peaks_type="hdf5", camera=0, t_arrival=None):
        """Queue one frame; returns its FitRequest."""
        h, w = image.shape[0], image.shape[1]
        if h * w > self.max_shape[0] * self.max_shape[1] or h > self.max_shape[0] or w > self.max_shape[1]:
            raise MultiFitterException("frame %dx%d exceeds the slot size %s" % (h, w, self.max_shape))
        if rows is not None and rows.shape[0] > self.max_peaks:
            raise MultiFitterException("%d peaks exceed the slot's %d" % (rows.shape[0], self.max_peaks))
        r = FitRequest(image, background, rows, peaks_type, camera,
                       time.perf_counter() if t_arrival is None else t_arrival)
        bad = _invalid_input(image, background, rows)
        if bad is not None:
            r.error = MultiFitterException("request refused: " + bad)
            r.t_start = r.t_done = time.perf_counter()
            r.batch_size = 0
            with self._lock:
                self.completed.append(r)
            r.done.set()
            return r
        self.queue.put(r)
        return r

    def setWaitLimit(self, polls):
        """Polls after which a scheduler wait gives up and fails its frame's request (0: default)."""
        for bf in self.fitters:
            bf.setWaitLimit(polls)

    def _work(self, w):
        bf = self.fitters[w]
        used = 0
        while True:
            first = self.queue.get()
            if first is None:
                return
            batch = [first]
            while len(batch) < self.slots:
                try:
                    r = self.queue.get_nowait()
                except queue.Empty:
                    break
                if r is None:
                    self.queue.put(None)
                    break
                batch.append(r)
            t0 = time.perf_counter()
            try:
                failed = set()
                for r in batch:
                    r.t_start = t0
                # one load call per peak format (normally one)
                for ptype in sorted(set(r.peaks_type for r in batch)):
                    ks = [k for k, r in enumerate(batch) if r.peaks_type == ptype]
                    bf.loadFrames(ks, [batch[k].image for k in ks], [batch[k].background for k in ks],
                                  [batch[k].rows for k in ks], ptype)
                bf.clearFrames(list(range(len(batch), used)))
                used = len(batch)
                try:
                    bf.doFit(self.max_iterations)
                except dao_batch.FrameFitError as e:
                    # only these slots' requests fail; the others were fitted
                    failed = set(e.frames)
                    for k in failed:
                        if k < len(batch):
                            batch[k].error = e
                rows, status, counts = bf.getAllPeaks()
                at = numpy.concatenate([[0], numpy.cumsum(counts)])
                for k, r in enumerate(batch):
                    if k in failed:
                        continue
                    a, b = int(at[k]), int(at[k + 1])
                    res = {c: rows[a:b, i].copy() for i, c in enumerate(dao_batch.PEAK_COLUMNS)}
                    res["iterations"] = res["iterations"].astype(numpy.int32)
                    res["status"] = status[a:b].copy()
                    r.result = res
            except Exception as e:  # the launch itself failed: every request of it
                for r in batch:
                    if r.error is None:
                        r.error = e
            t1 = time.perf_counter()
            with self._lock:
                self.launches += 1
                for r in batch:
                    r.t_done = t1
                    r.worker = w
                    r.batch_size = len(batch)
                    self.completed.append(r)
            for r in batch:
                r.done.set()

    def close(self):
        for _ in self.threads:
            self.queue.put(None)
        for t in self.threads:
            t.join()
        for bf in self.fitters:
            bf.cleanup()
        self.fitters = []

    def summary(self, t_begin=None):
        """Latency percentiles (ms), throughput and batching figures of the completed requests."""
        with self._lock:
            done = list(self.completed)
        if not done:
            return {"requests": 0}
        lat = numpy.array([r.latency() for r in done]) * 1e3
        t0 = min(r.t_arrival for r in done) if t_begin is None else t_begin
        t1 = max(r.t_done for r in done)
        conv = sum(int((r.result["status"] == 1).sum()) for r in done if r.result is not None)
        return {"requests": len(done), "p50_ms": float(numpy.percentile(lat, 50)),
                "p99_ms": float(numpy.percentile(lat, 99)), "max_ms": float(lat.max()),
                "by_size": latency_by_size(done),
                "converged": conv, "fits_per_s": conv / (t1 - t0), "frames_per_s": len(done) / (t1 - t0),
                "launches": self.launches, "mean_batch": float(numpy.mean([r.batch_size for r in done])),
                "workers": self.workers, "slots": self.slots, "slot_bytes": self.slot_bytes,
                "errors": sum(1 for r in done if r.error is not None)}


# ---------------------------------------------------------------------------
# the resident fitter: frames admitted into a running kernel
# ---------------------------------------------------------------------------

_res_lib = None


def loadResidentC():
    """libdao_fit.so with the daoResident* entry points typed (include/dao_resident_amd.h)."""
    global _res_lib
    if _res_lib is None:
        lib = dao_batch.loadBatchC()
        vp = ctypes.c_void_p
        i32p = ndpointer(dtype=numpy.int32)
        f64p = ndpointer(dtype=numpy.float64)
        lib.daoResidentCreate.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        lib.daoResidentCreate.restype = vp
        lib.daoResidentDestroy.argtypes = [vp]
        lib.daoResidentSubmit.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, ctypes.c_int,
                                          ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
        lib.daoResidentSubmit.restype = ctypes.c_int
        lib.daoResidentWait.argtypes = [vp, ctypes.c_double, ctypes.c_int, i32p]
        lib.daoResidentWait.restype = ctypes.c_int
        lib.daoResidentResult.argtypes = [vp, ctypes.c_int, f64p, i32p, i32p, f64p]
        lib.daoResidentResult.restype = ctypes.c_int
        lib.daoResidentStats.argtypes = [vp, ndpointer(dtype=numpy.int64)]
        _res_lib = raiseOnLibraryError(lib, ("daoResidentSubmit", "daoResidentWait"))
    return _res_lib


def resident_workgroups():
    """Workgroups of a resident kernel by default: one per compute unit but 8."""
    import torch
    return max(1, torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count - 8)


class ResidentServer(object):
    """
    Continuous batching inside one running kernel.  `slots` frame slots of
    max_shape (height, width) and max_peaks peaks (fewer if hbm_budget_gb
    holds fewer) on one device -- by default one per resident workgroup, so
    that no request waits for a slot while a workgroup is free; a resident kernel (include/dao_resident_amd.h)
    whose workgroups take each request as soon as one is free, fit it
    (new image, background, start peaks, doFit) and write its peaks to
    pinned host memory.  A request waits in a host queue only while every
    slot is taken; submitting is a few host memory writes, with no launch.
    team > 0 fits a frame of n peaks with ceil(n / team) workgroups (at
    most 64) for lower latency; 0 with one workgroup per frame.
    Frames share no state: each result equals the frame fitted alone
    (BatchFitter) bit for bit.  MLE only; the remaining keywords are
    BatchFitter's (model, precision, roi_size, wx_params, ...).
    """
    def __init__(self, slots=None, max_shape=(1024, 1024), max_peaks=4000, hbm_budget_gb=None, max_iterations=2000,
                 margin=-1, workgroups=0, idle_seconds=0.5, wait_limit=0, team=0, **fitter_kw):
        self.clib = loadResidentC()
        self.max_shape = tuple(max_shape)
        self.max_peaks = max_peaks
        roi_size = fitter_kw.get("roi_size", 18)
        if slots is None:
            slots = workgroups if workgroups > 0 else resident_workgroups()
        _, slots = pool_size(1, slots, max_shape, max_peaks, hbm_budget_gb, roi_size)
        self.slots = slots
        self.workers = 1
        self.slot_bytes = slot_bytes(max_shape, max_peaks, roi_size)
        self.bf = dao_batch.BatchFitter([self.max_shape] * slots, [max_peaks] * slots, **fitter_kw)
        self.bf.setTuning(0, margin, 0)
        self.bf.setWaitLimit(wait_limit)
        self.team = team
        self.handle = self.clib.daoResidentCreate(self.bf.handle, workgroups, max_iterations, idle_seconds, team)
        if not self.handle:
            code, msg = lastLibraryError(self.clib)
            self.bf.cleanup()
            raise MultiFitterException("daoResidentCreate failed (code %d): %s" % (code, msg))
        self._lock = threading.Lock()
        self.free = collections.deque(range(slots))
        self.pending = collections.deque()
        self.inflight = {}
        self.completed = []
        self.launches = 0
        self._stop = False
        self._poller = threading.Thread(target=self._poll, daemon=True)
        self._poller.start()

    def submit(self, image, background, rows, peaks_type="hdf5", camera=0, t_arrival=None):
        """Queue one frame; returns its FitRequest (results as BatchFitter.getPeaks)."""
        h, w = image.shape[0], image.shape[1]
        if h > self.max_shape[0] or w > self.max_shape[1]:
            raise MultiFitterException("frame %dx%d exceeds the slot size %s" % (h, w, self.max_shape))
        if isinstance(rows, dict):
            rows = formatPeaksGaussianPSF(rows, peaks_type)
        if rows is not None and rows.shape[0] > self.max_peaks:
            raise MultiFitterException("%d peaks exceed the slot's %d" % (rows.shape[0], self.max_peaks))
        r = FitRequest(image, background, rows, peaks_type, camera,
                       time.perf_counter() if t_arrival is None else t_arrival)
        # device arrays are checked by the kernel as it loads them (a check
        # here would be a kernel of its own, queued behind the resident one)
        bad = None if hasattr(image, "is_cuda") and image.is_cuda else _invalid_input(image, background, rows)
        if bad is not None:
            r.error = MultiFitterException("request refused: " + bad)
            r.t_start = r.t_done = time.perf_counter()
            r.batch_size = 0
            with self._lock:
                self.completed.append(r)
            r.done.set()
            return r
        with self._lock:
            if self.free:
                slot = self.free.popleft()
                if not self._start(slot, r):
                    self.free.append(slot)
            else:
                self.pending.append(r)
        return r

    def _next(self, slot):
        # with self._lock held: the slot's next request, or back to the pool
        while self.pending:
            if self._start(slot, self.pending.popleft()):
                return
        self.free.append(slot)

    def setWaitLimit(self, polls):
        """Polls after which a scheduler wait gives up and fails its frame's request (0: default)."""
        self.bf.setWaitLimit(polls)

    def _start(self, slot, r):
        # with self._lock held; False (request failed, slot unused) if refused
        img, dev_i, ki = dao_batch._ptr(r.image, numpy.float64)
        bg, dev_b, kb = dao_batch._ptr(r.background, numpy.float64)
        n = 0 if r.rows is None else int(r.rows.shape[0])
        if n > 0:
            rw, dev_r, kr = dao_batch._ptr(r.rows, numpy.float64)
        else:
            rw, dev_r, kr = None, dev_i, None
        if not (dev_i == dev_b == dev_r):
            r.error = MultiFitterException("image, background and peaks must all be host or all device arrays")
            r.t_start = r.t_done = time.perf_counter()
            self.completed.append(r)
            r.done.set()
            return False
        r._keep = (ki, kb, kr)
        ticket = ctypes.c_int64(0)
        r.t_start = time.perf_counter()
        r.worker = 0
        r.batch_size = 1
        try:
            self.clib.daoResidentSubmit(self.handle, slot, r.image.shape[1], r.image.shape[0], img, bg, rw, n,
                                        r.peaks_type.encode(), dev_i, ctypes.byref(ticket))
        except MultiFitterException as e:  # this request only; the slot stays free
            r.error = e
            r.t_done = time.perf_counter()
            r._keep = None
            self.completed.append(r)
            r.done.set()
            return False
        self.inflight[slot] = r
        return True

    def _finish(self, slot):
        # with self._lock held
        r = self.inflight.pop(slot)
        cap = self.max_peaks
        out = numpy.zeros((max(cap, 1), 9))
        status = numpy.zeros(max(cap, 1), dtype=numpy.int32)
        counters = numpy.zeros(8, dtype=numpy.int32)
        info = numpy.zeros(4)
        n = self.clib.daoResidentResult(self.handle, slot, out, status, counters, info)
        if n < 0:
            code, msg = lastLibraryError(self.clib)
            r.error = dao_batch.FrameFitError("resident fit failed (code %d): %s" % (n, msg), [slot])
        else:
            res = {c: out[:n, i].copy() for i, c in enumerate(dao_batch.PEAK_COLUMNS)}
            res["iterations"] = res["iterations"].astype(numpy.int32)
            res["status"] = status[:n].copy()
            r.result = res
        r.counters = dict(zip(dao_batch.COUNTER_NAMES, counters.tolist()))
        r.sweeps = int(info[0])
        r.device_seconds = float(info[1])
        r.load_seconds = float(info[2])
        r.t_done = time.perf_counter()
        r._keep = None
        self.completed.append(r)
        r.done.set()
        self._next(slot)

    def _poll(self):
        slots = numpy.zeros(self.slots, dtype=numpy.int32)
        while True:
            with self._lock:
                if self._stop and not self.inflight and not self.pending:
                    return
            try:
                k = self.clib.daoResidentWait(self.handle, 0.05, self.slots, slots)
            except MultiFitterException as e:
                with self._lock:
                    for r in list(self.inflight.values()) + list(self.pending):
                        r.error = e
                        r.t_done = time.perf_counter()
                        self.completed.append(r)
                        r.done.set()
                    self.inflight.clear()
                    self.pending.clear()
                return
            if k > 0:
                with self._lock:
                    for slot in slots[:k].tolist():
                        self._finish(slot)

    def stats(self):
        out = numpy.zeros(2, dtype=numpy.int64)
        self.clib.daoResidentStats(self.handle, out)
        return {"requests": int(out[0]), "kernel_launches": int(out[1])}

    def close(self):
        with self._lock:
            self._stop = True
        self._poller.join()
        if self.handle:
            self.launches = self.stats()["kernel_launches"]
            self.clib.daoResidentDestroy(self.handle)
            self.handle = None
        self.bf.cleanup()

    def summary(self, t_begin=None):
        """Latency percentiles (ms), overall and per frame size, and throughput of the completed requests."""
        with self._lock:
            done = list(self.completed)
        if not done:
            return {"requests": 0}
        lat = numpy.array([r.latency() for r in done]) * 1e3
        t0 = min(r.t_arrival for r in done) if t_begin is None else t_begin
        t1 = max(r.t_done for r in done)
        conv = sum(int((r.result["status"] == 1).sum()) for r in done if r.result is not None)
        dev = [r.device_seconds * 1e3 for r in done if getattr(r, "device_seconds", None) is not None]
        return {"requests": len(done), "p50_ms": float(numpy.percentile(lat, 50)),
                "p99_ms": float(numpy.percentile(lat, 99)), "max_ms": float(lat.max()),
                "by_size": latency_by_size(done),
                "device_fit_ms_p50": float(numpy.percentile(dev, 50)) if dev else None,
                "device_fit_ms_p99": float(numpy.percentile(dev, 99)) if dev else None,
                "converged": conv, "fits_per_s": conv / (t1 - t0), "frames_per_s": len(done) / (t1 - t0),
                "kernel_launches": self.stats()["kernel_launches"] if self.handle else self.launches,
                "slots": self.slots, "slot_bytes": self.slot_bytes,
                "errors": sum(1 for r in done if r.error is not None)}
