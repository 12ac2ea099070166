// Non-template kernels of the MI355X fitter and the host-side launchers
// (the device code they share is in dao_ops.h; the k_fit_frames
// instantiations are compiled in dao_fit_fp64.hip and dao_fit_mixed.hip).
#include "dao_ops.h"

namespace dao {

// ---------------------------------------------------------------------------
// multi-kernel executor: predecessor lists and rounds
// ---------------------------------------------------------------------------


__device__ inline bool is_big(const Sched& sc, int4 b) {
  return (b.z - b.x) > sc.cell || (b.w - b.y) > sc.cell;
}

// Cell of a box (by its origin), clamped into the frame's grid: an empty box
// at the image edge may start at x = W.
__device__ inline int cell_of(const Sched& sc, const Frame& fr, int4 b) {
  int cx = min(max(b.x / sc.cell, 0), fr.ncx - 1);
  int cy = min(max(b.y / sc.cell, 0), fr.ncy - 1);
  return fr.cell0 + cy * fr.ncx + cx;
}

__device__ inline int op_frame(const Sched& sc, const Peaks& pk, int i) { return pk.rec[sc.op_peak[i]].frame; }

__global__ void k_bin_count(Sched sc, const Frame* frames, const Peaks pk) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int4 b = sc.box[i];
  if (is_big(sc, b)) {
    int q = atomicAdd(sc.n_big, 1);
    sc.big_ops[q] = i;
    return;
  }
  const Frame fr = frames[op_frame(sc, pk, i)];
  atomicAdd(&sc.cell_fill[cell_of(sc, fr, b)], 1);
}

// exclusive scan of cell_fill into cell_start (one block), zeroing cell_fill
__global__ void k_scan_cells(Sched sc, int ncells) {
  __shared__ int part[1024];
  int t = threadIdx.x;
  int per = (ncells + blockDim.x - 1) / blockDim.x;
  int a = t * per, b = a + per < ncells ? a + per : ncells;
  int s = 0;
  for (int i = a; i < b; i++) s += sc.cell_fill[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < (int)blockDim.x; o <<= 1) {
    int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int i = a; i < b; i++) {
    sc.cell_start[i] = run;
    run += sc.cell_fill[i];
    sc.cell_fill[i] = 0;
  }
  if (t == (int)blockDim.x - 1) sc.cell_start[ncells] = part[t];
}

__global__ void k_bin_scatter(Sched sc, const Frame* frames, const Peaks pk) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int4 b = sc.box[i];
  if (is_big(sc, b)) return;
  const Frame fr = frames[op_frame(sc, pk, i)];
  int c = cell_of(sc, fr, b);
  int q = atomicAdd(&sc.cell_fill[c], 1);
  int pos = sc.cell_start[c] + q;
  if (pos < sc.nops) sc.cell_ops[pos] = i;
}

// Predecessors of op i: earlier ops of the same frame with overlapping boxes
// (ops are ordered by frame, then peak index).
__global__ void k_preds(Sched sc, const Frame* frames, const Peaks pk) {
  int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  int lane = lane_id();
  if (i >= sc.nops) return;
  int fi = op_frame(sc, pk, i);
  const Frame fr = frames[fi];
  int4 b = sc.box[i];
  int count = 0;
  auto take = [&](int j) {
    bool hit = (j < i) && (op_frame(sc, pk, j) == fi) && boxes_overlap(b, sc.box[j]);
    unsigned long long mk = __ballot(hit);
    if (hit) {
      int pos = count + __popcll(mk & ((1ull << lane) - 1ull));
      if (pos < kMaxPred) sc.pred[(long long)i * kMaxPred + pos] = j;
    }
    count += __popcll(mk);
  };
  int cx0 = (b.x - sc.cell + 1) >= 0 ? (b.x - sc.cell + 1) / sc.cell : 0;
  int cy0 = (b.y - sc.cell + 1) >= 0 ? (b.y - sc.cell + 1) / sc.cell : 0;
  int cx1 = min((b.z - 1) / sc.cell, fr.ncx - 1);
  int cy1 = min((b.w - 1) / sc.cell, fr.ncy - 1);
  for (int cy = cy0; cy <= cy1; cy++) {
    for (int cx = cx0; cx <= cx1; cx++) {
      int c = fr.cell0 + cy * fr.ncx + cx;
      int s0 = sc.cell_start[c], s1 = sc.cell_start[c + 1];
      for (int q = s0; q < s1; q += kWave) take(q + lane < s1 ? sc.cell_ops[q + lane] : i);
    }
  }
  int nb = *sc.n_big;
  for (int q = 0; q < nb; q += kWave) take(q + lane < nb ? sc.big_ops[q + lane] : i);
  if (lane == 0) {
    sc.pred_n[i] = count;
    sc.done[i] = -1;
  }
}

// Op i may run in round r when every predecessor finished in an earlier
// round (a predecessor finishing in this same launch counts as not done).
__device__ bool op_ready(const Sched& sc, const Frame* frames, const Peaks& pk, int i, int round) {
  int lane = lane_id();
  int n = sc.pred_n[i];
  bool wait = false;
  auto pending = [&](int j) {
    int d = __hip_atomic_load(&sc.done[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return d < 0 || d >= round;
  };
  if (n <= kMaxPred) {
    for (int q = lane; q < n; q += kWave)
      if (pending(sc.pred[(long long)i * kMaxPred + q])) wait = true;
  } else {
    // too many to store: rescan the cells
    int fi = op_frame(sc, pk, i);
    const Frame fr = frames[fi];
    int4 b = sc.box[i];
    auto chk = [&](int j) {
      if (j < i && op_frame(sc, pk, j) == fi && boxes_overlap(b, sc.box[j]) && pending(j)) wait = true;
    };
    int cx0 = (b.x - sc.cell + 1) >= 0 ? (b.x - sc.cell + 1) / sc.cell : 0;
    int cy0 = (b.y - sc.cell + 1) >= 0 ? (b.y - sc.cell + 1) / sc.cell : 0;
    int cx1 = min((b.z - 1) / sc.cell, fr.ncx - 1);
    int cy1 = min((b.w - 1) / sc.cell, fr.ncy - 1);
    for (int cy = cy0; cy <= cy1; cy++)
      for (int cx = cx0; cx <= cx1; cx++) {
        int c = fr.cell0 + cy * fr.ncx + cx;
        for (int q = sc.cell_start[c] + lane; q < sc.cell_start[c + 1]; q += kWave) chk(sc.cell_ops[q]);
      }
    int nb = *sc.n_big;
    for (int q = lane; q < nb; q += kWave) chk(sc.big_ops[q]);
  }
  return !__any(wait);
}

__device__ inline void op_finish(const Sched& sc, int i, int round, int flags, const int* d) {
  if (lane_id() == 0) {
#pragma unroll
    for (int q = 0; q < C_COUNT; q++) sc.op_delta[(long long)i * C_COUNT + q] = d[q];
    sc.op_flags[i] = flags;
    __hip_atomic_store(&sc.done[i], round, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicSub(sc.remaining, 1);
  }
}

// One executor round: every pending op whose predecessors are done runs.
template <int PPL, int SHAPE>
__global__ __launch_bounds__(256) void k_round(OpArgs a, int kind, const double* rows, int fmt) {
  __shared__ double lds[kWavesPerBlock][RoiLds<PPL>::kDoubles];
  const int wv = threadIdx.x >> 6;
  const int i = blockIdx.x * kWavesPerBlock + wv;
  if (i >= a.sc.nops) return;
  if (a.sc.done[i] >= 0) return;
  if (!op_ready(a.sc, a.frames, a.pk, i, a.round)) return;
  const int p = a.sc.op_peak[i];
  const int fi = a.pk.rec[p].frame;
  const Frame fr = a.frames[fi];
  int d[C_COUNT];
#pragma unroll
  for (int q = 0; q < C_COUNT; q++) d[q] = 0;
  int r = 1;
  double* slice = lds[wv];
  switch (kind) {
    case OP_LM: {
      int nsnap = 0;
      r = lm_turn<PPL, SHAPE>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, a.sc.box[i], d, nullptr, nsnap);
      if (r == 2 && lane_id() == 0) {
        // re-run in the next attempt with the whole frame as its box
        a.sc.esc_box[i] = make_int4(0, 0, fr.W, fr.H);
        a.sc.frame_bad[fi] = 1;
      }
      break;
    }
    case OP_RECENTER:
      recenter_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, d);
      break;
    case OP_NEWPEAK:
      newpeak_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, rows + (fmt == PF_TEXT ? 7LL : 4LL) * i, fmt,
                        i, d);
      break;
    case OP_SUBTRACT:
      subtract_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p);
      break;
    default:
      peakerror_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, d);
      break;
  }
  op_finish(a.sc, i, a.round, r, d);
}

// ---------------------------------------------------------------------------
// per-pixel and per-peak kernels
// ---------------------------------------------------------------------------

// mFitNewImage (multi_fit.c:1382-1408) for pixels [pix0, pix0 + n).
__global__ void k_new_image(Pix px, long long pix0, long long n, const double* img) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long k = pix0 + i;
  PixRO r = px.ro[k];
  r.x = img[i] * r.rqe + r.scm;
  px.ro[k] = r;
  PixRW w;
  w.f = 0.0;
  w.bg = 0.0;
  w.tfi = 0.0;
  w.cnt = 0;
  w.stale = 1;
  px.rw[k] = w;
}

// mFitNewBackground (multi_fit.c:1360-1370).
__global__ void k_new_background(Pix px, long long pix0, long long n, const double* bg) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long k = pix0 + i;
  px.ro[k].bgest = bg[i] * px.ro[k].rqe;
  px.rw[k].stale = 1;
}

// mFitNewImage and mFitNewBackground together (k_new_image, then
// k_new_background, in one pass over the pixel records).
__global__ void k_new_frame_pixels(Pix px, long long pix0, long long n, const double* img, const double* bg) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long k = pix0 + i;
  PixRO r = px.ro[k];
  r.x = img[i] * r.rqe + r.scm;
  r.bgest = bg[i] * r.rqe;
  px.ro[k] = r;
  PixRW w;
  w.f = 0.0;
  w.bg = 0.0;
  w.tfi = 0.0;
  w.cnt = 0;
  w.stale = 1;
  px.rw[k] = w;
}

// Calibration of pixels [pix0, pix0 + n) (mFitInitialize, multi_fit.c:971-981);
// the next image is scaled with it.
__global__ void k_set_calibration(Pix px, long long pix0, long long n, const double* rqe, const double* scm) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  px.ro[pix0 + i].rqe = rqe[i];
  px.ro[pix0 + i].scm = scm[i];
}

// Uniform calibration of pixels [pix0, pix0 + n).
__global__ void k_set_calibration_const(Pix px, long long pix0, long long n, double rqe, double scm) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  px.ro[pix0 + i].rqe = rqe;
  px.ro[pix0 + i].scm = scm;
}

// mFitAnscombeTransformImage (multi_fit.c:90-97).
__global__ void k_anscombe(Pix px, long long pix0, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  px.asx[pix0 + i] = anscombe(px.ro[pix0 + i].x);
}

// mFitGetFitImage (f / rqe) and mFitGetResidual (x - f).
__global__ void k_fit_image(Pix px, long long pix0, long long n, double* out, int residual) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long k = pix0 + i;
  out[i] = residual ? px.ro[k].x - px.rw[k].f : px.rw[k].f / px.ro[k].rqe;
}

// Reserved boxes of LM ops: the ROI at the peak's origin plus `margin`,
// grown to `override_box` where set; also backs up the peak record.
__global__ void k_box_lm(Sched sc, Peaks pk, const Frame* frames, Model m, const int4* override_box) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int p = sc.op_peak[i];
  const PeakRec r = pk.rec[p];
  const Frame fr = frames[r.frame];
  int4 b;
  b.x = max(0, r.xi - m.margin);
  b.y = max(0, r.yi - m.margin);
  b.z = min(fr.W, r.xi + m.roi + m.margin);
  b.w = min(fr.H, r.yi + m.roi + m.margin);
  if (override_box != nullptr) {
    int4 o = override_box[p];
    if (o.z > o.x) {
      b.x = max(0, min(b.x, o.x));
      b.y = max(0, min(b.y, o.y));
      b.z = min(fr.W, max(b.z, o.z));
      b.w = min(fr.H, max(b.w, o.w));
    }
  }
  sc.box[i] = b;
  pk.bk[p] = r;
  sc.op_flags[i] = 0;
}

__device__ inline int4 clamp_box(int x0, int y0, int x1, int y1, const Frame& fr) {
  int4 b;
  b.x = max(0, min(fr.W, x0));
  b.y = max(0, min(fr.H, y0));
  b.z = max(b.x, min(fr.W, x1));
  b.w = max(b.y, min(fr.H, y1));
  return b;
}

// Box = the ROI at the peak's origin (subtraction, initial errors).
__global__ void k_box_roi(Sched sc, const Peaks pk, const Frame* frames, Model m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  const PeakRec r = pk.rec[sc.op_peak[i]];
  sc.box[i] = clamp_box(r.xi, r.yi, r.xi + m.roi, r.yi + m.roi, frames[r.frame]);
  sc.op_flags[i] = 0;
}

// Box of a new peak from its start row (floor of the start position).
__global__ void k_box_newpeak(Sched sc, const Peaks pk, const Frame* frames, Model m, const double* rows,
                              int fmt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  const Frame fr = frames[pk.rec[sc.op_peak[i]].frame];
  const double* q = rows + (fmt == PF_TEXT ? 7LL : 4LL) * i;
  int xi = (int)floor(q[0] - m.xoff);
  int yi = (int)floor(q[1] - m.xoff);
  sc.box[i] = clamp_box(xi, yi, xi + m.roi, yi + m.roi, fr);
  sc.op_flags[i] = 0;
}

// Recenter: the box covers the old and the new ROI origin.
__global__ void k_box_recenter(Sched sc, const Peaks pk, const Frame* frames, Model m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  const PeakRec r = pk.rec[sc.op_peak[i]];
  int nx = (int)floor(r.prm[P_XC]);
  int ny = (int)floor(r.prm[P_YC]);
  sc.box[i] = clamp_box(min(r.xi, nx), min(r.yi, ny), max(r.xi, nx) + m.roi, max(r.yi, ny) + m.roi,
                        frames[r.frame]);
  sc.op_flags[i] = 0;
}

// Per-frame counter totals of the ops that completed in good frames.
__global__ void k_commit_counters(Sched sc, const Peaks pk, int* frame_counters) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  int fi = op_frame(sc, pk, i);
  if (sc.frame_bad[fi] || (sc.op_flags[i] & 1) == 0) return;
  for (int q = 0; q < C_COUNT; q++) {
    int v = sc.op_delta[(long long)i * C_COUNT + q];
    if (v) atomicAdd(&frame_counters[fi * C_COUNT + q], v);
  }
}

// Escaped ops leave the box they need for the next attempt.
__global__ void k_collect_escapes(Sched sc, const Peaks pk, int4* override_box) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sc.nops) return;
  if (sc.op_flags[i] & 2) override_box[sc.op_peak[i]] = sc.esc_box[i];
}

// Per-peak derived properties: bg_sum (what 0), fg_sum (1), sum (2)
// (multi_fit.c:1518-1632), significance fg_sum / sqrt(bg_sum) (3; the
// "compound" property of dao_fit_c.py:398-439).  Peak i is slot slot0 + i,
// or slots[i] when a slot list is given.
template <int PPL>
__global__ __launch_bounds__(256) void k_peak_sums(Model m, Pix px, Peaks pk, const Frame* frames,
                                                   const int* roi_tab, int slot0, const int* slots, int n, int what,
                                                   double* out) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= n) return;
  const int p = slots ? slots[i] : slot0 + i;
  PeakState w;
  load_state(pk.rec + p, w);
  const Frame fr = frames[pk.rec[p].frame];
  Roi<PPL> c;
  c.init(m, roi_tab, fr, nullptr);
  double t, e;
  lane_shape(m, w, t, e);
  double ps = 0.0, a = 0.0, a2 = 0.0;
#pragma unroll
  for (int s = 0; s < PPL; s++) {
    double ex = __shfl(e, c.rx(s), kWave);
    double ey = __shfl(e, 32 + c.ry(s), kWave);
    if (c.ok(s)) {
      double psf = ey * ex;
      ps += psf;
      int gx = w.xi + c.rx(s), gy = w.yi + c.ry(s);
      if ((what == 0 || what == 3) && gx >= 0 && gy >= 0 && gx < fr.W && gy < fr.H) {
        const PixRO r = px.ro[fr.pix0 + (long long)gy * fr.W + gx];
        double b = (r.x - r.scm) / r.rqe - w.prm[P_HEIGHT] * psf;
        a += b * psf * psf;
      }
      if (what == 1 || what == 3) a2 += psf * psf;
    }
  }
  ps = wsum(ps);
  a = wsum(a);
  a2 = wsum(a2);
  if (lane_id() == 0) {
    const double bg_sum = ((a > 0.0) && (ps > 0.0)) ? a / (ps * ps) : 1.0;
    const double fg_sum = ps > 0.0 ? (w.prm[P_HEIGHT] * a2) / ps : 1.0;
    double v;
    if (what == 0) {
      v = bg_sum;
    } else if (what == 1) {
      v = fg_sum;
    } else if (what == 3) {
      v = fg_sum / sqrt(bg_sum);
    } else {
      v = ps * w.prm[P_HEIGHT];
    }
    out[i] = v;
  }
}

// The slots of every frame's live peaks in frame order (off[f] = the
// frame's first row): one block per frame.
__global__ void k_packed_slots(const Frame* frames, const int* off, int* slots) {
  const Frame fr = frames[blockIdx.x];
  for (int i = threadIdx.x; i < fr.nfit; i += blockDim.x) slots[off[blockIdx.x] + i] = fr.peak0 + i;
}

// mFitSetPeakStatus (multi_fit.c:1809-1831) for every frame: flag[slot]
// marks a live peak that the new status (rows in frame order) makes ERROR
// while it is in the fit image (it must be subtracted first).
__global__ void k_status_flag(Peaks pk, const Frame* frames, const int* off, long long nslots, const int* status,
                              int* flag) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const PeakRec r = pk.rec[t];
  int sel = 0;
  if (r.frame >= 0) {
    const Frame fr = frames[r.frame];
    const long long j = t - fr.peak0;
    if (j >= 0 && j < fr.nfit) sel = (status[off[r.frame] + j] == ST_ERROR && r.added > 0) ? 1 : 0;
  }
  flag[t] = sel;
}

__global__ void k_apply_status(Peaks pk, const Frame* frames, const int* off, long long nslots, const int* status) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const int fi = pk.rec[t].frame;
  if (fi < 0) return;
  const Frame fr = frames[fi];
  const long long j = t - fr.peak0;
  if (j >= 0 && j < fr.nfit) pk.rec[t].status = status[off[fi] + j];
}

// mFitRemoveErrorPeaks / mFitRemoveRunningPeaks (multi_fit.c:1704-1779) for
// every frame at once: one block per frame keeps, in order, the live peaks
// whose status is not `drop` (gathered through `bk`, then copied back) and
// writes the frame's new count; removed RUNNING peaks count as
// n_non_converged.
__global__ __launch_bounds__(1024) void k_compact_frames(Peaks pk, const Frame* frames, int drop, int* counts,
                                                         int* counters) {
  __shared__ int part[1024];
  const int f = blockIdx.x;
  const Frame fr = frames[f];
  const int n = fr.nfit;
  const int t = threadIdx.x;
  const int per = (n + (int)blockDim.x - 1) / (int)blockDim.x;
  const int a = min(t * per, n), b = min(a + per, n);
  int mine = 0;
  for (int i = a; i < b; i++) mine += pk.rec[fr.peak0 + i].status != drop;
  part[t] = mine;
  __syncthreads();
  for (int d = 1; d < (int)blockDim.x; d <<= 1) {
    const int v = t >= d ? part[t - d] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int at = part[t] - mine;
  for (int i = a; i < b; i++) {
    const PeakRec r = pk.rec[fr.peak0 + i];
    if (r.status != drop) pk.bk[fr.peak0 + at++] = r;
  }
  const int kept = part[blockDim.x - 1];
  __syncthreads();
  for (int i = t; i < kept; i += blockDim.x) pk.rec[fr.peak0 + i] = pk.bk[fr.peak0 + i];
  if (t == 0) {
    counts[f] = kept;
    if (drop == ST_RUNNING && n > kept) atomicAdd(&counters[f * C_COUNT + C_NON_CONV], n - kept);
  }
}

// "jacobian" property: the 'b' vector at each peak's current state.
template <int PPL, int SHAPE>
__global__ __launch_bounds__(256) void k_peak_jacobian(Model m, Pix px, Peaks pk, const Frame* frames,
                                                       const int* roi_tab, int slot0, int n, double* out) {
  constexpr int N = SHAPE == SH_2DFIXED ? 4 : (SHAPE == SH_3D ? 6 : 5);
  __shared__ double lds[kWavesPerBlock][RoiLds<PPL>::kDoubles];
  const int wv = threadIdx.x >> 6;
  const int i = blockIdx.x * kWavesPerBlock + wv;
  if (i >= n) return;
  const int p = slot0 + i;
  PeakState w;
  load_state(pk.rec + p, w);
  const Frame fr = frames[pk.rec[p].frame];
  if (!origin_ok(m, fr, w.xi, w.yi)) {
    if (lane_id() == 0)
      for (int q = 0; q < N; q++) out[(long long)i * N + q] = 0.0;
    return;
  }
  Roi<PPL> c;
  c.init(m, roi_tab, fr, lds[wv]);
  c.load(px, fr, w.xi, w.yi);
  c.table(m, w);
  calc_jh<PPL, SHAPE>(m, c, w, px.asx);
  const double* J = c.hj();
  if (lane_id() == 0)
    for (int q = 0; q < N; q++) out[(long long)i * N + q] = J[q];
}

// Gather peak records: slot0+t takes slot0+map[t] (stable compaction).
__global__ void k_gather_peaks(PeakRec* dst, const PeakRec* src, const int* map, int slot0, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  dst[slot0 + t] = src[slot0 + map[t]];
}

__global__ void k_set_status(Peaks pk, int slot0, int n, const int* status) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  pk.rec[slot0 + t].status = status[t];
}

// Empty records owned by `frame` in slots [s0, s1).
__global__ void k_set_owner(Peaks pk, long long s0, long long s1, int frame) {
  long long t = s0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= s1) return;
  PeakRec r;
  for (int q = 0; q < kNFit; q++) r.prm[q] = 0.0;
  r.lam = 1.0;
  r.err = 0.0;
  r.xi = 0;
  r.yi = 0;
  r.status = ST_ERROR;
  r.iter = 0;
  r.added = 0;
  r.frame = frame;
  pk.rec[t] = r;
}

// Selection over peak slots (flag[slot] = 1 when selected).
__global__ void k_flag(Peaks pk, const Frame* frames, long long nslots, int mode, const int* frame_bad, int* flag) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const PeakRec r = pk.rec[t];
  int sel = 0;
  if (r.frame >= 0) {
    const Frame fr = frames[r.frame];
    bool live = (t - fr.peak0) < fr.nfit;
    if (live && r.status == ST_RUNNING) {
      if (mode == SEL_RUNNING) {
        sel = 1;
      } else if (mode == SEL_RUNNING_BAD) {
        sel = frame_bad[r.frame] ? 1 : 0;
      } else {
        // mFitRecenterPeaks test (multi_fit.c:1658-1661)
        double dx = r.prm[P_XC] - r.xi;
        double dy = r.prm[P_YC] - r.yi;
        sel = ((dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25)) ? 1 : 0;
      }
    }
  }
  flag[t] = sel;
}

// Order-preserving compaction of flagged slots into out[], count in *count
// (one block; each thread owns a contiguous chunk).
__global__ void k_compact(const int* flag, long long n, int* out, int* count) {
  __shared__ int part[1024];
  int t = threadIdx.x;
  long long per = (n + blockDim.x - 1) / blockDim.x;
  long long a = t * per, b = a + per < n ? a + per : n;
  int s = 0;
  for (long long i = a; i < b; i++) s += flag[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < (int)blockDim.x; o <<= 1) {
    int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (long long i = a; i < b; i++)
    if (flag[i]) out[run++] = (int)i;
  if (t == (int)blockDim.x - 1) *count = part[t];
}

__global__ void k_fill_i32(int* p, long long n, int v) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ void k_iota(int* p, int n, int base) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = base + t;
}

// Live peaks of every frame as rows of 9 doubles {x, y, z, height,
// background, xsigma, ysigma, error, iterations} + status, frame f's rows
// from offset[f].
__global__ void k_pack_peaks(Peaks pk, const Frame* frames, const int* offset, Model m, long long nslots,
                             double* out, int* status) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const PeakRec r = pk.rec[t];
  if (r.frame < 0) return;
  const Frame fr = frames[r.frame];
  long long j = t - fr.peak0;
  if (j < 0 || j >= fr.nfit) return;
  long long row = offset[r.frame] + j;
  double* o = out + 9 * row;
  o[0] = r.prm[P_XC] + m.xoff;
  o[1] = r.prm[P_YC] + m.xoff;
  o[2] = r.prm[P_ZC] + m.zoff;
  o[3] = r.prm[P_HEIGHT];
  o[4] = r.prm[P_BG];
  o[5] = sqrt(1.0 / (2.0 * r.prm[P_XW]));
  o[6] = sqrt(1.0 / (2.0 * r.prm[P_YW]));
  o[7] = r.err;
  o[8] = (double)r.iter;
  status[row] = r.status;
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------


void launch_op_round(int kind, const OpArgs& a, hipStream_t st, const double* rows, int fmt, int) {
  if (a.sc.nops == 0) return;
  dim3 g = waves_grid(a.sc.nops), b(256);
  if (kind == OP_LM) {
    DAO_PPL_SWITCH(a.m.nroi, DAO_SHAPE_SWITCH(a.m.shape, (k_round<PPL, SHP><<<g, b, 0, st>>>(a, kind, rows, fmt))));
  } else {
    // the shape only matters for LM turns
    DAO_PPL_SWITCH(a.m.nroi, (k_round<PPL, SH_2DFIXED><<<g, b, 0, st>>>(a, kind, rows, fmt)));
  }
}


void launch_binning(const Sched& sc, const Frame* frames, const Peaks& pk, int ncells, hipStream_t st) {
  if (sc.nops == 0) return;
  // cell_fill counts ops per cell, is zeroed by the scan and counts again as
  // the scatter's cursor: clear it first
  k_fill_i32<<<flat_grid(ncells), 256, 0, st>>>(sc.cell_fill, ncells, 0);
  k_bin_count<<<flat_grid(sc.nops), 256, 0, st>>>(sc, frames, pk);
  k_scan_cells<<<1, 1024, 0, st>>>(sc, ncells);
  k_bin_scatter<<<flat_grid(sc.nops), 256, 0, st>>>(sc, frames, pk);
  k_preds<<<waves_grid(sc.nops), 256, 0, st>>>(sc, frames, pk);
}

void launch_box(int kind, const Sched& sc, const Peaks& pk, const Frame* frames, const Model& m,
                const int4* override_box, const double* rows, int fmt, int, hipStream_t st) {
  if (sc.nops == 0) return;
  dim3 g = flat_grid(sc.nops);
  switch (kind) {
    case OP_LM: k_box_lm<<<g, 256, 0, st>>>(sc, pk, frames, m, override_box); break;
    case OP_RECENTER: k_box_recenter<<<g, 256, 0, st>>>(sc, pk, frames, m); break;
    case OP_NEWPEAK: k_box_newpeak<<<g, 256, 0, st>>>(sc, pk, frames, m, rows, fmt); break;
    default: k_box_roi<<<g, 256, 0, st>>>(sc, pk, frames, m); break;
  }
}

void launch_commit(const Sched& sc, const Peaks& pk, int* frame_counters, hipStream_t st) {
  if (sc.nops == 0) return;
  k_commit_counters<<<flat_grid(sc.nops), 256, 0, st>>>(sc, pk, frame_counters);
}

void launch_collect_escapes(const Sched& sc, const Peaks& pk, int4* override_box, hipStream_t st) {
  if (sc.nops == 0) return;
  k_collect_escapes<<<flat_grid(sc.nops), 256, 0, st>>>(sc, pk, override_box);
}

void launch_new_image(const Pix& px, long long pix0, long long n, const double* img, hipStream_t st) {
  k_new_image<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, img);
}
void launch_new_frame_pixels(const Pix& px, long long pix0, long long n, const double* img, const double* bg,
                             hipStream_t st) {
  if (n <= 0) return;
  k_new_frame_pixels<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, img, bg);
}

void launch_new_background(const Pix& px, long long pix0, long long n, const double* bg, hipStream_t st) {
  k_new_background<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, bg);
}
void launch_set_calibration(const Pix& px, long long pix0, long long n, const double* rqe, const double* scm,
                            hipStream_t st) {
  k_set_calibration<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, rqe, scm);
}
void launch_set_calibration_const(const Pix& px, long long pix0, long long n, double rqe, double scm,
                                  hipStream_t st) {
  k_set_calibration_const<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, rqe, scm);
}
void launch_anscombe(const Pix& px, long long pix0, long long n, hipStream_t st) {
  k_anscombe<<<flat_grid(n), 256, 0, st>>>(px, pix0, n);
}
void launch_fit_image(const Pix& px, long long pix0, long long n, double* out, int residual, hipStream_t st) {
  k_fit_image<<<flat_grid(n), 256, 0, st>>>(px, pix0, n, out, residual);
}
void launch_peak_sums(const Model& m, const Pix& px, const Peaks& pk, const Frame* frames, const int* roi_tab,
                      int slot0, int n, int what, double* out, hipStream_t st, const int* slots) {
  if (n == 0) return;
  DAO_PPL_SWITCH(m.nroi, (k_peak_sums<PPL><<<waves_grid(n), 256, 0, st>>>(m, px, pk, frames, roi_tab, slot0,
                                                                         slots, n, what, out)));
}
void launch_packed_slots(const Frame* frames, int nframes, const int* off, int* slots, hipStream_t st) {
  if (nframes > 0) k_packed_slots<<<nframes, 256, 0, st>>>(frames, off, slots);
}
void launch_status_flag(const Peaks& pk, const Frame* frames, const int* off, long long nslots, const int* status,
                        int* flag, hipStream_t st) {
  if (nslots > 0) k_status_flag<<<flat_grid(nslots), 256, 0, st>>>(pk, frames, off, nslots, status, flag);
}
void launch_apply_status(const Peaks& pk, const Frame* frames, const int* off, long long nslots, const int* status,
                         hipStream_t st) {
  if (nslots > 0) k_apply_status<<<flat_grid(nslots), 256, 0, st>>>(pk, frames, off, nslots, status);
}
void launch_compact_list(const int* flag, long long n, int* out, int* count, hipStream_t st) {
  k_compact<<<1, 1024, 0, st>>>(flag, n, out, count);
}
void launch_compact_frames(const Peaks& pk, const Frame* frames, int nframes, int drop, int* counts, int* counters,
                           hipStream_t st) {
  if (nframes > 0) k_compact_frames<<<nframes, 1024, 0, st>>>(pk, frames, drop, counts, counters);
}
void launch_peak_jacobian(const Model& m, const Pix& px, const Peaks& pk, const Frame* frames,
                          const int* roi_tab, int slot0, int n, double* out, hipStream_t st) {
  if (n == 0) return;
  DAO_PPL_SWITCH(m.nroi, DAO_SHAPE_SWITCH(m.shape, (k_peak_jacobian<PPL, SHP><<<waves_grid(n), 256, 0, st>>>(
                                                       m, px, pk, frames, roi_tab, slot0, n, out))));
}
void launch_gather_peaks(PeakRec* dst, const PeakRec* src, const int* map, int slot0, int n, hipStream_t st) {
  if (n == 0) return;
  k_gather_peaks<<<flat_grid(n), 256, 0, st>>>(dst, src, map, slot0, n);
}
void launch_set_status(const Peaks& pk, int slot0, int n, const int* status, hipStream_t st) {
  if (n == 0) return;
  k_set_status<<<flat_grid(n), 256, 0, st>>>(pk, slot0, n, status);
}
void launch_set_owner(const Peaks& pk, long long s0, long long s1, int frame, hipStream_t st) {
  if (s1 <= s0) return;
  k_set_owner<<<flat_grid(s1 - s0), 256, 0, st>>>(pk, s0, s1, frame);
}
void launch_select(const Peaks& pk, const Frame* frames, long long nslots, int mode, const int* frame_bad,
                   int* flag, int* out, int* count, hipStream_t st) {
  k_flag<<<flat_grid(nslots), 256, 0, st>>>(pk, frames, nslots, mode, frame_bad, flag);
  k_compact<<<1, 1024, 0, st>>>(flag, nslots, out, count);
}
void launch_fill_i32(int* p, long long n, int v, hipStream_t st) {
  if (n > 0) k_fill_i32<<<flat_grid(n), 256, 0, st>>>(p, n, v);
}
void launch_iota(int* p, int n, int base, hipStream_t st) {
  if (n > 0) k_iota<<<flat_grid(n), 256, 0, st>>>(p, n, base);
}
void launch_pack_peaks(const Peaks& pk, const Frame* frames, const int* offset, const Model& m, long long nslots,
                       double* out, int* status, hipStream_t st) {
  k_pack_peaks<<<flat_grid(nslots), 256, 0, st>>>(pk, frames, offset, m, nslots, out, status);
}

// Team tables before a k_fit_team launch: every frame at generation 0 with
// no escape, every seat idle.
__global__ void k_team_init(TeamCtl* ctl, int nframes, int* cur, int nseats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nframes) {
    TeamCtl c = {};
    c.esc = kNoEscape;
    ctl[i] = c;
  }
  if (i < nseats) cur[i] = kIdle;
}

// Frames flagged bad return to the copy of their pixel fit state and peak
// records taken at the sweep's start: block (f, c) copies chunk c of frame f.
__global__ void k_restore_bad_frames(const Frame* frames, const int* bad, PixRW* rw, const PixRW* rw_copy,
                                     PeakRec* rec, const PeakRec* rec_copy) {
  if (!bad[blockIdx.x]) return;
  const Frame fr = frames[blockIdx.x];
  constexpr int kQ = sizeof(PixRW) / sizeof(int4);
  const long long n = (long long)fr.W * fr.H * kQ;
  const long long per = (n + gridDim.y - 1) / gridDim.y;
  const long long lo = per * blockIdx.y, hi = lo + per < n ? lo + per : n;
  const int4* src = reinterpret_cast<const int4*>(rw_copy + fr.pix0);
  int4* dst = reinterpret_cast<int4*>(rw + fr.pix0);
  for (long long i = lo + threadIdx.x; i < hi; i += blockDim.x) dst[i] = src[i];
  if (blockIdx.y == 0) {
    constexpr int kR = sizeof(PeakRec) / sizeof(long long);
    const long long* rs = reinterpret_cast<const long long*>(rec_copy + fr.peak0);
    long long* rd = reinterpret_cast<long long*>(rec + fr.peak0);
    for (long long i = threadIdx.x; i < (long long)fr.nfit * kR; i += blockDim.x) rd[i] = rs[i];
  }
}

void launch_restore_bad_frames(const Frame* frames, int nframes, const int* bad, PixRW* rw, const PixRW* rw_copy,
                               PeakRec* rec, const PeakRec* rec_copy, hipStream_t st) {
  if (nframes > 0) k_restore_bad_frames<<<dim3(nframes, 16), 256, 0, st>>>(frames, bad, rw, rw_copy, rec, rec_copy);
}

void launch_team_init(TeamCtl* ctl, int nframes, int* cur, int nseats, hipStream_t st) {
  const int n = nframes > nseats ? nframes : nseats;
  if (n > 0) k_team_init<<<flat_grid(n), 256, 0, st>>>(ctl, nframes, cur, nseats);
}

void launch_add_peaks(const FitArgs& a, const AddPeaksJob* jobs, int njobs, int fmt, int nw, hipStream_t st) {
  if (njobs == 0) return;
  DAO_PPL_SWITCH(a.m.nroi, ({
                   constexpr int NW = PPL <= 8 ? 8 : 4;
                   launch_add_peaks_t<PPL, NW>(a, jobs, njobs, fmt, nw, st);
                 }));
}

// k_fit_team over nblocks workgroups, FP64 or mixed precision (MLE only);
// instantiations in dao_team_fp64.hip and dao_team_mixed.hip.
void launch_fit_team(const FitArgs& a, int nblocks, hipStream_t st) {
  if (nblocks == 0) return;
  if (a.m.prec == PREC_MIXED && a.m.errm == EM_MLE) {
    launch_fit_team_mixed(a, nblocks, st);
  } else {
    launch_fit_team_fp64(a, nblocks, st);
  }
}

// k_fit_frames, FP64 or mixed precision (MLE only); the instantiations are
// compiled in dao_fit_fp64.hip and dao_fit_mixed.hip.
void launch_fit_frames(const FitArgs& a, int nframes, hipStream_t st) {
  if (nframes == 0) return;
  if (a.m.prec == PREC_MIXED && a.m.errm == EM_MLE) {
    launch_fit_frames_mixed(a, nframes, st);
  } else {
    launch_fit_frames_fp64(a, nframes, st);
  }
}

}  // namespace dao
