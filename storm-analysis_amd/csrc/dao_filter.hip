// Post-fit peak filters on the device: markDimmerPeaks,
// markLowSignificancePeaks and runningIfHasNeighbors
// (storm_analysis/sa_library/ia_utilities.c:221-430) for a batch of frames.
//
// The reference visits the peaks in index order and asks a kd-tree for each
// one's neighbours.  Its result does not depend on that order: a peak is only
// set to ERROR by its own visit, its neighbours only move CONVERGED -> RUNNING,
// and the tests that decide removal read heights, significances and the
// initial ERROR flags, none of which the loop changes.  So each filter is two
// data-parallel passes over a frame's peaks, with one workgroup per frame:
//   A. flag every peak the reference removes (a brighter peak within
//      r_removal, or significance <= the minimum; ERROR peaks are skipped);
//   B. removed peaks -> ERROR, and a CONVERGED peak with a flagged peak within
//      r_neighbors -> RUNNING.
// runningIfHasNeighbors is pass B alone, with the new peaks as the searched set.
//
// Neighbour search: a uniform grid over the frame's searched peaks, built in
// LDS per call (bounding box, counting sort of the peaks by cell, their
// positions and values staged in LDS for frames of up to 2048 peaks), with
// cells at least 1.001 x the largest radius wide, so every peak within the
// radius of a query lies in the 3 x 3 cells around it.  The distance test is
// the kd-tree's, operation for operation: (x_k - x_i)^2 + (y_k - y_i)^2 <= r^2
// (storm_analysis/dbscan/kdtree.c:269-273), with no fused multiply-add.  Frames
// whose positions are not all finite, or a radius that is not a positive
// finite number, are searched by brute force with the same test.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/dao_filter_amd.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kGridMax = 64;  // cells per axis
constexpr int kCellsMax = kGridMax * kGridMax;
constexpr int kLdsPeaks = 2048;  // frames up to this many searched peaks are staged in LDS
constexpr int32_t kRunning = 0, kConverged = 1, kError = 2;

enum Mode { kDimmer = 0, kLowSig = 1, kRunningIf = 2 };

struct FilterArgs {
  const double* tx;  // searched peaks (the reference's kd-tree points)
  const double* ty;
  const int* toff;
  const double* qx;  // query peaks (the searched ones, except for runningIfHasNeighbors)
  const double* qy;
  const int* qoff;
  const double* val;  // heights (dimmer) or significances (low significance)
  int32_t* status;    // of the query peaks
  uint8_t* flag;      // scratch: removed flag per query peak
  int* sorted;        // scratch: searched peaks' indices, by cell
  int* removed;       // per frame (dimmer, low significance)
  double ra;          // r_removal (dimmer), min_sig (low significance), radius (runningIf)
  double rb;          // r_neighbors (dimmer, low significance)
};

// The kd-tree's squared distance of point k from query i, unfused
// (contraction off: the __d*_rn intrinsics are plain operators without
// OCML's rounded operations, so they would not stop an FMA).
__device__ inline double dist_sq(double xk, double yk, double xi, double yi) {
#pragma clang fp contract(off)
  const double dx = xk - xi;
  const double dy = yk - yi;
  return dx * dx + dy * dy;
}

__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

__device__ inline int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Inclusive scan of one int per thread across the workgroup.
__device__ inline int block_scan(int v, int* s_w) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  if (lane == 63) s_w[w] = v;
  __syncthreads();
  int add = 0;
  for (int j = 0; j < w; ++j) add += s_w[j];
  __syncthreads();
  return v + add;
}

// The grid over one frame's searched peaks.
struct Grid {
  double x0, y0, cw, ch;
  int gx, gy;
  bool brute;
};

__device__ inline int cell_of(const Grid& g, double x, double y) {
  const int cx = min(max((int)floor((x - g.x0) / g.cw), 0), g.gx - 1);
  const int cy = min(max((int)floor((y - g.y0) / g.ch), 0), g.gy - 1);
  return cy * g.gx + cx;
}

// A frame's searched peaks t0 .. t0+n-1 in search order s = 0 .. n-1: cell by
// cell after the counting sort (peak order for the brute-force search).  A
// frame of at most kLdsPeaks peaks on the grid is staged in LDS (position,
// value, peak index, removed flag per s), so the neighbour scans read LDS
// only; a larger one is read through a.sorted from global memory.
struct Frame {
  const FilterArgs* a;
  double* sx;
  double* sy;
  double* sv;
  int* sk;
  uint8_t* sf;
  int t0, n;
  bool lds, brute;
  __device__ int peak(int s) const { return lds ? sk[s] : (brute ? t0 + s : a->sorted[t0 + s]); }
  __device__ double v(int s) const { return lds ? sv[s] : a->val[peak(s)]; }
  __device__ bool flag(int s) const { return lds ? sf[s] != 0 : a->flag[peak(s)] != 0; }
  __device__ void set_flag(int s, bool b) const {
    if (lds)
      sf[s] = b ? 1 : 0;
    else
      a->flag[peak(s)] = b ? 1 : 0;
  }
  __device__ void pos(int s, double& x, double& y) const {
    if (lds) {
      x = sx[s];
      y = sy[s];
    } else {
      const int k = peak(s);
      x = a->tx[k];
      y = a->ty[k];
    }
  }
};

// Bounding box, cell sizes and the counting sort of the searched peaks into
// search order, cell c holding s = cell[c] .. cell[c+1]-1.
__device__ Grid build_grid(const FilterArgs& a, Frame& fr, double rmax, bool with_val, int* cell, double* s_d,
                           int* s_i) {
  const int w = threadIdx.x >> 6;
  const int t0 = fr.t0, t1 = fr.t0 + fr.n;
  double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
  int bad = 0;
  for (int k = t0 + threadIdx.x; k < t1; k += kThreads) {
    const double x = a.tx[k], y = a.ty[k];
    if (!isfinite(x) || !isfinite(y)) bad = 1;
    xmin = fmin(xmin, x);
    xmax = fmax(xmax, x);
    ymin = fmin(ymin, y);
    ymax = fmax(ymax, y);
  }
  xmin = wave_min(xmin);
  xmax = wave_max(xmax);
  ymin = wave_min(ymin);
  ymax = wave_max(ymax);
  bad = wave_sum(bad);
  if ((threadIdx.x & 63) == 0) {
    s_d[4 * w + 0] = xmin;
    s_d[4 * w + 1] = xmax;
    s_d[4 * w + 2] = ymin;
    s_d[4 * w + 3] = ymax;
    s_i[w] = bad;
  }
  __syncthreads();
  Grid g;
  xmin = s_d[0], xmax = s_d[1], ymin = s_d[2], ymax = s_d[3];
  bad = s_i[0];
  for (int j = 1; j < kWaves; ++j) {
    xmin = fmin(xmin, s_d[4 * j + 0]);
    xmax = fmax(xmax, s_d[4 * j + 1]);
    ymin = fmin(ymin, s_d[4 * j + 2]);
    ymax = fmax(ymax, s_d[4 * j + 3]);
    bad += s_i[j];
  }
  __syncthreads();
  const double wx = xmax - xmin, wy = ymax - ymin;
  g.brute = bad != 0 || !(rmax > 0.0) || !isfinite(rmax) || !isfinite(wx) || !isfinite(wy);
  fr.brute = g.brute;
  fr.lds = !g.brute && fr.n <= kLdsPeaks;
  if (g.brute) return g;
  g.x0 = xmin;
  g.y0 = ymin;
  g.cw = fmax(1.001 * rmax, wx / (kGridMax - 1));
  g.ch = fmax(1.001 * rmax, wy / (kGridMax - 1));
  g.gx = min((int)floor(wx / g.cw) + 1, kGridMax);
  g.gy = min((int)floor(wy / g.ch) + 1, kGridMax);
  const int nc = g.gx * g.gy;
  for (int c = threadIdx.x; c <= nc; c += kThreads) cell[c] = 0;
  __syncthreads();
  for (int k = t0 + threadIdx.x; k < t1; k += kThreads) atomicAdd(&cell[cell_of(g, a.tx[k], a.ty[k])], 1);
  __syncthreads();
  // inclusive prefix sum of the counts, each thread owning a run of cells
  const int per = (nc + kThreads - 1) / kThreads;
  const int c0 = min(threadIdx.x * per, nc), c1 = min(c0 + per, nc);
  int run = 0;
  for (int c = c0; c < c1; ++c) run += cell[c];
  int acc = block_scan(run, s_i) - run;
  for (int c = c0; c < c1; ++c) {
    acc += cell[c];
    cell[c] = acc;
  }
  if (threadIdx.x == 0) cell[nc] = fr.n;
  __syncthreads();
  // scatter: counting down leaves cell[c] at the start of cell c
  for (int k = t0 + threadIdx.x; k < t1; k += kThreads) {
    const double x = a.tx[k], y = a.ty[k];
    const int s = atomicSub(&cell[cell_of(g, x, y)], 1) - 1;
    if (fr.lds) {
      fr.sk[s] = k;
      fr.sx[s] = x;
      fr.sy[s] = y;
      if (with_val) fr.sv[s] = a.val[k];
    } else {
      a.sorted[t0 + s] = k;
    }
  }
  __syncthreads();
  return g;
}

// True when some searched peak s with pred(s) lies within sqrt(r2) of (x, y).
template <class Pred>
__device__ inline bool any_near(const Frame& fr, const Grid& g, const int* cell, double x, double y, double r2,
                                Pred pred) {
  if (g.brute) {
    for (int s = 0; s < fr.n; ++s) {
      double px, py;
      fr.pos(s, px, py);
      if (dist_sq(px, py, x, y) <= r2 && pred(s)) return true;
    }
    return false;
  }
  if (!isfinite(x) || !isfinite(y)) return false;  // no distance from it is <= r2
  const double fx = floor((x - g.x0) / g.cw), fy = floor((y - g.y0) / g.ch);
  // clamped before the conversion: a query far outside the grid maps to an empty range
  const int xlo = (int)fmin(fmax(fx - 1.0, 0.0), (double)g.gx), xhi = (int)fmax(fmin(fx + 1.0, g.gx - 1.0), -1.0);
  const int ylo = (int)fmin(fmax(fy - 1.0, 0.0), (double)g.gy), yhi = (int)fmax(fmin(fy + 1.0, g.gy - 1.0), -1.0);
  if (xlo > xhi || ylo > yhi) return false;
  for (int cy = ylo; cy <= yhi; ++cy) {
    const int s1 = cell[cy * g.gx + xhi + 1];
    for (int s = cell[cy * g.gx + xlo]; s < s1; ++s) {
      double px, py;
      fr.pos(s, px, py);
      if (dist_sq(px, py, x, y) <= r2 && pred(s)) return true;
    }
  }
  return false;
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_filter(FilterArgs a) {
  __shared__ int cell[kCellsMax + 1];
  __shared__ double s_x[kLdsPeaks], s_y[kLdsPeaks], s_v[MODE == kRunningIf ? 1 : kLdsPeaks];
  __shared__ int s_k[kLdsPeaks];
  __shared__ uint8_t s_f[MODE == kRunningIf ? 1 : kLdsPeaks];
  __shared__ double s_d[4 * kWaves];
  __shared__ int s_i[kWaves];
  const int f = blockIdx.x;
  const int t0 = a.toff[f], t1 = a.toff[f + 1];
  const int q0 = a.qoff[f], q1 = a.qoff[f + 1];
  if (t1 <= t0) {  // no peaks to search: nothing changes
    if (MODE != kRunningIf && threadIdx.x == 0) a.removed[f] = 0;
    return;
  }
  Frame fr{&a, s_x, s_y, s_v, s_k, s_f, t0, t1 - t0, false, false};
  if (MODE == kRunningIf) {
    const Grid g = build_grid(a, fr, fabs(a.ra), false, cell, s_d, s_i);
    const double r2 = __dmul_rn(a.ra, a.ra);
    for (int i = q0 + threadIdx.x; i < q1; i += kThreads) {
      const int32_t st = a.status[i];
      if (st == kRunning || st == kError) continue;
      if (any_near(fr, g, cell, a.qx[i], a.qy[i], r2, [](int) { return true; })) a.status[i] = kRunning;
    }
    return;
  }
  // pass A: flag the peaks the reference removes
  const Grid g = build_grid(a, fr, MODE == kDimmer ? fmax(fabs(a.ra), fabs(a.rb)) : fabs(a.rb), MODE == kDimmer,
                            cell, s_d, s_i);
  if (MODE == kDimmer) {
    const double r2 = __dmul_rn(a.ra, a.ra);
    for (int s = threadIdx.x; s < fr.n; s += kThreads) {
      bool rm = false;
      if (a.status[fr.peak(s)] != kError) {
        const double hs = fr.v(s);
        double x, y;
        fr.pos(s, x, y);
        rm = any_near(fr, g, cell, x, y, r2, [&](int j) { return fr.v(j) > hs; });
      }
      fr.set_flag(s, rm);
    }
  } else {
    for (int s = threadIdx.x; s < fr.n; s += kThreads) {
      const int k = fr.peak(s);
      fr.set_flag(s, a.status[k] != kError && !(a.val[k] > a.ra));
    }
  }
  __syncthreads();
  // pass B: removed -> ERROR; CONVERGED near a removed peak -> RUNNING
  const double r2 = __dmul_rn(a.rb, a.rb);
  int n_removed = 0;
  for (int s = threadIdx.x; s < fr.n; s += kThreads) {
    const int k = fr.peak(s);
    if (fr.flag(s)) {
      a.status[k] = kError;
      ++n_removed;
    } else if (a.status[k] == kConverged) {
      double x, y;
      fr.pos(s, x, y);
      if (any_near(fr, g, cell, x, y, r2, [&](int j) { return fr.flag(j); })) a.status[k] = kRunning;
    }
  }
  n_removed = wave_sum(n_removed);
  if ((threadIdx.x & 63) == 0) s_i[threadIdx.x >> 6] = n_removed;
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int j = 0; j < kWaves; ++j) total += s_i[j];
    a.removed[f] = total;
  }
}

hipStream_t stream() {
  static hipStream_t st = nullptr;
  if (st == nullptr && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
  return st;
}

// Device scratch kept between calls, grown on demand.
void* scratch(size_t bytes) {
  static void* p = nullptr;
  static size_t cap = 0;
  if (bytes > cap) {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = bytes < (1u << 20) ? (1u << 20) : bytes + bytes / 4;
    if (hipMalloc(&p, want) != hipSuccess) return nullptr;
    cap = want;
  }
  return p;
}

inline size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// Total peak count of a batch (offsets[nframes] from the device).
int total_of(const int* d_off, int nframes, hipStream_t st) {
  int n = -1;
  if (hipMemcpyAsync(&n, d_off + nframes, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  if (hipStreamSynchronize(st) != hipSuccess) return -1;
  return n;
}

template <int MODE>
int launch(FilterArgs a, int nframes, int nq, int nt, hipStream_t st) {
  // scratch: flags (nq bytes) then sorted indices (nt ints)
  char* s = static_cast<char*>(scratch(up256((size_t)nq) + sizeof(int) * (size_t)nt + 256));
  if (s == nullptr) return -2;
  a.flag = reinterpret_cast<uint8_t*>(s);
  a.sorted = reinterpret_cast<int*>(s + up256((size_t)nq));
  k_filter<MODE><<<nframes, kThreads, 0, st>>>(a);
  if (hipGetLastError() != hipSuccess) return -2;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -2;
}

template <int MODE>
int batch_mark(const double* x, const double* y, const double* v, int32_t* status, const int* off, int nframes,
               double ra, double rb, int* removed) {
  if (x == nullptr || y == nullptr || v == nullptr || status == nullptr || off == nullptr || removed == nullptr ||
      nframes <= 0)
    return -1;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  const int n = total_of(off, nframes, st);
  if (n < 0) return -2;
  FilterArgs a{x, y, off, x, y, off, v, status, nullptr, nullptr, removed, ra, rb};
  return launch<MODE>(a, nframes, n, n, st);
}

// One host peak list through the batched kernel: copies in, filter, status back.
template <int MODE>
int host_mark(const double* x, const double* y, const double* v, int32_t* status, double ra, double rb, int np) {
  if (np < 0 || (np > 0 && (x == nullptr || y == nullptr || v == nullptr || status == nullptr))) return -1;
  if (np == 0) return 0;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  const size_t nd = sizeof(double) * (size_t)np;
  char* buf = nullptr;
  if (hipMalloc(&buf, 3 * up256(nd) + up256(sizeof(int32_t) * np) + 256) != hipSuccess) return -2;
  double* dx = reinterpret_cast<double*>(buf);
  double* dy = reinterpret_cast<double*>(buf + up256(nd));
  double* dv = reinterpret_cast<double*>(buf + 2 * up256(nd));
  int32_t* ds = reinterpret_cast<int32_t*>(buf + 3 * up256(nd));
  int* doff = reinterpret_cast<int*>(buf + 3 * up256(nd) + up256(sizeof(int32_t) * np));  // offsets, then removed
  const int hoff[2] = {0, np};
  int removed = 0;
  int rc = 0;
  if (hipMemcpyAsync(dx, x, nd, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dy, y, nd, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dv, v, nd, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(ds, status, sizeof(int32_t) * np, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(doff, hoff, sizeof(hoff), hipMemcpyHostToDevice, st) != hipSuccess) {
    rc = -2;
  } else {
    FilterArgs a{dx, dy, doff, dx, dy, doff, dv, ds, nullptr, nullptr, doff + 2, ra, rb};
    rc = launch<MODE>(a, 1, np, np, st);
  }
  if (rc == 0 && (hipMemcpyAsync(status, ds, sizeof(int32_t) * np, hipMemcpyDeviceToHost, st) != hipSuccess ||
                  hipMemcpyAsync(&removed, doff + 2, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
                  hipStreamSynchronize(st) != hipSuccess))
    rc = -2;
  (void)hipFree(buf);
  return rc == 0 ? removed : rc;
}

}  // namespace

extern "C" {

int markDimmerPeaks(double* x, double* y, double* h, int32_t* status, double r_removal, double r_neighbors, int np) {
  return host_mark<kDimmer>(x, y, h, status, r_removal, r_neighbors, np);
}

int markLowSignificancePeaks(double* x, double* y, double* sig, int32_t* status, double min_sig, double r_neighbors,
                             int np) {
  return host_mark<kLowSig>(x, y, sig, status, min_sig, r_neighbors, np);
}

void runningIfHasNeighbors(double* c_x, double* c_y, double* n_x, double* n_y, int32_t* status, double radius, int nc,
                           int nn) {
  if (nc <= 0 || nn <= 0 || c_x == nullptr || c_y == nullptr || n_x == nullptr || n_y == nullptr ||
      status == nullptr)
    return;
  hipStream_t st = stream();
  if (st == nullptr) return;
  const size_t bc = up256(sizeof(double) * (size_t)nc), bn = up256(sizeof(double) * (size_t)nn);
  char* buf = nullptr;
  if (hipMalloc(&buf, 2 * bc + 2 * bn + up256(sizeof(int32_t) * nc) + 256) != hipSuccess) return;
  double* dcx = reinterpret_cast<double*>(buf);
  double* dcy = reinterpret_cast<double*>(buf + bc);
  double* dnx = reinterpret_cast<double*>(buf + 2 * bc);
  double* dny = reinterpret_cast<double*>(buf + 2 * bc + bn);
  int32_t* ds = reinterpret_cast<int32_t*>(buf + 2 * bc + 2 * bn);
  int* doff = reinterpret_cast<int*>(buf + 2 * bc + 2 * bn + up256(sizeof(int32_t) * nc));  // c then n offsets
  const int hoff[4] = {0, nc, 0, nn};
  if (hipMemcpyAsync(dcx, c_x, sizeof(double) * nc, hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(dcy, c_y, sizeof(double) * nc, hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(dnx, n_x, sizeof(double) * nn, hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(dny, n_y, sizeof(double) * nn, hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(ds, status, sizeof(int32_t) * nc, hipMemcpyHostToDevice, st) == hipSuccess &&
      hipMemcpyAsync(doff, hoff, sizeof(hoff), hipMemcpyHostToDevice, st) == hipSuccess) {
    FilterArgs a{dnx, dny, doff + 2, dcx, dcy, doff, nullptr, ds, nullptr, nullptr, nullptr, radius, 0.0};
    if (launch<kRunningIf>(a, 1, nc, nn, st) == 0)
      (void)hipMemcpyAsync(status, ds, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, st);
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(buf);
}

int daoBatchMarkDimmerPeaks(const double* x, const double* y, const double* h, int32_t* status, const int* offsets,
                            int nframes, double r_removal, double r_neighbors, int* removed) {
  return batch_mark<kDimmer>(x, y, h, status, offsets, nframes, r_removal, r_neighbors, removed);
}

int daoBatchMarkLowSignificancePeaks(const double* x, const double* y, const double* sig, int32_t* status,
                                     const int* offsets, int nframes, double min_sig, double r_neighbors,
                                     int* removed) {
  return batch_mark<kLowSig>(x, y, sig, status, offsets, nframes, min_sig, r_neighbors, removed);
}

int daoBatchRunningIfHasNeighbors(const double* c_x, const double* c_y, const int* c_offsets, const double* n_x,
                                  const double* n_y, const int* n_offsets, int32_t* status, int nframes,
                                  double radius) {
  if (c_x == nullptr || c_y == nullptr || c_offsets == nullptr || n_x == nullptr || n_y == nullptr ||
      n_offsets == nullptr || status == nullptr || nframes <= 0)
    return -1;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  const int nc = total_of(c_offsets, nframes, st), nn = total_of(n_offsets, nframes, st);
  if (nc < 0 || nn < 0) return -2;
  FilterArgs a{n_x, n_y, n_offsets, c_x, c_y, c_offsets, nullptr, status, nullptr, nullptr, nullptr, radius, 0.0};
  return launch<kRunningIf>(a, nframes, nc, nn, st);
}

}  // extern "C"
