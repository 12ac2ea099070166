// Local-maxima peak finder on the device: findLocalMaxima and calcMaxPeaks
// (storm_analysis/sa_library/ia_utilities.c:52-220).
//
// The reference walks the pixels in (plane, row, column) order and emits every
// pixel that is above threshold, not yet taken and a local maximum, stopping
// after n_peaks.  Whether a pixel is a maximum reads only the images, and the
// walk changes only the taken counter of the pixel it emits, which it never
// visits again; so every pixel can be decided at once.  Three launches:
//   k_mark  one workgroup per (row, plane): the reference's test per pixel,
//           its search window computed with the reference's integer
//           conversions, flags in a byte mask and a count per row;
//   k_scan_rows, k_scan  exclusive prefix sums of the row counts within each
//           plane (one workgroup per plane), then of the plane totals;
//   k_emit  one workgroup per (row, plane): each flagged pixel's rank in the
//           row (wave ballots) + the row's offset = its index in the
//           reference's list; the first max_peaks are written and their taken
//           counters incremented.
// Bound: HBM for the image read and the mask (8 + 1 B per pixel, the window's
// re-reads hitting the caches), plus the per-candidate window scan.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include <mutex>

#include "../../include/dao_finder_amd.h"

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr double kMaxRadius = 1.0e6;  // keeps the window arithmetic in int range
constexpr int kMaxZRange = 1 << 20;

struct FlmArgs {
  const double* img;  // zs x ys x xs
  int32_t* taken;     // zs x ys x xs
  int zs, ys, xs, margin, z_range;
  double radius, threshold;
  uint8_t* flag;  // zs x ys x xs
  int* rowcnt;    // zs x ys row counts (then offsets within the plane), zs + 1 plane offsets
  const double* zv;
  double *oz, *oy, *ox, *oh;
  int max_peaks;
};

// isLocalMaxima (ia_utilities.c:175-214) over the window the caller computed.
// (Testing the whole window without the early return, to issue its loads
// together, measured slower: most candidates are vetoed by a near neighbour.)
//
// The eight nearest pixels of the candidate's own plane are tested first,
// their loads issued together, when all eight lie in the window: most
// candidates are vetoed there, and the test's outcome is the same in any
// order (a pixel is a maximum iff no window pixel vetoes it).
__device__ bool is_local_max(const FlmArgs& a, double cur, int cz, int sz, int ez, int sy, int cy, int ey, int sx,
                             int cx, int ex, int rr) {
  if (rr >= 2 && sx <= cx - 1 && ex >= cx + 1 && sy <= cy - 1 && ey >= cy + 1) {
    const double* c = a.img + ((size_t)cz * a.ys + cy) * a.xs + cx;
    const double* u = c - a.xs;
    const double* d = c + a.xs;
    // rows above and the pixel to the left: veto on a strictly larger value;
    // to the right and below: on a larger or equal value
    // (ia_utilities.c:199-206)
    const double v0 = u[-1], v1 = u[0], v2 = u[1], v3 = c[-1], v4 = c[1], v5 = d[-1], v6 = d[0], v7 = d[1];
    if (v0 > cur || v1 > cur || v2 >= cur || v3 > cur || v4 >= cur || v5 >= cur || v6 >= cur || v7 >= cur)
      return false;
  }
  for (int zi = sz; zi <= ez; ++zi) {
    const double* plane = a.img + (size_t)zi * a.ys * a.xs;
    for (int yi = sy; yi <= ey; ++yi) {
      const int dy = (yi - cy) * (yi - cy);
      const double* row = plane + (size_t)yi * a.xs;
      for (int xi = sx; xi <= ex; ++xi) {
        const int dx = (xi - cx) * (xi - cx);
        if (dx + dy <= rr) {
          const double v = row[xi];
          if (yi <= cy && xi <= cx) {
            if (v > cur) return false;
          } else if (v >= cur) {
            return false;
          }
        }
      }
    }
  }
  return true;
}

__device__ inline int wave_sum(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// findLocalMaxima's test (ia_utilities.c:115-150) for every pixel of one row.
// count_only: calcMaxPeaks's test (above threshold and not taken) instead.
template <bool COUNT_ONLY>
__global__ __launch_bounds__(kThreads) void k_mark(FlmArgs a) {
  __shared__ int s_w[kWaves];
  const int yi = blockIdx.x, zi = blockIdx.y;
  const size_t rowbase = ((size_t)zi * a.ys + yi) * a.xs;
  const bool row_in = yi >= a.margin && yi < a.ys - a.margin;
  // z and y search ranges (ia_utilities.c:120-133): assigned from double to int
  int sz = zi - a.z_range, ez = zi + a.z_range;
  if (sz < 0) sz = 0;
  if (ez >= a.zs) ez = a.zs - 1;
  int sy = (int)((double)yi - a.radius), ey = (int)((double)yi + a.radius);
  if (sy < 0) sy = 0;
  if (ey >= a.ys) ey = a.ys - 1;
  const double cr = ceil(a.radius);
  const int rr = (int)(a.radius * a.radius);
  int n = 0;
  for (int xi = threadIdx.x; xi < a.xs; xi += kThreads) {
    bool f = false;
    if (row_in && xi >= a.margin && xi < a.xs - a.margin) {
      const double cur = a.img[rowbase + xi];
      if (cur > a.threshold && a.taken[rowbase + xi] < 1) {
        if (COUNT_ONLY) {
          f = true;
        } else {
          int sx = (int)((double)xi - cr), ex = (int)((double)xi + cr);
          if (sx < 0) sx = 0;
          if (ex >= a.xs) ex = a.xs - 1;
          f = is_local_max(a, cur, zi, sz, ez, sy, yi, ey, sx, xi, ex, rr);
        }
      }
    }
    if (!COUNT_ONLY) a.flag[rowbase + xi] = f ? 1 : 0;
    n += f ? 1 : 0;
  }
  n = wave_sum(n);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int j = 0; j < kWaves; ++j) t += s_w[j];
    a.rowcnt[zi * a.ys + yi] = t;
  }
}

// Exclusive prefix sum of n counts in place by one workgroup, each thread
// owning a run of entries; returns the total (to every thread).
__device__ int scan_block(int* cnt, int n, int* s_w) {
  const int per = (n + kThreads - 1) / kThreads;
  const int r0 = min((int)threadIdx.x * per, n), r1 = min(r0 + per, n);
  int run = 0;
  for (int r = r0; r < r1; ++r) run += cnt[r];
  // inclusive scan of the runs across the workgroup
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int v = run;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  if (lane == 63) s_w[w] = v;
  __syncthreads();
  int acc = v - run, total = 0;
  for (int j = 0; j < kWaves; ++j) {
    if (j < w) acc += s_w[j];
    total += s_w[j];
  }
  for (int r = r0; r < r1; ++r) {
    const int c = cnt[r];
    cnt[r] = acc;
    acc += c;
  }
  return total;
}

// Row counts -> offsets within each plane (one workgroup per plane); the
// plane's total goes to tot[plane].
__global__ __launch_bounds__(kThreads) void k_scan_rows(int* cnt, int ys, int* tot) {
  __shared__ int s_w[kWaves];
  const int t = scan_block(cnt + (size_t)blockIdx.x * ys, ys, s_w);
  if (threadIdx.x == 0) tot[blockIdx.x] = t;
}

// Plane totals -> plane offsets; tot[n] = the number of maxima.
__global__ __launch_bounds__(kThreads) void k_scan(int* tot, int n) {
  __shared__ int s_w[kWaves];
  const int t = scan_block(tot, n, s_w);
  if (threadIdx.x == 0) tot[n] = t;
}

// Writes the flagged pixels of one row at their index in the reference's list.
__global__ __launch_bounds__(kThreads) void k_emit(FlmArgs a) {
  __shared__ int s_w[kWaves];
  const int yi = blockIdx.x, zi = blockIdx.y;
  const size_t rowbase = ((size_t)zi * a.ys + yi) * a.xs;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int base = a.rowcnt[(size_t)a.zs * a.ys + zi] + a.rowcnt[zi * a.ys + yi];
  if (base >= a.max_peaks) return;
  for (int x0 = 0; x0 < a.xs; x0 += kThreads) {
    const int xi = x0 + threadIdx.x;
    const bool f = xi < a.xs && a.flag[rowbase + xi] != 0;
    const unsigned long long m = __ballot(f);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_w[w] = __popcll(m);
    __syncthreads();
    int off = base + before;
    int tile = 0;
    for (int j = 0; j < kWaves; ++j) {
      if (j < w) off += s_w[j];
      tile += s_w[j];
    }
    if (f && off < a.max_peaks) {
      a.oz[off] = a.zv[zi];
      a.oy[off] = (double)yi;
      a.ox[off] = (double)xi;
      a.oh[off] = a.img[rowbase + xi];
      a.taken[rowbase + xi] += 1;
    }
    base += tile;
    __syncthreads();
  }
}

// n_found = min(maxima, max_peaks) and the maxima count, on the device.
__global__ void k_counts(const int* tot, int zs, int max_peaks, int* out) {
  const int total = tot[zs];
  out[0] = total < max_peaks ? total : (max_peaks > 0 ? max_peaks : 0);
  out[1] = total;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point: a file-scope device array of
// fixed size (every GPU the library runs on holds one until the process
// ends).  Host code never uses the array's name as an address (on the host
// it is a stand-in symbol); each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current.
// k_mark writes every mask byte and row count of the call's frames before
// k_scan_rows / k_scan / k_emit read them, so nothing needs clearing.
constexpr long long kNoAllocMaxPixels = 1ll << 24;  // mask bytes: zsize * ysize * xsize
constexpr long long kNoAllocMaxRows = 1ll << 16;    // row counts: zsize * ysize
constexpr int kNoAllocMaxPlanes = 4096;             // plane offsets: zsize + 1
constexpr long long kNoAllocCountsOff = kNoAllocMaxPixels;
constexpr long long kNoAllocBytes = kNoAllocMaxPixels + 4 * (kNoAllocMaxRows + kNoAllocMaxPlanes + 64);
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch_noalloc[kNoAllocBytes];

std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kMaxDevices] = {};

// This device's address of scratch_noalloc (nullptr on failure).
void* noalloc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (g_noalloc_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch_noalloc)) != hipSuccess) return nullptr;
    g_noalloc_ptr[dev] = p;
  }
  return g_noalloc_ptr[dev];
}

hipStream_t stream() {
  static hipStream_t st = nullptr;
  if (st == nullptr && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
  return st;
}

bool args_ok(int zs, int ys, int xs, int z_range, double radius) {
  return zs > 0 && ys > 0 && xs > 0 && zs <= 65535 && (long long)zs * ys * xs < (1LL << 40) &&
         (long long)zs * ys < (1LL << 30) && radius == radius && fabs(radius) <= kMaxRadius && z_range <= kMaxZRange &&
         z_range >= -kMaxZRange;
}

// Device scratch (mask, row counts, z values) kept between calls.
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

// The search on device arrays; *total receives the number of maxima (before
// the max_peaks cut).
int find_device(const double* img, int32_t* taken, int zs, int ys, int xs, int margin, int z_range, double radius,
                double threshold, const double* z_values, int max_peaks, double* z, double* y, double* x, double* h,
                int* n_found, hipStream_t st) {
  const size_t npx = (size_t)zs * ys * xs, nrows = (size_t)zs * ys;
  char* s = static_cast<char*>(scratch(up256(npx) + up256(sizeof(int) * (nrows + zs + 1)) + sizeof(double) * zs));
  if (s == nullptr) return -2;
  FlmArgs a;
  a.img = img;
  a.taken = taken;
  a.zs = zs;
  a.ys = ys;
  a.xs = xs;
  a.margin = margin;
  a.z_range = z_range;
  a.radius = radius;
  a.threshold = threshold;
  a.flag = reinterpret_cast<uint8_t*>(s);
  a.rowcnt = reinterpret_cast<int*>(s + up256(npx));
  double* dzv = reinterpret_cast<double*>(s + up256(npx) + up256(sizeof(int) * (nrows + zs + 1)));
  a.zv = dzv;
  a.oz = z;
  a.oy = y;
  a.ox = x;
  a.oh = h;
  a.max_peaks = max_peaks;
  if (hipMemcpyAsync(dzv, z_values, sizeof(double) * zs, hipMemcpyHostToDevice, st) != hipSuccess) return -2;
  const dim3 grid((unsigned)ys, (unsigned)zs);
  k_mark<false><<<grid, kThreads, 0, st>>>(a);
  k_scan_rows<<<zs, kThreads, 0, st>>>(a.rowcnt, ys, a.rowcnt + nrows);
  k_scan<<<1, kThreads, 0, st>>>(a.rowcnt + nrows, zs);
  if (max_peaks > 0) k_emit<<<grid, kThreads, 0, st>>>(a);
  if (hipGetLastError() != hipSuccess) return -2;
  int total = 0;
  if (hipMemcpyAsync(&total, a.rowcnt + nrows + zs, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return -2;
  *n_found = total < max_peaks ? total : (max_peaks > 0 ? max_peaks : 0);
  return total;
}

}  // namespace

extern "C" {

int daoFindLocalMaxima(const double* images, int32_t* taken, int zsize, int ysize, int xsize, int margin,
                       int z_range, double radius, double threshold, const double* z_values, int max_peaks,
                       double* z, double* y, double* x, double* h, int* n_found) {
  if (!args_ok(zsize, ysize, xsize, z_range, radius) || images == nullptr || taken == nullptr ||
      z_values == nullptr || n_found == nullptr ||
      (max_peaks > 0 && (z == nullptr || y == nullptr || x == nullptr || h == nullptr)))
    return -1;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  const int rc = find_device(images, taken, zsize, ysize, xsize, margin, z_range, radius, threshold, z_values,
                             max_peaks, z, y, x, h, n_found, st);
  return rc < 0 ? rc : 0;
}

// daoFindLocalMaxima with no allocation: the scratch is this device's
// scratch_noalloc, z values and counts live in device memory, and the work
// is queued on the caller's stream without a host synchronisation.
int daoFindLocalMaximaNoAlloc(const double* images, int32_t* taken, int zsize, int ysize, int xsize, int margin,
                              int z_range, double radius, double threshold, const double* z_values, int max_peaks,
                              double* z, double* y, double* x, double* h, int* counts, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!args_ok(zsize, ysize, xsize, z_range, radius) || images == nullptr || taken == nullptr ||
      z_values == nullptr || counts == nullptr ||
      (max_peaks > 0 && (z == nullptr || y == nullptr || x == nullptr || h == nullptr)))
    return -1;
  const long long npx = (long long)zsize * ysize * xsize, nrows = (long long)zsize * ysize;
  if (npx > kNoAllocMaxPixels || nrows > kNoAllocMaxRows || zsize > kNoAllocMaxPlanes) {
    fprintf(stderr,
            "daoFindLocalMaximaNoAlloc: %d x %d x %d pixels exceed the scratch (at most %lld pixels, %lld rows, "
            "%d planes)\n",
            zsize, ysize, xsize, kNoAllocMaxPixels, kNoAllocMaxRows, kNoAllocMaxPlanes);
    return -1;
  }
  uint8_t* s = static_cast<uint8_t*>(noalloc_scratch());
  if (s == nullptr) return -2;
  FlmArgs a;
  a.img = images;
  a.taken = taken;
  a.zs = zsize;
  a.ys = ysize;
  a.xs = xsize;
  a.margin = margin;
  a.z_range = z_range;
  a.radius = radius;
  a.threshold = threshold;
  a.flag = s;
  a.rowcnt = reinterpret_cast<int*>(s + kNoAllocCountsOff);
  a.zv = z_values;
  a.oz = z;
  a.oy = y;
  a.ox = x;
  a.oh = h;
  a.max_peaks = max_peaks;
  const dim3 grid((unsigned)ysize, (unsigned)zsize);
  k_mark<false><<<grid, kThreads, 0, st>>>(a);
  k_scan_rows<<<zsize, kThreads, 0, st>>>(a.rowcnt, ysize, a.rowcnt + nrows);
  k_scan<<<1, kThreads, 0, st>>>(a.rowcnt + nrows, zsize);
  if (max_peaks > 0) k_emit<<<grid, kThreads, 0, st>>>(a);
  k_counts<<<1, 1, 0, st>>>(a.rowcnt + nrows, zsize, max_peaks, counts);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int calcMaxPeaks(flmData* f) {
  if (f == nullptr || f->images == nullptr || f->taken == nullptr ||
      !args_ok(f->zsize, f->ysize, f->xsize, f->z_range, f->radius))
    return -1;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  const size_t plane = (size_t)f->ysize * f->xsize, npx = plane * f->zsize, nrows = (size_t)f->zsize * f->ysize;
  char* buf = nullptr;
  if (hipMalloc(&buf, up256(sizeof(double) * npx) + up256(sizeof(int32_t) * npx) + sizeof(int) * (nrows + f->zsize + 1)) !=
      hipSuccess)
    return -2;
  double* dimg = reinterpret_cast<double*>(buf);
  int32_t* dtk = reinterpret_cast<int32_t*>(buf + up256(sizeof(double) * npx));
  int* cnt = reinterpret_cast<int*>(buf + up256(sizeof(double) * npx) + up256(sizeof(int32_t) * npx));
  int rc = 0;
  for (int zi = 0; zi < f->zsize && rc == 0; ++zi)
    if (hipMemcpyAsync(dimg + zi * plane, f->images[zi], sizeof(double) * plane, hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(dtk + zi * plane, f->taken[zi], sizeof(int32_t) * plane, hipMemcpyHostToDevice, st) !=
            hipSuccess)
      rc = -2;
  int total = 0;
  if (rc == 0) {
    FlmArgs a{};
    a.img = dimg;
    a.taken = dtk;
    a.zs = f->zsize;
    a.ys = f->ysize;
    a.xs = f->xsize;
    a.margin = f->margin;
    a.z_range = f->z_range;
    a.radius = f->radius;
    a.threshold = f->threshold;
    a.rowcnt = cnt;
    k_mark<true><<<dim3((unsigned)f->ysize, (unsigned)f->zsize), kThreads, 0, st>>>(a);
    k_scan_rows<<<f->zsize, kThreads, 0, st>>>(cnt, f->ysize, cnt + nrows);
    k_scan<<<1, kThreads, 0, st>>>(cnt + nrows, f->zsize);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&total, cnt + nrows + f->zsize, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
      rc = -2;
  }
  if (hipStreamSynchronize(st) != hipSuccess) rc = -2;
  (void)hipFree(buf);
  return rc == 0 ? total : rc;
}

void findLocalMaxima(flmData* f, double* z, double* y, double* x, double* h) {
  if (f == nullptr) return;
  const int max_peaks = f->n_peaks;
  f->n_peaks = 0;
  if (f->images == nullptr || f->taken == nullptr || f->z_values == nullptr ||
      !args_ok(f->zsize, f->ysize, f->xsize, f->z_range, f->radius) ||
      (max_peaks > 0 && (z == nullptr || y == nullptr || x == nullptr || h == nullptr)))
    return;
  hipStream_t st = stream();
  if (st == nullptr) return;
  const size_t plane = (size_t)f->ysize * f->xsize, npx = plane * f->zsize;
  const size_t np = max_peaks > 0 ? (size_t)max_peaks : 1;
  const size_t b_img = up256(sizeof(double) * npx), b_tk = up256(sizeof(int32_t) * npx);
  const size_t b_out = up256(sizeof(double) * np);
  char* buf = nullptr;
  if (hipMalloc(&buf, b_img + b_tk + 4 * b_out) != hipSuccess) return;
  double* dimg = reinterpret_cast<double*>(buf);
  int32_t* dtk = reinterpret_cast<int32_t*>(buf + b_img);
  double* dz = reinterpret_cast<double*>(buf + b_img + b_tk);
  double* dy = reinterpret_cast<double*>(buf + b_img + b_tk + b_out);
  double* dx = reinterpret_cast<double*>(buf + b_img + b_tk + 2 * b_out);
  double* dh = reinterpret_cast<double*>(buf + b_img + b_tk + 3 * b_out);
  bool ok = true;
  for (int zi = 0; zi < f->zsize && ok; ++zi)
    ok = hipMemcpyAsync(dimg + zi * plane, f->images[zi], sizeof(double) * plane, hipMemcpyHostToDevice, st) ==
             hipSuccess &&
         hipMemcpyAsync(dtk + zi * plane, f->taken[zi], sizeof(int32_t) * plane, hipMemcpyHostToDevice, st) ==
             hipSuccess;
  int n = 0, total = -2;
  if (ok)
    total = find_device(dimg, dtk, f->zsize, f->ysize, f->xsize, f->margin, f->z_range, f->radius, f->threshold,
                        f->z_values, max_peaks, dz, dy, dx, dh, &n, st);
  if (total >= 0) {
    for (int zi = 0; zi < f->zsize && ok; ++zi)
      ok = hipMemcpyAsync(f->taken[zi], dtk + zi * plane, sizeof(int32_t) * plane, hipMemcpyDeviceToHost, st) ==
           hipSuccess;
    if (n > 0)
      ok = ok && hipMemcpyAsync(z, dz, sizeof(double) * n, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipMemcpyAsync(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost, st) == hipSuccess &&
           hipMemcpyAsync(h, dh, sizeof(double) * n, hipMemcpyDeviceToHost, st) == hipSuccess;
    ok = hipStreamSynchronize(st) == hipSuccess && ok;
    if (ok) {
      f->n_peaks = n;
      if (max_peaks > 0 && total >= max_peaks) printf("Warning! Found maximum number of peaks!\n");
    }
  }
  (void)hipStreamSynchronize(st);
  (void)hipFree(buf);
}

}  // extern "C"
