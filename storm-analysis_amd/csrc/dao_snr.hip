// The peak finder's per-pixel arithmetic on the device
// (PeakFinderGaussian.peakFinder, storm_analysis/sa_library/fitting.py:444-501):
// the background variance and foreground images before the matched filters,
// and the signal-to-noise image after them.  One pass over HBM each, the
// reference's operations in its order, so results match numpy bit for bit.
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/dao_snr_amd.h"

namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_prepare(const double* image, const double* bg, const double* fit,
                                                      const double* cam, long long npix, long long total,
                                                      double* bg_var, double* fg) {
  for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < total;
       i += (long long)gridDim.x * kThreads) {
    const double b = bg[i], f = fit[i];
    double v = b + f;  // bg_var = self.background + fit_peaks_image
    if (cam) v = v + cam[i % npix];  // bg_var += self.camera_variance
    bg_var[i] = v;
    fg[i] = (image[i] - b) - f;  // foreground = self.image - self.background - fit_peaks_image
  }
}

__global__ __launch_bounds__(kThreads) void k_snr(const double* fg, const double* bg_var, const double* mask,
                                                  long long npix, long long total, double* snr) {
  for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < total;
       i += (long long)gridDim.x * kThreads) {
    double v = bg_var[i];
    if (v <= 0.1) v = 0.1;                 // bg_var[mask] = 0.1
    double s = fg[i] / sqrt(v);            // foreground / numpy.sqrt(bg_var)
    if (mask) s = s * mask[i % npix];      // foreground * self.peak_mask
    snr[i] = s;
  }
}

hipStream_t stream() {
  static hipStream_t st = nullptr;
  if (st == nullptr && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
  return st;
}

unsigned grid_for(long long n) {
  const long long g = (n + kThreads - 1) / kThreads;
  return (unsigned)(g < 8192 ? (g < 1 ? 1 : g) : 8192);
}

int finish(hipStream_t st) {
  if (hipGetLastError() != hipSuccess) return -2;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -2;
}

}  // namespace

extern "C" {

int daoFinderPrepare(const double* image, const double* background, const double* fit, const double* camera_variance,
                     int n, long long npix, double* bg_var, double* foreground) {
  if (image == nullptr || background == nullptr || fit == nullptr || bg_var == nullptr || foreground == nullptr ||
      n <= 0 || npix <= 0)
    return -1;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  const long long total = (long long)n * npix;
  k_prepare<<<grid_for(total), kThreads, 0, st>>>(image, background, fit, camera_variance, npix, total, bg_var,
                                                  foreground);
  return finish(st);
}

int daoFinderSNR(const double* foreground, const double* bg_var, const double* peak_mask, int n, long long npix,
                 double* snr) {
  if (foreground == nullptr || bg_var == nullptr || snr == nullptr || n <= 0 || npix <= 0) return -1;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  const long long total = (long long)n * npix;
  k_snr<<<grid_for(total), kThreads, 0, st>>>(foreground, bg_var, peak_mask, npix, total, snr);
  return finish(st);
}

}  // extern "C"
