// FFT image convolution on the device (libmatched_filter.so): the reference's
// matched filter (storm_analysis/sa_library/matched_filter.c:60-230) with the
// transforms in hipFFT (double precision, real-to-complex and back) and the
// spectrum product in a HIP kernel.  The filter's arrays stay in device
// memory; the host-array entry points copy the image in and the result out.
//
// convolve: spec = FFT(image); spec *= psf_fft (psf_fft = FFT(psf) / (X Y),
// stored at initialize); result = inverse FFT(spec), unnormalised -- FFTW's
// c2r does not scale, and neither does hipFFT's.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <stdint.h>
#include <stdlib.h>

#include "../../include/dao_mfilter_amd.h"

struct daoMatchedFilter {
  int x_size, y_size;
  size_t image_size, fft_size;  // x * y reals, x * (y / 2 + 1) complex
  double max_diff;
  hipStream_t st;
  hipfftHandle fwd, bwd;  // one image
  hipfftHandle fwd_n, bwd_n;  // batch_n images
  int batch_n;
  double* psf_fft;  // fft_size complex (interleaved)
  double* spec;     // fft_size complex
  double* in;       // image_size: the input image
  double* vec;      // image_size: the last result (the PSF before any convolution)
  double* old;      // image_size: the previous image (memoizing filters)
  int* flag;        // 1 int: some pixel differs
  double* spec_n;   // batch_n spectra
};

namespace {

constexpr int kThreads = 256;

// a *= b (complex), one spectrum of `period` values per blockIdx.y; the
// product as FFTW-style code writes it (real: ar br - ai bi, imaginary:
// ar bi + ai br).
__global__ __launch_bounds__(kThreads) void k_cmul(double2* a, const double2* b, size_t period) {
  double2* ai = a + period * blockIdx.y;
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < period; i += (size_t)gridDim.x * kThreads) {
    const double2 x = ai[i], y = b[i];
    ai[i] = make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
  }
}

// b = a * s (complex, interleaved), n values.
__global__ __launch_bounds__(kThreads) void k_cscale(const double* a, double* b, double s, size_t n) {
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < 2 * n; i += (size_t)gridDim.x * kThreads)
    b[i] = a[i] * s;
}

// *flag = 1 when some |a - b| > max_diff (flag cleared by the caller's kernel).
__global__ __launch_bounds__(kThreads) void k_differs(const double* a, const double* b, size_t n, double max_diff,
                                                      int* flag) {
  bool d = false;
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
    d = d || fabs(a[i] - b[i]) > max_diff;
  if (d) atomicOr(flag, 1);
}

__global__ void k_zero_int(int* p) { *p = 0; }

unsigned grid_for(size_t n) {
  const size_t g = (n + kThreads - 1) / kThreads;
  return (unsigned)(g < 4096 ? (g < 1 ? 1 : g) : 4096);
}

bool ok(hipError_t e) { return e == hipSuccess; }
bool ok(hipfftResult r) { return r == HIPFFT_SUCCESS; }

// spec = FFT(in); spec *= psf_fft; out = inverse FFT(spec).
// `in` is the transform's input and may be overwritten by it (rocFFT's
// out-of-place transforms are allowed to use their input as work space):
// callers pass a buffer of their own, never the caller's images.
bool run(filter* f, hipfftHandle fwd, hipfftHandle bwd, double* in, double* spec, double* out, int n) {
  if (!ok(hipfftExecD2Z(fwd, in, reinterpret_cast<hipfftDoubleComplex*>(spec)))) return false;
  k_cmul<<<dim3(grid_for(f->fft_size), (unsigned)n), kThreads, 0, f->st>>>(
      reinterpret_cast<double2*>(spec), reinterpret_cast<const double2*>(f->psf_fft), f->fft_size);
  if (!ok(hipGetLastError())) return false;
  return ok(hipfftExecZ2D(bwd, reinterpret_cast<hipfftDoubleComplex*>(spec), out));
}

bool plan(hipfftHandle* p, int x, int y, hipfftType t, int n, hipStream_t st) {
  int dims[2] = {x, y};
  if (!ok(hipfftPlanMany(p, 2, dims, nullptr, 1, 0, nullptr, 1, 0, t, n))) return false;
  return ok(hipfftSetStream(*p, st));
}

void release(filter* f) {
  for (hipfftHandle* h : {&f->fwd, &f->bwd, &f->fwd_n, &f->bwd_n})
    if (*h) (void)hipfftDestroy(*h);
  for (void* p : {(void*)f->psf_fft, (void*)f->spec, (void*)f->in, (void*)f->vec, (void*)f->old, (void*)f->flag,
                  (void*)f->spec_n})
    if (p) (void)hipFree(p);
  if (f->st) (void)hipStreamDestroy(f->st);
  free(f);
}

}  // namespace

extern "C" {

filter* initialize(double* psf, double max_diff, int x_size, int y_size, int estimate) {
  (void)estimate;
  if (psf == nullptr || x_size <= 0 || y_size <= 0 || (long long)x_size * y_size >= (1LL << 31)) return nullptr;
  filter* f = static_cast<filter*>(calloc(1, sizeof(filter)));
  if (f == nullptr) return nullptr;
  f->x_size = x_size;
  f->y_size = y_size;
  f->image_size = (size_t)x_size * y_size;
  f->fft_size = (size_t)x_size * (y_size / 2 + 1);
  f->max_diff = max_diff > 0.0 ? max_diff : 0.0;
  const double normalization = 1.0 / ((double)(x_size * y_size));
  bool good = ok(hipStreamCreateWithFlags(&f->st, hipStreamNonBlocking)) &&
              ok(hipMalloc(&f->psf_fft, sizeof(double) * 2 * f->fft_size)) &&
              ok(hipMalloc(&f->spec, sizeof(double) * 2 * f->fft_size)) &&
              ok(hipMalloc(&f->in, sizeof(double) * f->image_size)) &&
              ok(hipMalloc(&f->vec, sizeof(double) * f->image_size)) && ok(hipMalloc(&f->flag, sizeof(int))) &&
              plan(&f->fwd, x_size, y_size, HIPFFT_D2Z, 1, f->st) && plan(&f->bwd, x_size, y_size, HIPFFT_Z2D, 1, f->st);
  if (good && f->max_diff > 0.0)
    good = ok(hipMalloc(&f->old, sizeof(double) * f->image_size)) &&
           ok(hipMemsetAsync(f->old, 0, sizeof(double) * f->image_size, f->st));
  // FFT of the PSF, normalised; the work buffer keeps the PSF (matched_filter.c:222-227)
  if (good)
    good = ok(hipMemcpyAsync(f->vec, psf, sizeof(double) * f->image_size, hipMemcpyHostToDevice, f->st)) &&
           ok(hipMemcpyAsync(f->in, psf, sizeof(double) * f->image_size, hipMemcpyHostToDevice, f->st)) &&
           ok(hipfftExecD2Z(f->fwd, f->in, reinterpret_cast<hipfftDoubleComplex*>(f->spec)));
  if (good) {
    k_cscale<<<grid_for(2 * f->fft_size), kThreads, 0, f->st>>>(f->spec, f->psf_fft, normalization, f->fft_size);
    good = ok(hipGetLastError()) && ok(hipStreamSynchronize(f->st));
  }
  if (!good) {
    release(f);
    return nullptr;
  }
  return f;
}

void convolve(filter* f, double* image, double* result) {
  if (f == nullptr || image == nullptr || result == nullptr) return;
  const size_t b = sizeof(double) * f->image_size;
  if (ok(hipMemcpyAsync(f->in, image, b, hipMemcpyHostToDevice, f->st)) && run(f, f->fwd, f->bwd, f->in, f->spec, f->vec, 1))
    (void)hipMemcpyAsync(result, f->vec, b, hipMemcpyDeviceToHost, f->st);
  (void)hipStreamSynchronize(f->st);
}

void convolveMemo(filter* f, double* image, double* result) {
  if (f == nullptr || image == nullptr || result == nullptr) return;
  if (f->old == nullptr) {  // not configured for memoization (the reference has no previous image to read)
    convolve(f, image, result);
    return;
  }
  const size_t b = sizeof(double) * f->image_size;
  int differs = 0;
  bool good = ok(hipMemcpyAsync(f->in, image, b, hipMemcpyHostToDevice, f->st));
  if (good) {
    k_zero_int<<<1, 1, 0, f->st>>>(f->flag);
    k_differs<<<grid_for(f->image_size), kThreads, 0, f->st>>>(f->in, f->old, f->image_size, f->max_diff, f->flag);
    good = ok(hipGetLastError()) && ok(hipMemcpyAsync(&differs, f->flag, sizeof(int), hipMemcpyDeviceToHost, f->st)) &&
           ok(hipStreamSynchronize(f->st));
  }
  if (good && differs)
    good = ok(hipMemcpyAsync(f->old, f->in, b, hipMemcpyDeviceToDevice, f->st)) &&
           run(f, f->fwd, f->bwd, f->in, f->spec, f->vec, 1);
  if (good) (void)hipMemcpyAsync(result, f->vec, b, hipMemcpyDeviceToHost, f->st);
  (void)hipStreamSynchronize(f->st);
}

void cleanup(filter* f) {
  if (f == nullptr) return;
  (void)hipStreamSynchronize(f->st);
  release(f);
}

int daoMatchedFilterConvolveBatch(filter* f, const double* images, int n, double* result) {
  if (f == nullptr || images == nullptr || result == nullptr || n <= 0 || n > 65535) return -1;
  if (n != f->batch_n) {
    if (f->fwd_n) (void)hipfftDestroy(f->fwd_n);
    if (f->bwd_n) (void)hipfftDestroy(f->bwd_n);
    if (f->spec_n) (void)hipFree(f->spec_n);
    f->fwd_n = f->bwd_n = 0;
    f->spec_n = nullptr;
    f->batch_n = 0;
    if (!ok(hipMalloc(&f->spec_n, sizeof(double) * 2 * f->fft_size * (size_t)n)) ||
        !plan(&f->fwd_n, f->x_size, f->y_size, HIPFFT_D2Z, n, f->st) ||
        !plan(&f->bwd_n, f->x_size, f->y_size, HIPFFT_Z2D, n, f->st))
      return -2;
    f->batch_n = n;
  }
  // the forward transform may overwrite its input: it reads a copy of the
  // images in `result`, which the inverse transform then fills
  if (result != images &&
      !ok(hipMemcpyAsync(result, images, sizeof(double) * f->image_size * (size_t)n, hipMemcpyDeviceToDevice, f->st)))
    return -2;
  if (!run(f, f->fwd_n, f->bwd_n, result, f->spec_n, result, n)) return -2;
  return ok(hipStreamSynchronize(f->st)) ? 0 : -2;
}

}  // extern "C"
