// Frame ingest on the device: raw camera frames to the fitter's input.
//
// The reference prepares each frame on the host, one at a time:
//   - FrameReader.loadAFrame (storm_analysis/sa_library/analysis_io.py:243-260):
//     frame = (frame - offset) * (gain * rqe), values below 1.0 set to 1.0
//     (gain = 1 / camera gain, rqe = 1 / relative QE; scalars for a standard
//     camera, per-pixel maps for sCMOS, analysis_io.py:263-295);
//   - DaxReader.loadAFrame (storm_analysis/sa_library/datareader.py:234-243):
//     uint16 pixels, byte-swapped when the movie is big-endian;
//   - padArray (storm_analysis/sa_library/fitting.py:112-132): the frame grows
//     by `pad` pixels on every side, filled by mirroring the frame (the edge
//     pixel repeated: rows first, then columns, which for a separable mirror is
//     the same as mirroring each axis).
// Here one pass over HBM does all three for a batch of frames: each output
// pixel is computed from its mirrored source pixel, a workgroup per four
// padded rows.  The
// arithmetic is the reference's, operation for operation (one subtraction,
// the gain * rqe product, one multiplication, the clamp), so the results are
// bit-identical.  Bytes per output pixel: 2 (raw, shared by up to 4 padded
// pixels) + 8 written; the per-pixel calibration maps are read once per frame
// and stay in the caches across frames.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dao_ingest_amd.h"

namespace {

// Source index of padded coordinate o (0 <= o < n + 2 pad) of an axis of
// length n: the reference's flipud / fliplr mirror (edge pixel repeated).
__device__ inline int mirror(int o, int n, int pad) {
  const int i = o - pad;
  if (i < 0) return -1 - i;
  if (i >= n) return 2 * n - 1 - i;
  return i;
}

constexpr int kRowsPerBlock = 4;

// mode 0: raw uint16 counts -> photo-electrons; mode 1: float64 frames,
// padded only.  Workgroup (row block, frame) writes kRowsPerBlock padded
// output rows; its threads stride along each, so the stores are contiguous
// and no thread divides.
template <int MODE>
__global__ __launch_bounds__(256) void k_ingest(const void* in, int W, int H, int big_endian, const double* off_map,
                                                const double* gain_map, const double* rqe_map, double off,
                                                double gain, double rqe, int pad, double* out) {
  const int Wo = W + 2 * pad, Ho = H + 2 * pad;
  const long long f = blockIdx.y;
  for (int r = 0; r < kRowsPerBlock; ++r) {
    const int oy = blockIdx.x * kRowsPerBlock + r;
    if (oy >= Ho) break;
    const int sy = mirror(oy, H, pad);
    double* orow = out + (f * Ho + oy) * (long long)Wo;
    const long long srow = (long long)sy * W;
    for (int ox = threadIdx.x; ox < Wo; ox += blockDim.x) {
      const long long s = srow + mirror(ox, W, pad);
      double v;
      if (MODE == 0) {
        uint16_t a = static_cast<const uint16_t*>(in)[f * W * H + s];
        if (big_endian) a = (uint16_t)((a >> 8) | (a << 8));
        const double o = off_map ? off_map[s] : off;
        const double g = gain_map ? gain_map[s] : gain;
        const double q = rqe_map ? rqe_map[s] : rqe;
        v = ((double)a - o) * (g * q);
        if (v < 1.0) v = 1.0;
      } else {
        v = static_cast<const double*>(in)[f * W * H + s];
      }
      orow[ox] = v;
    }
  }
}

hipStream_t stream() {
  static hipStream_t st = nullptr;
  if (st == nullptr && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
  return st;
}

// Device copy of a host or device array of `bytes` (nullptr in -> nullptr).
const void* on_device(const void* p, size_t bytes, int is_device, void** tmp, hipStream_t st) {
  *tmp = nullptr;
  if (p == nullptr || is_device) return p;
  if (hipMalloc(tmp, bytes) != hipSuccess) return nullptr;
  if (hipMemcpyAsync(*tmp, p, bytes, hipMemcpyHostToDevice, st) != hipSuccess) return nullptr;
  return *tmp;
}

bool sizes_ok(int n, int W, int H, int pad) {
  // the grid holds one workgroup per padded row (x) and frame (y)
  return n > 0 && n <= 65535 && W > 0 && H > 0 && pad >= 0 && pad <= W && pad <= H &&
         (long long)H + 2LL * pad <= 0x7fffffffLL;
}

// One workgroup per kRowsPerBlock padded output rows of every frame.
dim3 grid_for(int H, int n, int pad) {
  return dim3((unsigned)((H + 2 * pad + kRowsPerBlock - 1) / kRowsPerBlock), (unsigned)n);
}

}  // namespace

extern "C" {

int daoIngestFrames(const uint16_t* raw, int raw_on_device, int n, int width, int height, int big_endian,
                    const double* offset, const double* gain, const double* rqe, int per_pixel,
                    int calib_on_device, int pad, double* out) {
  if (!sizes_ok(n, width, height, pad) || raw == nullptr || out == nullptr || offset == nullptr ||
      gain == nullptr || rqe == nullptr)
    return -1;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  const size_t npx = (size_t)width * height;
  void *t_raw = nullptr, *t_o = nullptr, *t_g = nullptr, *t_q = nullptr;
  const void* d_raw = on_device(raw, sizeof(uint16_t) * npx * n, raw_on_device, &t_raw, st);
  const double *d_o = nullptr, *d_g = nullptr, *d_q = nullptr;
  double so = 0.0, sg = 0.0, sq = 0.0;
  if (per_pixel) {
    d_o = static_cast<const double*>(on_device(offset, sizeof(double) * npx, calib_on_device, &t_o, st));
    d_g = static_cast<const double*>(on_device(gain, sizeof(double) * npx, calib_on_device, &t_g, st));
    d_q = static_cast<const double*>(on_device(rqe, sizeof(double) * npx, calib_on_device, &t_q, st));
  } else if (calib_on_device) {
    if (hipMemcpy(&so, offset, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&sg, gain, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&sq, rqe, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return -2;
  } else {
    so = *offset;
    sg = *gain;
    sq = *rqe;
  }
  int rc = 0;
  if (d_raw == nullptr || (per_pixel && (d_o == nullptr || d_g == nullptr || d_q == nullptr))) {
    rc = -2;
  } else {
    k_ingest<0><<<grid_for(height, n, pad), 256, 0, st>>>(d_raw, width, height, big_endian, d_o, d_g, d_q, so,
                                                            sg, sq, pad, out);
    if (hipGetLastError() != hipSuccess) rc = -2;
  }
  if (hipStreamSynchronize(st) != hipSuccess) rc = -2;
  for (void* p : {t_raw, t_o, t_g, t_q})
    if (p) (void)hipFree(p);
  return rc;
}

int daoPadFrames(const double* in, int in_on_device, int n, int width, int height, int pad, double* out) {
  if (!sizes_ok(n, width, height, pad) || in == nullptr || out == nullptr) return -1;
  hipStream_t st = stream();
  if (st == nullptr) return -2;
  void* t_in = nullptr;
  const void* d_in = on_device(in, sizeof(double) * width * height * (size_t)n, in_on_device, &t_in, st);
  int rc = 0;
  if (d_in == nullptr) {
    rc = -2;
  } else {
    k_ingest<1><<<grid_for(height, n, pad), 256, 0, st>>>(d_in, width, height, 0, nullptr, nullptr, nullptr, 0.0,
                                                            0.0, 0.0, pad, out);
    if (hipGetLastError() != hipSuccess) rc = -2;
  }
  if (hipStreamSynchronize(st) != hipSuccess) rc = -2;
  if (t_in) (void)hipFree(t_in);
  return rc;
}

}  // extern "C"
