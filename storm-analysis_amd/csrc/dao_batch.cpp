// Batch entry points of libdao_fit.so (declared in include/dao_batch_amd.h):
// many frames resident on the GPU, fitted by one executor.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "dao_engine.h"
#include "../../include/dao_batch_amd.h"

using dao::Engine;

struct daoBatch {
  std::unique_ptr<Engine> eng;
  std::vector<int> counters;  // [F][8], host totals not yet handed out
  double* d_tmp = nullptr;
  size_t tmp_bytes = 0;
};

namespace {

void pull_counters(daoBatch* b) {
  std::vector<int> c;
  b->eng->read_counters(c);
  for (size_t i = 0; i < c.size(); i++) b->counters[i] += c[i];
}

double* tmp(daoBatch* b, size_t bytes) {
  if (bytes > b->tmp_bytes) {
    if (b->d_tmp) DAO_CHECK(hipFree(b->d_tmp));
    DAO_CHECK(hipMalloc((void**)&b->d_tmp, bytes));
    b->tmp_bytes = bytes;
  }
  return b->d_tmp;
}

}  // namespace

extern "C" {

daoBatch* daoBatchCreate(const daoBatchModel* bm, int n_frames, const int* widths, const int* heights,
                         const int* peak_caps) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "dao_fit (HIP): no GPU device is available\n");
    return nullptr;
  }
  if (bm->roi_size < 1 || bm->roi_size > 32 || n_frames < 1) return nullptr;
  dao::Model m;
  std::memset(&m, 0, sizeof(m));
  m.shape = bm->shape;
  m.errm = bm->error_model;
  m.roi = bm->roi_size;
  m.jac = bm->shape == dao::SH_2DFIXED ? 4 : (bm->shape == dao::SH_3D ? 6 : 5);
  m.margin = bm->margin > 0 ? bm->margin : 2;
  m.prec = bm->precision == dao::PREC_MIXED ? dao::PREC_MIXED : dao::PREC_FP64;
  m.xoff = 0.5 * (double)bm->roi_size - 1.0;
  m.zoff = 0.0;
  m.tol = bm->tolerance;
  m.min_height = -1.0;
  m.wmin = bm->width_min;
  m.wmax = bm->width_max;
  for (int i = 0; i < 5; i++) {
    m.wxz[i] = bm->wx[i];
    m.wyz[i] = bm->wy[i];
  }
  m.wxz[0] = m.wxz[0] * m.wxz[0];
  m.wyz[0] = m.wyz[0] * m.wyz[0];
  m.zmin = bm->z_min;
  m.zmax = bm->z_max;
  std::vector<dao::FrameSpec> spec(n_frames);
  for (int f = 0; f < n_frames; f++) {
    if (widths[f] < 1 || heights[f] < 1 || peak_caps[f] < 0) return nullptr;
    spec[f].W = widths[f];
    spec[f].H = heights[f];
    spec[f].cap = peak_caps[f];
  }
  daoBatch* b = new daoBatch();
  b->eng.reset(new Engine(m, spec));
  b->counters.assign((size_t)n_frames * dao::C_COUNT, 0);
  // default calibration: rqe = 1, scmos = 0
  std::vector<double> ones;
  for (int f = 0; f < n_frames; f++) daoBatchSetCalibration(b, f, nullptr, nullptr);
  return b;
}

void daoBatchDestroy(daoBatch* b) {
  if (!b) return;
  b->eng.reset();
  if (b->d_tmp) (void)hipFree(b->d_tmp);
  delete b;
}

void daoBatchSetCalibration(daoBatch* b, int f, const double* rqe, const double* scmos) {
  Engine& e = *b->eng;
  size_t n = (size_t)e.fr[f].W * e.fr[f].H;
  std::vector<double> r, s;
  if (!rqe) {
    r.assign(n, 1.0);
    rqe = r.data();
  }
  if (!scmos) {
    s.assign(n, 0.0);
    scmos = s.data();
  }
  e.set_calibration(f, rqe, scmos);
}

void daoBatchNewImages(daoBatch* b, const double* images, int on_device) {
  Engine& e = *b->eng;
  const double* d = images;
  if (!on_device) {
    double* t = tmp(b, sizeof(double) * e.npix);
    DAO_CHECK(hipMemcpyAsync(t, images, sizeof(double) * e.npix, hipMemcpyHostToDevice, e.st));
    d = t;
  }
  e.new_images_all_dev(d);
  e.sync();
}

void daoBatchNewBackgrounds(daoBatch* b, const double* bgs, int on_device) {
  Engine& e = *b->eng;
  const double* d = bgs;
  if (!on_device) {
    double* t = tmp(b, sizeof(double) * e.npix);
    DAO_CHECK(hipMemcpyAsync(t, bgs, sizeof(double) * e.npix, hipMemcpyHostToDevice, e.st));
    d = t;
  }
  e.new_backgrounds_all_dev(d);
  e.sync();
}

void daoBatchAnscombe(daoBatch* b) {
  Engine& e = *b->eng;
  dao::launch_anscombe(e.px, 0, e.npix, e.st);
  e.sync();
}

void daoBatchNewPeaks(daoBatch* b, const double* rows, const int* counts, const char* p_type, int on_device) {
  Engine& e = *b->eng;
  int fmt;
  if (!std::strcmp(p_type, "finder")) {
    fmt = dao::PF_FINDER;
  } else if (!std::strcmp(p_type, "testing")) {
    fmt = dao::PF_TESTING;
  } else if (!std::strcmp(p_type, "text") || !std::strcmp(p_type, "hdf5")) {
    fmt = dao::PF_TEXT;
  } else {
    std::printf("Unrecognized peak type '%s'!\n", p_type);
    return;
  }
  int width = fmt == dao::PF_TEXT ? 7 : 4;
  long long total = 0;
  for (int f = 0; f < e.F; f++) total += counts[f];
  if (total == 0) return;
  const double* d = rows;
  if (!on_device) {
    double* t = tmp(b, sizeof(double) * total * width);
    DAO_CHECK(hipMemcpyAsync(t, rows, sizeof(double) * total * width, hipMemcpyHostToDevice, e.st));
    d = t;
  }
  std::vector<int> frames, cnt;
  std::vector<const double*> ptrs;
  long long at = 0;
  for (int f = 0; f < e.F; f++) {
    frames.push_back(f);
    cnt.push_back(counts[f]);
    ptrs.push_back(d + at * width);
    at += counts[f];
  }
  e.new_peaks_multi(frames, ptrs, cnt, fmt);
  e.sync();
  pull_counters(b);
}

int daoBatchFit(daoBatch* b, int max_iterations) {
  int s = b->eng->fit(max_iterations);
  pull_counters(b);
  return s;
}

int daoBatchNFit(daoBatch* b, int f) { return b->eng->fr[f].nfit; }

void daoBatchGetPeaks(daoBatch* b, int f, double* out, int32_t* status) {
  Engine& e = *b->eng;
  int n = e.fr[f].nfit;
  std::vector<dao::PeakRec> r;
  e.get_records(f, r);
  for (int i = 0; i < n; i++) {
    const dao::PeakRec& p = r[i];
    double* o = out + (size_t)i * 9;
    o[0] = p.prm[dao::P_XC] + e.m.xoff;
    o[1] = p.prm[dao::P_YC] + e.m.xoff;
    o[2] = p.prm[dao::P_ZC] + e.m.zoff;
    o[3] = p.prm[dao::P_HEIGHT];
    o[4] = p.prm[dao::P_BG];
    o[5] = std::sqrt(1.0 / (2.0 * p.prm[dao::P_XW]));
    o[6] = std::sqrt(1.0 / (2.0 * p.prm[dao::P_YW]));
    o[7] = p.err;
    o[8] = (double)p.iter;
    status[i] = p.status;
  }
}

void daoBatchGetAllPeaks(daoBatch* b, double* out, int32_t* status) { b->eng->pack_all(out, status); }

void daoBatchCounters(daoBatch* b, int f, int32_t* out8) {
  for (int q = 0; q < dao::C_COUNT; q++) {
    out8[q] = b->counters[(size_t)f * dao::C_COUNT + q];
    b->counters[(size_t)f * dao::C_COUNT + q] = 0;
  }
}

void daoBatchGetFitImage(daoBatch* b, int f, double* out) { b->eng->get_image(f, out, 0); }

void daoBatchRemoveErrorPeaks(daoBatch* b) {
  Engine& e = *b->eng;
  for (int f = 0; f < e.F; f++) {
    std::vector<int32_t> st, keep;
    e.get_status(f, st);
    for (int i = 0; i < (int)st.size(); i++)
      if (st[i] != dao::ST_ERROR) keep.push_back(i);
    e.compact(f, keep);
  }
}

void daoBatchSetTiming(daoBatch* b, int on) { b->eng->timing = on != 0; }

void daoBatchStats(daoBatch* b, double* out) {
  Engine& e = *b->eng;
  out[0] = e.lm_ms;
  out[1] = (double)e.lm_launches;
  out[2] = (double)e.sweeps_total;
  out[3] = (double)e.rounds_total;
  out[4] = (double)e.escapes_total;
  out[5] = (double)e.max_attempts_seen;
}

}  // extern "C"
