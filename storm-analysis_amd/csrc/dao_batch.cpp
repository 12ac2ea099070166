// Batch entry points of libdao_fit.so (declared in include/dao_batch_amd.h):
// many frames resident on the GPU, fitted by one executor.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "dao_batch_handle.h"

using dao::Engine;

namespace {

void pull_counters(daoBatch* b) {
  std::vector<int> c;
  b->eng->read_counters(c);
  for (size_t i = 0; i < c.size(); i++) b->counters[i] += c[i];
}

double* tmp(daoBatch* b, size_t bytes) {
  if (bytes > b->tmp_bytes) {
    if (b->d_tmp) (void)hipFree(b->d_tmp);
    b->d_tmp = nullptr;
    b->tmp_bytes = 0;
    DAO_CHECK(hipMalloc((void**)&b->d_tmp, bytes));
    b->tmp_bytes = bytes;
  }
  return b->d_tmp;
}

void check_frame_index(const Engine& e, int f) {
  if (f < 0 || f >= e.F) dao::fail(dao::kErrArg, "frame index out of range");
}

}  // namespace

int dao_peak_format(const char* p_type) {
  if (p_type != nullptr) {
    if (!std::strcmp(p_type, "finder")) return dao::PF_FINDER;
    if (!std::strcmp(p_type, "testing")) return dao::PF_TESTING;
    if (!std::strcmp(p_type, "text") || !std::strcmp(p_type, "hdf5")) return dao::PF_TEXT;
  }
  std::printf("Unrecognized peak type '%s'!\n", p_type ? p_type : "(null)");
  return -1;
}

namespace {

int peak_format(const char* p_type) { return dao_peak_format(p_type); }

}  // namespace

extern "C" {

// Every entry point runs through dao::guarded: 0 (or a count) on success, a
// negative dao::ErrCode on failure, with the message kept as the calling
// thread's last error (daoGetLastError).  Nothing aborts the process.

daoBatch* daoBatchCreate(const daoBatchModel* bm, int n_frames, const int* widths, const int* heights,
                         const int* peak_caps) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "dao_fit (HIP): no GPU device is available\n");
    dao::set_last_error(dao::kErrHip, "no GPU device is available");
    return nullptr;
  }
  if (bm == nullptr || bm->roi_size < 1 || bm->roi_size > 32 || n_frames < 1) {
    dao::set_last_error(dao::kErrArg, "unsupported model or frame count (roi_size must be 1..32)");
    return nullptr;
  }
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
    if (widths[f] < 1 || heights[f] < 1 || peak_caps[f] < 0) {
      dao::set_last_error(dao::kErrArg, "frame sizes must be positive and capacities non-negative");
      return nullptr;
    }
    spec[f].W = widths[f];
    spec[f].H = heights[f];
    spec[f].cap = peak_caps[f];
  }
  daoBatch* b = new daoBatch();
  const int rc = dao::guarded([&] {
    b->eng.reset(new Engine(m, spec));
    b->counters.assign((size_t)n_frames * dao::C_COUNT, 0);
    // default calibration: rqe = 1, scmos = 0
    for (int f = 0; f < n_frames; f++) b->eng->set_calibration_const(f, 1.0, 0.0);
    b->eng->sync();
  });
  if (rc != dao::kOk) {
    char msg[256];
    const int code = dao::take_last_error(msg, sizeof(msg));
    std::fprintf(stderr, "dao_fit (HIP): daoBatchCreate: %s\n", msg);
    dao::set_last_error(code, msg);
    daoBatchDestroy(b);
    return nullptr;
  }
  return b;
}

void daoBatchDestroy(daoBatch* b) {
  if (!b) return;
  b->eng.reset();
  if (b->d_tmp) (void)hipFree(b->d_tmp);
  delete b;
}

int daoBatchSetCalibration(daoBatch* b, int f, const double* rqe, const double* scmos) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    check_frame_index(e, f);
    if (!rqe && !scmos) {
      e.set_calibration_const(f, 1.0, 0.0);
      e.sync();
      return;
    }
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
  });
}

int daoBatchNewImages(daoBatch* b, const double* images, int on_device) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    const double* d = images;
    if (!on_device) {
      const long long n = e.frame_pixels_total();
      double* t = tmp(b, sizeof(double) * n);
      DAO_CHECK(hipMemcpyAsync(t, images, sizeof(double) * n, hipMemcpyHostToDevice, e.st));
      d = t;
    }
    e.new_images_all_dev(d);
    e.sync();
  });
}

int daoBatchNewBackgrounds(daoBatch* b, const double* bgs, int on_device) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    const double* d = bgs;
    if (!on_device) {
      const long long n = e.frame_pixels_total();
      double* t = tmp(b, sizeof(double) * n);
      DAO_CHECK(hipMemcpyAsync(t, bgs, sizeof(double) * n, hipMemcpyHostToDevice, e.st));
      d = t;
    }
    e.new_backgrounds_all_dev(d);
    e.sync();
  });
}

int daoBatchNewFrames(daoBatch* b, const double* images, const double* backgrounds, int on_device) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    const double *di = images, *db = backgrounds;
    if (!on_device) {
      const long long n = e.frame_pixels_total();
      double* t = tmp(b, 2 * sizeof(double) * n);
      DAO_CHECK(hipMemcpyAsync(t, images, sizeof(double) * n, hipMemcpyHostToDevice, e.st));
      DAO_CHECK(hipMemcpyAsync(t + n, backgrounds, sizeof(double) * n, hipMemcpyHostToDevice, e.st));
      di = t;
      db = t + n;
    }
    e.new_frames_all_dev(di, db);
    e.sync();
  });
}

int daoBatchAnscombe(daoBatch* b) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    dao::launch_anscombe(e.px, 0, e.npix, e.st);
    e.sync();
  });
}

int daoBatchLoadFrame(daoBatch* b, int f, int width, int height, const double* image, const double* background,
                      const double* rqe, const double* scmos_cal, const double* rows, int n_rows, const char* p_type,
                      int on_device) {
  int ok = 0;
  const int rc = dao::guarded([&] {
    Engine& e = *b->eng;
    const int fmt = peak_format(p_type);
    if (f < 0 || f >= e.F || fmt < 0 || n_rows < 0 || n_rows > e.fr[f].cap) return;
    if (!e.set_frame_shape(f, width, height)) return;
    const long long n = (long long)width * height;
    const int wid = fmt == dao::PF_TEXT ? 7 : 4;
    // staging: image, background, calibration, rows
    const double *img = image, *bg = background, *rq = rqe, *sc = scmos_cal, *rw = rows;
    if (!on_device) {
      double* t = tmp(b, sizeof(double) * (4 * n + (long long)n_rows * wid));
      DAO_CHECK(hipMemcpyAsync(t, image, sizeof(double) * n, hipMemcpyHostToDevice, e.st));
      DAO_CHECK(hipMemcpyAsync(t + n, background, sizeof(double) * n, hipMemcpyHostToDevice, e.st));
      img = t;
      bg = t + n;
      if (rqe) {
        DAO_CHECK(hipMemcpyAsync(t + 2 * n, rqe, sizeof(double) * n, hipMemcpyHostToDevice, e.st));
        rq = t + 2 * n;
      }
      if (scmos_cal) {
        DAO_CHECK(hipMemcpyAsync(t + 3 * n, scmos_cal, sizeof(double) * n, hipMemcpyHostToDevice, e.st));
        sc = t + 3 * n;
      }
      if (n_rows > 0) {
        DAO_CHECK(hipMemcpyAsync(t + 4 * n, rows, sizeof(double) * n_rows * wid, hipMemcpyHostToDevice, e.st));
        rw = t + 4 * n;
      }
    }
    if (rq && sc) {
      dao::launch_set_calibration(e.px, e.fr[f].pix0, n, rq, sc, e.st);
    } else {
      e.set_calibration_const(f, 1.0, 0.0);
      if (rq || sc) std::printf("daoBatchLoadFrame: rqe and scmos_cal go together; using rqe = 1, scmos = 0\n");
    }
    e.new_image_dev(f, img);
    e.new_background_dev(f, bg);
    if (e.m.errm == dao::EM_ALS) dao::launch_anscombe(e.px, e.fr[f].pix0, n, e.st);
    if (n_rows > 0) e.new_peaks_multi({f}, {rw}, {n_rows}, fmt);
    e.sync();
    pull_counters(b);
    ok = 1;
  });
  return rc != dao::kOk ? rc : ok;
}

int daoBatchLoadFrames(daoBatch* b, int n, const int* frames, const int* widths, const int* heights,
                       const double* const* images, const double* const* backgrounds, const double* const* rows,
                       const int* n_rows, const char* p_type, int on_device) {
  int ok = 0;
  const int rc = dao::guarded([&] {
    Engine& e = *b->eng;
    const int fmt = peak_format(p_type);
    if (fmt < 0 || n < 0) return;
    const int wid = fmt == dao::PF_TEXT ? 7 : 4;
    for (int k = 0; k < n; k++) {
      const int f = frames[k];
      if (f < 0 || f >= e.F || n_rows[k] < 0 || n_rows[k] > e.fr[f].cap) return;
      const int cell = e.sc.cell;
      const long long px = (long long)widths[k] * heights[k];
      const int cells = ((widths[k] + cell - 1) / cell) * ((heights[k] + cell - 1) / cell);
      if (widths[k] < 1 || heights[k] < 1 || px > e.pix_cap[f] || cells > e.cell_cap[f]) return;
    }
    // host arrays go through one staging buffer
    std::vector<const double*> img(images, images + n), bg(backgrounds, backgrounds + n), rw(rows, rows + n);
    if (!on_device) {
      long long total = 0;
      for (int k = 0; k < n; k++) total += 2LL * widths[k] * heights[k] + (long long)n_rows[k] * wid;
      double* t = tmp(b, sizeof(double) * (total + 1));
      for (int k = 0; k < n; k++) {
        const long long px = (long long)widths[k] * heights[k];
        DAO_CHECK(hipMemcpyAsync(t, images[k], sizeof(double) * px, hipMemcpyHostToDevice, e.st));
        img[k] = t;
        t += px;
        DAO_CHECK(hipMemcpyAsync(t, backgrounds[k], sizeof(double) * px, hipMemcpyHostToDevice, e.st));
        bg[k] = t;
        t += px;
        if (n_rows[k] > 0) {
          DAO_CHECK(hipMemcpyAsync(t, rows[k], sizeof(double) * n_rows[k] * wid, hipMemcpyHostToDevice, e.st));
          rw[k] = t;
          t += (long long)n_rows[k] * wid;
        }
      }
    }
    for (int k = 0; k < n; k++) e.set_frame_shape(frames[k], widths[k], heights[k], false);
    e.upload_frames();
    std::vector<int> fs, cnt;
    std::vector<const double*> rp;
    for (int k = 0; k < n; k++) {
      const int f = frames[k];
      const long long px = (long long)widths[k] * heights[k];
      e.set_calibration_const(f, 1.0, 0.0);
      dao::launch_new_image(e.px, e.fr[f].pix0, px, img[k], e.st);
      dao::launch_new_background(e.px, e.fr[f].pix0, px, bg[k], e.st);
      if (e.m.errm == dao::EM_ALS) dao::launch_anscombe(e.px, e.fr[f].pix0, px, e.st);
      if (n_rows[k] > 0) {
        fs.push_back(f);
        cnt.push_back(n_rows[k]);
        rp.push_back(rw[k]);
      }
    }
    if (!fs.empty()) e.new_peaks_multi(fs, rp, cnt, fmt);
    e.sync();
    pull_counters(b);
    ok = 1;
  });
  return rc != dao::kOk ? rc : ok;
}

int daoBatchClearFrames(daoBatch* b, int n, const int* frames) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    for (int k = 0; k < n; k++)
      if (frames[k] >= 0 && frames[k] < e.F) e.fr[frames[k]].nfit = 0;
    e.upload_frames();
  });
}

void daoBatchSetTuning(daoBatch* b, int waves, int margin, int team) {
  Engine& e = *b->eng;
  dao::Model m = e.m;
  m.waves = waves < 0 ? 0 : waves;
  if (margin >= 0) m.margin = margin;
  if (team >= 0) m.team = team;
  e.set_model(m);
}

void daoBatchSetWaitLimit(daoBatch* b, long long polls) { b->eng->spin_limit = polls > 0 ? polls : 0; }

void daoBatchSetTeamHandoff(daoBatch* b, int ops) { b->eng->team_solo = ops < 0 ? dao::kTeamSolo : ops; }

int daoBatchClearFrame(daoBatch* b, int f) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    check_frame_index(e, f);
    e.clear_frame(f);
  });
}

int daoBatchNewPeaks(daoBatch* b, const double* rows, const int* counts, const char* p_type, int on_device) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    const int fmt = peak_format(p_type);
    if (fmt < 0) dao::fail(dao::kErrArg, "unrecognized peak type");
    int width = fmt == dao::PF_TEXT ? 7 : 4;
    long long total = 0;
    for (int f = 0; f < e.F; f++) {
      if (counts[f] < 0) dao::fail(dao::kErrArg, "negative peak count");
      // checked before any work: a refused call changes nothing
      if ((long long)e.fr[f].nfit + counts[f] > e.fr[f].cap) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "frame %d: %d peaks held + %d new exceed its capacity of %d", f,
                      e.fr[f].nfit, counts[f], e.fr[f].cap);
        dao::fail(dao::kErrCapacity, msg);
      }
      total += counts[f];
    }
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
  });
}

int daoBatchFit(daoBatch* b, int max_iterations) {
  int s = 0;
  const int rc = dao::guarded([&] {
    s = b->eng->fit(max_iterations);
    pull_counters(b);
    if (b->eng->failed_count > 0) {
      char msg[160];
      std::snprintf(msg, sizeof(msg),
                    "%d frame(s) failed: a scheduler wait exceeded its bound; their peaks were dropped "
                    "(daoBatchFailedFrames names them)",
                    b->eng->failed_count);
      dao::fail(dao::kErrFrame, msg);
    }
  });
  return rc != dao::kOk ? rc : s;
}

int daoBatchFailedFrames(daoBatch* b, int32_t* flags) {
  const Engine& e = *b->eng;
  for (int f = 0; f < e.F; f++) flags[f] = (int)e.failed.size() == e.F ? e.failed[f] : 0;
  return e.failed_count;
}

int daoBatchNFrames(daoBatch* b) { return b->eng->F; }

int daoBatchNFit(daoBatch* b, int f) {
  if (f < 0 || f >= b->eng->F) return dao::kErrArg;
  return b->eng->fr[f].nfit;
}

int daoBatchGetPeaks(daoBatch* b, int f, double* out, int32_t* status) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    check_frame_index(e, f);
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
  });
}

int daoBatchGetAllPeaks(daoBatch* b, double* out, int32_t* status) {
  return dao::guarded([&] { b->eng->pack_all(out, status, false); });
}

int daoBatchGetAllPeaksDevice(daoBatch* b, double* out, int32_t* status) {
  return dao::guarded([&] { b->eng->pack_all(out, status, true); });
}

void daoBatchCounters(daoBatch* b, int f, int32_t* out8) {
  if (f < 0 || f >= b->eng->F) return;
  for (int q = 0; q < dao::C_COUNT; q++) {
    out8[q] = b->counters[(size_t)f * dao::C_COUNT + q];
    b->counters[(size_t)f * dao::C_COUNT + q] = 0;
  }
}

int daoBatchGetFitImage(daoBatch* b, int f, double* out) {
  return dao::guarded([&] {
    check_frame_index(*b->eng, f);
    b->eng->get_image(f, out, 0);
  });
}

int daoBatchGetFitImages(daoBatch* b, double* out, int residual, int on_device) {
  return dao::guarded([&] { b->eng->get_images_all(out, residual, on_device != 0); });
}

int daoBatchRemoveErrorPeaks(daoBatch* b) {
  return dao::guarded([&] { b->eng->remove_all(dao::ST_ERROR); });
}

int daoBatchRemoveRunningPeaks(daoBatch* b) {
  return dao::guarded([&] {
    b->eng->remove_all(dao::ST_RUNNING);
    pull_counters(b);
  });
}

int daoBatchSetPeakStatus(daoBatch* b, const int32_t* status, int on_device) {
  return dao::guarded([&] {
    Engine& e = *b->eng;
    const long long n = e.total_nfit();
    if (n == 0) return;
    if (status == nullptr) dao::fail(dao::kErrArg, "status is NULL");
    if (on_device) {
      e.set_status_all(status);
      return;
    }
    int32_t* d = reinterpret_cast<int32_t*>(tmp(b, sizeof(int32_t) * (size_t)n));
    DAO_CHECK(hipMemcpy(d, status, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice));
    e.set_status_all(d);
  });
}

int daoBatchGetSignificanceDevice(daoBatch* b, double* out) {
  return dao::guarded([&] {
    if (b->eng->total_nfit() > 0 && out == nullptr) dao::fail(dao::kErrArg, "out is NULL");
    b->eng->significance_all(out);
  });
}

int daoBatchFitActive(daoBatch* b, int max_iterations, const int32_t* active) {
  Engine& e = *b->eng;
  if (active == nullptr) return daoBatchFit(b, max_iterations);
  e.fit_mask.assign(e.F, 0);
  for (int f = 0; f < e.F; f++) e.fit_mask[f] = active[f] != 0;
  const int rc = daoBatchFit(b, max_iterations);
  e.fit_mask.clear();
  return rc;
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
