// C-ABI of libdao_fit.so: the entry points of the reference's libdao_fit
// (bound by storm_analysis/sa_library/dao_fit_c.py:147-238), each returning
// only when its GPU work is done.  The handle is a host struct whose leading
// fields have the reference's fitData layout (multi_fit.h:80-143) so that the
// Python surface can read the counters, nfit, max_nfit, jac_size and the ROI
// tables directly; the device state lives behind it in a dao::Engine.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "dao_handle.h"

using dao::Engine;

namespace {

constexpr int kPeakBlock = 500;  // INCNPEAKS (multi_fit.h:49)

using Handle = dao::FitHandle;

Handle* H(fitData* fd) { return dao::handle_of(fd); }

// Counter increments of the handle's frame; the other frames' increments (a
// multiplane fitter's other channels) are not read here, so they stay on
// the device for their own handles.
void add_counters(Handle* h) {
  std::vector<int> c;
  h->eng->read_counters(c, h->frame);
  dao::add_frame_counters(h->pub, c, 0);
}

void set_model(Handle* h, int shape, int errm, int jac) {
  h->model.shape = shape;
  h->model.errm = errm;
  h->model.jac = jac;
  h->pub.jac_size = jac;
  h->eng->set_model(h->model);
}


// The frame fitter leaves a frame whose scheduler wait gave up empty; for
// the single-frame handle that is the call's failure.
void check_frame(Handle* h) {
  if (h->eng->failed_count > 0) {
    h->pub.nfit = 0;
    dao::fail(dao::kErrFrame, "a scheduler wait of the fitter exceeded its bound; the frame's peaks were dropped");
  }
}

// Entry points run their body through dao::guarded: a failure becomes the
// calling thread's last error (daoGetLastError) and the call returns.
template <class Fn>
void run(Fn&& fn) {
  (void)dao::guarded(fn);
}
template <class T, class Fn>
T run_value(T on_error, Fn&& fn) {
  T v = on_error;
  if (dao::guarded([&] { v = fn(); }) != dao::kOk) return on_error;
  return v;
}

// A multiplane channel (a handle that does not own its engine) is fitted
// through mpIterateLM with its group; a single-channel sweep is refused.
void refuse_channel_view(Handle* h) {
  if (!h->own) dao::fail(dao::kErrArg, "a multiplane channel is fitted by mpIterateLM / mpDoFit");
}

int peak_format(const char* p_type) {
  if (!std::strcmp(p_type, "finder")) return dao::PF_FINDER;
  if (!std::strcmp(p_type, "testing")) return dao::PF_TESTING;
  if (!std::strcmp(p_type, "text") || !std::strcmp(p_type, "hdf5")) return dao::PF_TEXT;
  return -1;
}

}  // namespace

namespace dao {

// daoNewPeaks (dao_fit.c:1582-1766) on a handle's frame, with the storage
// policy of mFitNewPeaks (multi_fit.c:1428-1477): growing drops ERROR peaks
// and rounds the capacity to blocks of 500.  Throws on failure.
void handle_new_peaks(FitHandle* h, const double* peak_params, const char* p_type, int n_peaks) {
  fitData& p = h->pub;
  const int fmt = peak_format(p_type);
  if (p.nfit + n_peaks > p.max_nfit) {
    int n_alloc = kPeakBlock * ((p.nfit + n_peaks) / kPeakBlock + 1);
    std::vector<int32_t> st, keep;
    h->eng->get_status(h->frame, st);
    for (int i = 0; i < (int)st.size(); i++)
      if (st[i] != ST_ERROR) keep.push_back(i);
    h->eng->compact(h->frame, keep);
    p.max_nfit = n_alloc;
    p.nfit = (int)keep.size();
  }
  if (fmt < 0) {
    // the reference allocates and initialises the slots, then leaves them
    // untouched (dao_fit.c:1753-1755)
    std::printf("Unrecognized peak type '%s'!\n", p_type);
    return;
  }
  if (n_peaks > 0) {
    h->eng->new_peaks(h->frame, peak_params, fmt, n_peaks);
    p.nfit += n_peaks;
  }
  std::vector<int> c;
  h->eng->read_counters(c, h->frame);
  add_frame_counters(p, c, 0);
}

}  // namespace dao

extern "C" {

fitData* daoInitialize(double* rqe, double* scmos_cal, double tol, int im_size_x, int im_size_y, int roi_size) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "dao_fit (HIP): no GPU device is available\n");
    dao::set_last_error(dao::kErrHip, "no GPU device is available");
    return nullptr;
  }
  if (roi_size < 1 || roi_size > 32 || im_size_x < 1 || im_size_y < 1) {
    std::fprintf(stderr, "dao_fit (HIP): unsupported sizes (roi %d, image %d x %d)\n", roi_size, im_size_x,
                 im_size_y);
    dao::set_last_error(dao::kErrArg, "unsupported sizes (roi_size must be 1..32)");
    return nullptr;
  }
  Handle* h = new Handle();
  std::memset(&h->pub, 0, sizeof(fitData));
  fitData& p = h->pub;
  p.image_size_x = im_size_x;
  p.image_size_y = im_size_y;
  p.tolerance = tol;
  p.minimum_height = -1.0;
  p.fit_size_x = roi_size;
  p.fit_size_y = roi_size;
  p.xoff = 0.5 * (double)roi_size - 1.0;
  p.yoff = 0.5 * (double)roi_size - 1.0;
  p.zoff = 0.0;

  dao::Model& m = h->model;
  std::memset(&m, 0, sizeof(m));
  m.shape = dao::SH_2DFIXED;
  m.errm = dao::EM_MLE;
  m.roi = roi_size;
  m.jac = 4;
  m.margin = 2;
  m.xoff = p.xoff;
  m.zoff = 0.0;
  m.tol = tol;
  m.min_height = -1.0;

  dao::roi_table(roi_size, h->roi_x, h->roi_y);
  p.roi_n_index = (int)h->roi_x.size();
  p.roi_x_index = h->roi_x.data();
  p.roi_y_index = h->roi_y.data();

  std::vector<dao::FrameSpec> spec(1);
  spec[0].W = im_size_x;
  spec[0].H = im_size_y;
  spec[0].cap = 0;
  const int rc = dao::guarded([&] {
    h->own.reset(new Engine(m, spec));
    h->eng = h->own.get();
    h->frame = 0;
    h->eng->set_calibration(0, rqe, scmos_cal);
  });
  if (rc != dao::kOk) {
    char msg[256];
    const int code = dao::take_last_error(msg, sizeof(msg));
    std::fprintf(stderr, "dao_fit (HIP): daoInitialize: %s\n", msg);
    dao::set_last_error(code, msg);
    delete h;
    return nullptr;
  }
  return &h->pub;
}

void daoInitialize2DFixed(fitData* fd) { set_model(H(fd), dao::SH_2DFIXED, dao::EM_MLE, 4); }
void daoInitialize2DFixedALS(fitData* fd) { set_model(H(fd), dao::SH_2DFIXED, dao::EM_ALS, 4); }
void daoInitialize2DFixedLS(fitData* fd) { set_model(H(fd), dao::SH_2DFIXED, dao::EM_LS, 4); }
void daoInitialize2DFixedDWLS(fitData* fd) { set_model(H(fd), dao::SH_2DFIXED, dao::EM_DWLS, 4); }
void daoInitialize2DFixedFWLS(fitData* fd) { set_model(H(fd), dao::SH_2DFIXED, dao::EM_FWLS, 4); }

static void init_2d(fitData* fd, int errm, double wmin, double wmax) {
  Handle* h = H(fd);
  h->model.wmin = wmin;
  h->model.wmax = wmax;
  set_model(h, dao::SH_2D, errm, 5);
}
void daoInitialize2D(fitData* fd, double a, double b) { init_2d(fd, dao::EM_MLE, a, b); }
void daoInitialize2DALS(fitData* fd, double a, double b) { init_2d(fd, dao::EM_ALS, a, b); }
void daoInitialize2DLS(fitData* fd, double a, double b) { init_2d(fd, dao::EM_LS, a, b); }
void daoInitialize2DDWLS(fitData* fd, double a, double b) { init_2d(fd, dao::EM_DWLS, a, b); }
void daoInitialize2DFWLS(fitData* fd, double a, double b) { init_2d(fd, dao::EM_FWLS, a, b); }

void daoInitialize3D(fitData* fd, double wmin, double wmax) {
  Handle* h = H(fd);
  h->model.wmin = wmin;
  h->model.wmax = wmax;
  set_model(h, dao::SH_3D, dao::EM_MLE, 6);
}

void daoInitializeZ(fitData* fd, double* wx, double* wy, double z_min, double z_max) {
  Handle* h = H(fd);
  for (int i = 0; i < 5; i++) {
    h->model.wxz[i] = wx[i];
    h->model.wyz[i] = wy[i];
  }
  h->model.wxz[0] = h->model.wxz[0] * h->model.wxz[0];
  h->model.wyz[0] = h->model.wyz[0] * h->model.wyz[0];
  h->model.zmin = z_min;
  h->model.zmax = z_max;
  set_model(h, dao::SH_Z, dao::EM_MLE, 5);
}

void daoNewPeaks(fitData* fd, double* peak_params, char* p_type, int n_peaks) {
  run([&] { dao::handle_new_peaks(H(fd), peak_params, p_type, n_peaks); });
}

void daoCleanup(fitData* fd) {
  if (fd == nullptr) return;
  Handle* h = H(fd);
  if (!h->own) return;  // a view of a multiplane fitter's channel: mpCleanup frees it
  h->own.reset();
  delete h;
}

void mFitNewImage(fitData* fd, double* image) {
  run([&] {
    Handle* h = H(fd);
    h->pub.nfit = 0;
    h->eng->new_image(h->frame, image);
  });
}

void mFitNewBackground(fitData* fd, double* background) {
  run([&] { H(fd)->eng->new_background(H(fd)->frame, background); });
}

void mFitAnscombeTransformImage(fitData* fd) {
  run([&] { H(fd)->eng->anscombe(H(fd)->frame); });
}

void mFitIterateLM(fitData* fd) {
  run([&] {
    Handle* h = H(fd);
    refuse_channel_view(h);
    if (h->eng->use_rounds)
      h->eng->sweep();
    else
      h->eng->fit_frames(0);
    add_counters(h);
    check_frame(h);
  });
}

// doFit in one call: one sweep, then more while any peak is RUNNING, at most
// max_iterations more (dao_fit_c.py:332-356); the same result as calling
// mFitIterateLM / mFitGetUnconverged in that loop.  Returns the sweeps run
// (-1 on failure).
int mFitDoFit(fitData* fd, int max_iterations) {
  return run_value(-1, [&] {
    Handle* h = H(fd);
    refuse_channel_view(h);
    const int s = h->eng->use_rounds ? h->eng->fit(max_iterations) : h->eng->fit_frames(max_iterations);
    add_counters(h);
    check_frame(h);
    return s;
  });
}

int mFitGetUnconverged(fitData* fd) {
  return run_value(-1, [&] { return H(fd)->eng->count_status(H(fd)->frame, dao::ST_RUNNING); });
}

int mFitGetNError(fitData* fd) {
  return run_value(-1, [&] { return H(fd)->eng->count_status(H(fd)->frame, dao::ST_ERROR); });
}

void mFitGetFitImage(fitData* fd, double* out) {
  run([&] { H(fd)->eng->get_image(H(fd)->frame, out, 0); });
}

void mFitGetResidual(fitData* fd, double* out) {
  run([&] { H(fd)->eng->get_image(H(fd)->frame, out, 1); });
}

void mFitGetPeakPropertyDouble(fitData* fd, double* values, char* what) {
  run([&] {
    Handle* h = H(fd);
    Engine& e = *h->eng;
    const fitData& p = h->pub;
    std::vector<double> v;
    int n = p.nfit;
    auto put = [&](double off) {
      for (int i = 0; i < n; i++) values[i] = v[i] + off;
    };
    if (!std::strcmp(what, "background")) {
      e.get_param(h->frame, dao::P_BG, v);
      put(0.0);
    } else if (!std::strcmp(what, "bg_sum")) {
      e.get_sums(h->frame, 0, v);
      put(0.0);
    } else if (!std::strcmp(what, "error")) {
      e.get_error(h->frame, v);
      put(0.0);
    } else if (!std::strcmp(what, "fg_sum") || !std::strcmp(what, "fg_sum_sc")) {
      e.get_sums(h->frame, 1, v);
      put(0.0);
    } else if (!std::strcmp(what, "height")) {
      e.get_param(h->frame, dao::P_HEIGHT, v);
      put(0.0);
    } else if (!std::strcmp(what, "jacobian")) {
      e.get_jacobian(h->frame, v);
      for (size_t i = 0; i < v.size(); i++) values[i] = v[i];
    } else if (!std::strcmp(what, "sum")) {
      e.get_sums(h->frame, 2, v);
      put(0.0);
    } else if (!std::strcmp(what, "x")) {
      e.get_param(h->frame, dao::P_XC, v);
      put(p.xoff);
    } else if (!std::strcmp(what, "xsigma")) {
      e.get_param(h->frame, dao::P_XW, v);
      for (int i = 0; i < n; i++) values[i] = std::sqrt(1.0 / (2.0 * v[i]));
    } else if (!std::strcmp(what, "y")) {
      e.get_param(h->frame, dao::P_YC, v);
      put(p.yoff);
    } else if (!std::strcmp(what, "ysigma")) {
      e.get_param(h->frame, dao::P_YW, v);
      for (int i = 0; i < n; i++) values[i] = std::sqrt(1.0 / (2.0 * v[i]));
    } else if (!std::strcmp(what, "z")) {
      e.get_param(h->frame, dao::P_ZC, v);
      put(p.zoff);
    } else {
      std::printf("Unrecognized parameter '%s'!\n", what);
    }
  });
}

void mFitGetPeakPropertyInt(fitData* fd, int32_t* values, char* what) {
  run([&] {
    Handle* h = H(fd);
    std::vector<int32_t> v;
    if (!std::strcmp(what, "iterations")) {
      h->eng->get_iterations(h->frame, v);
    } else if (!std::strcmp(what, "status")) {
      h->eng->get_status(h->frame, v);
    } else {
      std::printf("Unrecognized parameter '%s'!\n", what);
      return;
    }
    for (size_t i = 0; i < v.size(); i++) values[i] = v[i];
  });
}

void mFitRemoveErrorPeaks(fitData* fd) {
  run([&] {
    Handle* h = H(fd);
    std::vector<int32_t> st, keep;
    h->eng->get_status(h->frame, st);
    for (int i = 0; i < (int)st.size(); i++)
      if (st[i] != dao::ST_ERROR) keep.push_back(i);
    h->eng->compact(h->frame, keep);
    h->pub.nfit = (int)keep.size();
  });
}

void mFitRemoveRunningPeaks(fitData* fd) {
  run([&] {
    Handle* h = H(fd);
    std::vector<int32_t> st, keep;
    h->eng->get_status(h->frame, st);
    int dropped = 0;
    for (int i = 0; i < (int)st.size(); i++) {
      if (st[i] != dao::ST_RUNNING)
        keep.push_back(i);
      else
        dropped++;
    }
    h->eng->compact(h->frame, keep);
    h->pub.n_non_converged += dropped;
    h->pub.nfit = (int)keep.size();
  });
}

void mFitSetPeakStatus(fitData* fd, int32_t* status) {
  run([&] { H(fd)->eng->set_status(H(fd)->frame, status); });
}

void daoSetTeam(fitData* fd, int peaks_per_workgroup) {
  Handle* h = H(fd);
  h->model.team = peaks_per_workgroup < 0 ? 0 : peaks_per_workgroup;
  h->eng->set_model(h->model);
}

int daoGetLastError(char* msg, int len) { return dao::take_last_error(msg, len); }

void daoSetWaitLimit(fitData* fd, long long polls) { H(fd)->eng->spin_limit = polls > 0 ? polls : 0; }

}  // extern "C"
