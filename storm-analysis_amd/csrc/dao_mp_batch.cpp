// Batched multiplane fitting (declared in include/mp_fit_amd.h): many
// frame-sets (a frame of every channel) in one engine, fitted by one k_fit_mp
// launch with one wavefront per frame-set.  The arithmetic is mpIterateLM's
// (dao_mp.hip), as for the single-set handle; frame f of the engine is
// channel f % C of set f / C.  No counterpart in the reference, which fits
// one frame-set at a time.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "dao_engine.h"
#include "../../include/mp_fit_amd.h"

using dao::Engine;

struct mpBatch {
  std::unique_ptr<Engine> eng;
  int S = 0, C = 0;
  std::vector<dao::Model> ch_model;  // per channel: the width range
  std::vector<double> xt0N, yt0N, xtN0, ytN0;
  std::vector<double> wmin, wmax, w_x, w_y;
  int n_weights = 0;
  double w_z_offset = 0.0, w_z_scale = 0.0;
  double tol = 1e-6;
  std::vector<int> counters;  // [F][C_COUNT], not yet handed out
  double* d_cpl = nullptr;
  int* d_set0 = nullptr;
  int* d_sweeps = nullptr;
  double* d_tmp = nullptr;
  size_t tmp_bytes = 0;

  ~mpBatch() {
    if (d_cpl) (void)hipFree(d_cpl);
    if (d_set0) (void)hipFree(d_set0);
    if (d_sweeps) (void)hipFree(d_sweeps);
    if (d_tmp) (void)hipFree(d_tmp);
  }
};

namespace {

double* tmp(mpBatch* b, size_t bytes) {
  if (bytes > b->tmp_bytes) {
    if (b->d_tmp) (void)hipFree(b->d_tmp);
    b->d_tmp = nullptr;
    b->tmp_bytes = 0;
    DAO_CHECK(hipMalloc((void**)&b->d_tmp, bytes));
    b->tmp_bytes = bytes;
  }
  return b->d_tmp;
}

void pull_counters(mpBatch* b) {
  std::vector<int> c;
  b->eng->read_counters(c);
  for (size_t i = 0; i < c.size(); i++) b->counters[i] += c[i];
}

int peak_format(const char* p_type) {
  if (p_type == nullptr) return -1;
  if (!std::strcmp(p_type, "finder")) return dao::PF_FINDER;
  if (!std::strcmp(p_type, "testing")) return dao::PF_TESTING;
  if (!std::strcmp(p_type, "text") || !std::strcmp(p_type, "hdf5")) return dao::PF_TEXT;
  return -1;
}

// mpCheckError over every group of every set (mp_fit.c:44-67): statuses are
// consistent across channels except for groups just added, so this is the
// reference's check of the new groups.
void sync_errors(mpBatch* b) {
  Engine& e = *b->eng;
  const long long total = e.total_nfit();
  if (total == 0) return;
  std::vector<double> rows((size_t)total * 9);
  std::vector<int32_t> st((size_t)total);
  e.pack_all(rows.data(), st.data(), false);
  std::vector<long long> off(e.F + 1, 0);
  for (int f = 0; f < e.F; f++) off[f + 1] = off[f] + e.fr[f].nfit;
  bool change = false;
  for (int s = 0; s < b->S; s++) {
    const int n = e.fr[s * b->C].nfit;
    for (int i = 0; i < n; i++) {
      bool bad = false;
      for (int c = 0; c < b->C; c++) bad |= st[off[s * b->C + c] + i] == dao::ST_ERROR;
      if (!bad) continue;
      for (int c = 0; c < b->C; c++) {
        int32_t& v = st[off[s * b->C + c] + i];
        if (v != dao::ST_ERROR) {
          v = dao::ST_ERROR;
          change = true;
        }
      }
    }
  }
  if (!change) return;
  int32_t* d = reinterpret_cast<int32_t*>(tmp(b, sizeof(int32_t) * (size_t)total));
  DAO_CHECK(hipMemcpy(d, st.data(), sizeof(int32_t) * (size_t)total, hipMemcpyHostToDevice));
  e.set_status_all(d);
}

void upload_coupling(mpBatch* b) {
  if (b->n_weights < 2) dao::fail(dao::kErrArg, "weights not set (mpBatchSetWeights)");
  std::vector<double> t;
  for (const std::vector<double>* v : {&b->xt0N, &b->yt0N, &b->xtN0, &b->ytN0, &b->wmin, &b->wmax, &b->w_x, &b->w_y})
    t.insert(t.end(), v->begin(), v->end());
  if (b->d_cpl) (void)hipFree(b->d_cpl);
  b->d_cpl = nullptr;
  DAO_CHECK(hipMalloc((void**)&b->d_cpl, sizeof(double) * t.size()));
  DAO_CHECK(hipMemcpy(b->d_cpl, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
}

template <class Fn>
int run(Fn&& fn) {
  return dao::guarded(fn);
}

}  // namespace

extern "C" {

mpBatch* mpBatchCreate(int n_sets, int n_channels, int im_size_x, int im_size_y, int roi_size, double tolerance,
                       int peak_cap) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    dao::set_last_error(dao::kErrHip, "no GPU device is available");
    return nullptr;
  }
  if (n_sets < 1 || n_channels < 1 || n_channels > dao::kMpMaxChannels || im_size_x < 1 || im_size_y < 1 ||
      roi_size < 1 || roi_size > 32 || peak_cap < 1) {
    dao::set_last_error(dao::kErrArg, "unsupported sets, channels (1..8), image size, roi_size (1..32) or capacity");
    return nullptr;
  }
  std::unique_ptr<mpBatch> b(new mpBatch());
  const int rc = dao::guarded([&] {
    dao::Model m;
    std::memset(&m, 0, sizeof(m));
    m.shape = dao::SH_2D;
    m.errm = dao::EM_MLE;
    m.roi = roi_size;
    m.jac = 5;
    m.margin = 2;
    m.xoff = 0.5 * (double)roi_size - 1.0;
    m.tol = tolerance;
    m.min_height = 1.0;  // mpDaoInitialize2DChannel (mp_fit_dao.c:59)
    std::vector<dao::FrameSpec> spec((size_t)n_sets * n_channels);
    for (auto& q : spec) {
      q.W = im_size_x;
      q.H = im_size_y;
      q.cap = peak_cap;
    }
    b->eng.reset(new Engine(m, spec));
    b->S = n_sets;
    b->C = n_channels;
    b->tol = tolerance;
    b->ch_model.assign(n_channels, b->eng->m);
    const size_t n3 = 3 * (size_t)n_channels;
    b->xt0N.assign(n3, 0.0);
    b->yt0N.assign(n3, 0.0);
    b->xtN0.assign(n3, 0.0);
    b->ytN0.assign(n3, 0.0);
    b->wmin.assign(n_channels, 0.0);
    b->wmax.assign(n_channels, 0.0);
    b->counters.assign((size_t)b->eng->F * dao::C_COUNT, 0);
    std::vector<int> set0(n_sets);
    for (int s = 0; s < n_sets; s++) set0[s] = s * n_channels;
    DAO_CHECK(hipMalloc((void**)&b->d_set0, sizeof(int) * n_sets));
    DAO_CHECK(hipMalloc((void**)&b->d_sweeps, sizeof(int) * (size_t)b->eng->F));
    DAO_CHECK(hipMemcpy(b->d_set0, set0.data(), sizeof(int) * n_sets, hipMemcpyHostToDevice));
  });
  return rc == dao::kOk ? b.release() : nullptr;
}

void mpBatchDestroy(mpBatch* b) { delete b; }

int mpBatchInitializeChannel(mpBatch* b, int channel, const double* rqe, const double* variance, double width_min,
                             double width_max) {
  return run([&] {
    if (channel < 0 || channel >= b->C) dao::fail(dao::kErrArg, "channel out of range");
    for (int s = 0; s < b->S; s++) b->eng->set_calibration(s * b->C + channel, rqe, variance);
    b->wmin[channel] = width_min;
    b->wmax[channel] = width_max;
    b->ch_model[channel].wmin = width_min;
    b->ch_model[channel].wmax = width_max;
  });
}

int mpBatchSetTransforms(mpBatch* b, const double* xt_0toN, const double* yt_0toN, const double* xt_Nto0,
                         const double* yt_Nto0) {
  return run([&] {
    const size_t n3 = 3 * (size_t)b->C;
    std::copy(xt_0toN, xt_0toN + n3, b->xt0N.begin());
    std::copy(yt_0toN, yt_0toN + n3, b->yt0N.begin());
    std::copy(xt_Nto0, xt_Nto0 + n3, b->xtN0.begin());
    std::copy(yt_Nto0, yt_Nto0 + n3, b->ytN0.begin());
  });
}

int mpBatchSetWeights(mpBatch* b, const double* w_x, const double* w_y, int z_size, double z_offset,
                      double z_scale) {
  return run([&] {
    if (z_size < 2) dao::fail(dao::kErrArg, "at least two weight rows");
    const size_t n = (size_t)z_size * b->C;
    b->w_x.assign(w_x, w_x + n);
    b->w_y.assign(w_y, w_y + n);
    b->n_weights = z_size;
    b->w_z_offset = z_offset;
    b->w_z_scale = z_scale;
  });
}

int mpBatchNewImages(mpBatch* b, const double* images, int on_device) {
  return run([&] {
    Engine& e = *b->eng;
    const double* d = images;
    if (!on_device) {
      const long long n = e.frame_pixels_total();
      double* t = tmp(b, sizeof(double) * n);
      DAO_CHECK(hipMemcpy(t, images, sizeof(double) * n, hipMemcpyHostToDevice));
      d = t;
    }
    e.new_images_all_dev(d);
    e.sync();
  });
}

int mpBatchNewBackgrounds(mpBatch* b, const double* bgs, int on_device) {
  return run([&] {
    Engine& e = *b->eng;
    const double* d = bgs;
    if (!on_device) {
      const long long n = e.frame_pixels_total();
      double* t = tmp(b, sizeof(double) * n);
      DAO_CHECK(hipMemcpy(t, bgs, sizeof(double) * n, hipMemcpyHostToDevice));
      d = t;
    }
    e.new_backgrounds_all_dev(d);
    e.sync();
  });
}

int mpBatchNewPeaks(mpBatch* b, const double* rows, const int* counts, const char* p_type, int on_device) {
  return run([&] {
    Engine& e = *b->eng;
    const int fmt = peak_format(p_type);
    if (fmt < 0) dao::fail(dao::kErrArg, "unrecognized peak type");
    const int w = fmt == dao::PF_TEXT ? 7 : 4;
    long long total = 0;
    for (int s = 0; s < b->S; s++) {
      if (counts[s] < 0) dao::fail(dao::kErrArg, "negative peak count");
      if ((long long)e.fr[s * b->C].nfit + counts[s] > e.fr[s * b->C].cap)
        dao::fail(dao::kErrCapacity, "more peaks than a frame-set's capacity");
      total += counts[s];
    }
    if (total == 0) return;
    // channel-0 rows on the host (they are few), mapped per channel
    // (mpDaoNewPeaks, mp_fit_dao.c:82-153)
    std::vector<double> host((size_t)total * w);
    if (on_device)
      DAO_CHECK(hipMemcpy(host.data(), rows, sizeof(double) * host.size(), hipMemcpyDeviceToHost));
    else
      std::copy(rows, rows + host.size(), host.begin());
    std::vector<double> mapped(host.size());
    double* d = tmp(b, sizeof(double) * host.size());
    for (int c = 0; c < b->C; c++) {
      const double* xt = &b->xt0N[3 * c];
      const double* yt = &b->yt0N[3 * c];
      for (long long j = 0; j < total; j++) {
        const double* q = &host[(size_t)j * w];
        double* o = &mapped[(size_t)j * w];
        std::copy(q, q + w, o);
        if (c > 0) {
          o[0] = xt[0] + q[0] * xt[1] + q[1] * xt[2];
          o[1] = yt[0] + q[0] * yt[1] + q[1] * yt[2];
        }
      }
      DAO_CHECK(hipMemcpy(d, mapped.data(), sizeof(double) * mapped.size(), hipMemcpyHostToDevice));
      std::vector<int> frames, cnt;
      std::vector<const double*> ptrs;
      long long at = 0;
      for (int s = 0; s < b->S; s++) {
        frames.push_back(s * b->C + c);
        cnt.push_back(counts[s]);
        ptrs.push_back(d + at * w);
        at += counts[s];
      }
      // daoNewPeaks checks each new peak with the channel's daoCheck2D
      e.set_model(b->ch_model[c]);
      e.new_peaks_multi(frames, ptrs, cnt, fmt);
    }
    pull_counters(b);
    sync_errors(b);
  });
}

int mpBatchFit(mpBatch* b, int max_iterations) {
  int sweeps = 0;
  const int rc = run([&] {
    Engine& e = *b->eng;
    for (int s = 0; s < b->S; s++)
      for (int c = 1; c < b->C; c++)
        if (e.fr[s * b->C + c].nfit != e.fr[s * b->C].nfit)
          dao::fail(dao::kErrArg, "a frame-set's channels hold different numbers of peaks");
    upload_coupling(b);
    dao::MpArgs a;
    a.m = e.m;
    a.px = e.px;
    a.pk = e.pk;
    a.frames = e.d_fr;
    a.roi_tab = e.d_roi;
    a.nch = b->C;
    a.nsets = b->S;
    a.set_frame0 = b->d_set0;
    a.cpl = b->d_cpl;
    a.n_weights = b->n_weights;
    a.w_z_offset = b->w_z_offset;
    a.w_z_scale = b->w_z_scale;
    a.tol = b->tol;
    a.max_more = max_iterations;
    a.frame_counters = e.d_counters;
    a.frame_sweeps = b->d_sweeps;
    DAO_CHECK(hipMemsetAsync(b->d_sweeps, 0, sizeof(int) * (size_t)e.F, e.st));
    dao::launch_fit_mp(a, e.st);
    DAO_CHECK(hipGetLastError());
    std::vector<int> sw(e.F);
    DAO_CHECK(hipMemcpyAsync(sw.data(), b->d_sweeps, sizeof(int) * (size_t)e.F, hipMemcpyDeviceToHost, e.st));
    e.sync();
    for (int v : sw) sweeps = std::max(sweeps, v);
    pull_counters(b);
  });
  return rc != dao::kOk ? rc : sweeps;
}

int mpBatchNFit(mpBatch* b, int set) {
  if (set < 0 || set >= b->S) return -1;
  return b->eng->fr[set * b->C].nfit;
}

int mpBatchGetAllPeaks(mpBatch* b, double* out, int32_t* status, int on_device) {
  return run([&] { b->eng->pack_all(out, status, on_device != 0); });
}

int mpBatchRemoveErrorPeaks(mpBatch* b) {
  return run([&] { b->eng->remove_all(dao::ST_ERROR); });
}

int mpBatchRemoveRunningPeaks(mpBatch* b) {
  return run([&] {
    b->eng->remove_all(dao::ST_RUNNING);
    pull_counters(b);
  });
}

void mpBatchCounters(mpBatch* b, int set, int channel, int32_t* out8) {
  if (set < 0 || set >= b->S || channel < 0 || channel >= b->C) return;
  const size_t f = (size_t)set * b->C + channel;
  for (int q = 0; q < dao::C_COUNT; q++) {
    out8[q] = b->counters[f * dao::C_COUNT + q];
    b->counters[f * dao::C_COUNT + q] = 0;
  }
}

}  // extern "C"
