// C-ABI of the multiplane DAO fitter (declared in include/mp_fit_amd.h): the
// entry points of the reference's libmp_fit_dao (mp_fit.c, mp_fit_dao.c,
// bound by storm_analysis/multi_plane/mp_fit_c.py:55-127 and
// mp_fit_dao_c.py:17-36).  One Engine holds the channels as frames; each
// channel's fitData handle is a view of its frame, so the single-channel
// mFit* entry points read properties and images of a channel as in the
// reference; k_fit_mp (dao_mp.hip) runs the coupled sweeps.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "dao_handle.h"
#include "../../include/mp_fit_amd.h"

using dao::Engine;
using dao::FitHandle;

namespace {

struct MpHandle {
  mpFit pub;  // must stay first
  std::unique_ptr<Engine> eng;
  std::vector<std::unique_ptr<FitHandle>> ch;
  std::vector<fitData*> ch_ptrs;
  std::vector<double> xt0N, yt0N, xtN0, ytN0;
  std::vector<double> w_bg, w_h, w_x, w_y, w_z, heights;
  std::vector<double> wmin, wmax;
  std::vector<int> roi_x, roi_y;
  int roi = 0;
  double* d_cpl = nullptr;
  size_t cpl_doubles = 0;
  bool cpl_dirty = true;
  int* d_set0 = nullptr;
  int* d_sweeps = nullptr;

  ~MpHandle() {
    if (d_cpl) (void)hipFree(d_cpl);
    if (d_set0) (void)hipFree(d_set0);
    if (d_sweeps) (void)hipFree(d_sweeps);
  }
};

MpHandle* M(mpFit* mp) { return reinterpret_cast<MpHandle*>(mp); }

template <class Fn>
void run(Fn&& fn) {
  (void)dao::guarded(fn);
}

// The engine, created with the first channel (its ROI size is the channels').
void ensure_engine(MpHandle* h, int roi_size) {
  if (h->eng) return;
  if (roi_size < 1 || roi_size > 32) dao::fail(dao::kErrArg, "roi_size must be in [1, 32]");
  const int nch = h->pub.n_channels;
  dao::Model m;
  std::memset(&m, 0, sizeof(m));
  m.shape = dao::SH_2D;
  m.errm = dao::EM_MLE;
  m.roi = roi_size;
  m.jac = 5;
  m.margin = 2;
  m.xoff = 0.5 * (double)roi_size - 1.0;
  m.tol = h->pub.tolerance;
  m.min_height = 1.0;  // mpDaoInitialize2DChannel (mp_fit_dao.c:59)
  std::vector<dao::FrameSpec> spec(nch);
  for (int c = 0; c < nch; c++) {
    spec[c].W = h->pub.im_size_x;
    spec[c].H = h->pub.im_size_y;
    spec[c].cap = 0;
  }
  h->eng.reset(new Engine(m, spec));
  h->eng->growable = true;
  h->roi = roi_size;
  dao::roi_table(roi_size, h->roi_x, h->roi_y);
  for (int c = 0; c < nch; c++) {
    std::unique_ptr<FitHandle> v(new FitHandle());
    std::memset(&v->pub, 0, sizeof(fitData));
    fitData& p = v->pub;
    p.image_size_x = h->pub.im_size_x;
    p.image_size_y = h->pub.im_size_y;
    p.tolerance = h->pub.tolerance;
    p.minimum_height = 1.0;
    p.fit_size_x = roi_size;
    p.fit_size_y = roi_size;
    p.xoff = m.xoff;
    p.yoff = m.xoff;
    p.jac_size = 5;
    p.roi_n_index = (int)h->roi_x.size();
    p.roi_x_index = h->roi_x.data();
    p.roi_y_index = h->roi_y.data();
    v->eng = h->eng.get();
    v->frame = c;
    v->model = m;
    h->ch_ptrs[c] = &v->pub;
    h->ch.push_back(std::move(v));
  }
}

void require_engine(MpHandle* h) {
  if (!h->eng) dao::fail(dao::kErrArg, "no channel initialised (mpDaoInitialize2DChannel)");
}

// The coupling table of k_fit_mp (MpArgs::cpl), uploaded when it changed.
void upload_coupling(MpHandle* h) {
  if (!h->cpl_dirty) return;
  const int nch = h->pub.n_channels;
  const int nw = h->pub.n_weights;
  if (nw < 2 || (int)h->w_x.size() < nw * nch) dao::fail(dao::kErrArg, "weights not set (mpSetWeights)");
  std::vector<double> t;
  for (const std::vector<double>* v : {&h->xt0N, &h->yt0N, &h->xtN0, &h->ytN0}) t.insert(t.end(), v->begin(), v->end());
  t.insert(t.end(), h->wmin.begin(), h->wmin.end());
  t.insert(t.end(), h->wmax.begin(), h->wmax.end());
  t.insert(t.end(), h->w_x.begin(), h->w_x.end());
  t.insert(t.end(), h->w_y.begin(), h->w_y.end());
  if (t.size() > h->cpl_doubles) {
    if (h->d_cpl) (void)hipFree(h->d_cpl);
    h->d_cpl = nullptr;
    h->cpl_doubles = 0;
    DAO_CHECK(hipMalloc((void**)&h->d_cpl, sizeof(double) * t.size()));
    h->cpl_doubles = t.size();
  }
  DAO_CHECK(hipMemcpy(h->d_cpl, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
  h->cpl_dirty = false;
}

// Counter increments of every channel to its handle.
void pull_counters(MpHandle* h) {
  std::vector<int> c;
  h->eng->read_counters(c);
  for (int k = 0; k < h->pub.n_channels; k++) dao::add_frame_counters(h->ch[k]->pub, c, k);
}

// mpCheckError for groups [i0, i1) (mp_fit.c:44-67), on the host's view of
// the statuses: a group in ERROR in some channel is set to ERROR (and
// subtracted) in all.
void sync_errors(MpHandle* h, int i0, int i1) {
  const int nch = h->pub.n_channels;
  std::vector<std::vector<int32_t>> st(nch);
  for (int c = 0; c < nch; c++) h->eng->get_status(c, st[c]);
  std::vector<char> bad(i1 > i0 ? i1 - i0 : 0, 0);
  bool any = false;
  for (int i = i0; i < i1; i++)
    for (int c = 0; c < nch; c++)
      if (st[c][i] == dao::ST_ERROR) {
        bad[i - i0] = 1;
        any = true;
      }
  if (!any) return;
  for (int c = 0; c < nch; c++) {
    std::vector<int32_t> s = st[c];
    bool change = false;
    for (int i = i0; i < i1; i++)
      if (bad[i - i0] && s[i] != dao::ST_ERROR) {
        s[i] = dao::ST_ERROR;
        change = true;
      }
    if (change) h->eng->set_status(c, s.data());
  }
}

int fit(MpHandle* h, int max_more) {
  require_engine(h);
  upload_coupling(h);
  Engine& e = *h->eng;
  // k_fit_mp walks channel 0's groups in every channel
  for (int c = 1; c < h->pub.n_channels; c++)
    if (e.fr[c].nfit != e.fr[0].nfit) dao::fail(dao::kErrArg, "the channels hold different numbers of peaks");
  if (!h->d_set0) {
    DAO_CHECK(hipMalloc((void**)&h->d_set0, sizeof(int)));
    DAO_CHECK(hipMalloc((void**)&h->d_sweeps, sizeof(int) * h->pub.n_channels));
    const int zero = 0;
    DAO_CHECK(hipMemcpy(h->d_set0, &zero, sizeof(int), hipMemcpyHostToDevice));
  }
  dao::MpArgs a;
  a.m = e.m;
  a.px = e.px;
  a.pk = e.pk;
  a.frames = e.d_fr;
  a.roi_tab = e.d_roi;
  a.nch = h->pub.n_channels;
  a.nsets = 1;
  a.set_frame0 = h->d_set0;
  a.cpl = h->d_cpl;
  a.n_weights = h->pub.n_weights;
  a.w_z_offset = h->pub.w_z_offset;
  a.w_z_scale = h->pub.w_z_scale;
  a.tol = h->pub.tolerance;
  a.max_more = max_more;
  a.frame_counters = e.d_counters;
  a.frame_sweeps = h->d_sweeps;
  dao::launch_fit_mp(a, e.st);
  DAO_CHECK(hipGetLastError());
  int sweeps = 0;
  DAO_CHECK(hipMemcpyAsync(&sweeps, h->d_sweeps, sizeof(int), hipMemcpyDeviceToHost, e.st));
  e.sync();
  pull_counters(h);
  return sweeps;
}

}  // namespace

extern "C" {

mpFit* mpInitialize(double tolerance, int n_channels, int independent_heights, int im_size_x, int im_size_y) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::fprintf(stderr, "mp_fit (HIP): no GPU device is available\n");
    dao::set_last_error(dao::kErrHip, "no GPU device is available");
    return nullptr;
  }
  if (n_channels < 1 || n_channels > dao::kMpMaxChannels || im_size_x < 1 || im_size_y < 1) {
    dao::set_last_error(dao::kErrArg, "unsupported channel count (1..8) or image size");
    return nullptr;
  }
  MpHandle* h = new MpHandle();
  std::memset(&h->pub, 0, sizeof(mpFit));
  mpFit& p = h->pub;
  p.im_size_x = im_size_x;
  p.im_size_y = im_size_y;
  p.independent_heights = independent_heights;
  p.n_channels = n_channels;
  p.tolerance = tolerance;
  const size_t n3 = 3 * (size_t)n_channels;
  h->xt0N.assign(n3, 0.0);
  h->yt0N.assign(n3, 0.0);
  h->xtN0.assign(n3, 0.0);
  h->ytN0.assign(n3, 0.0);
  h->wmin.assign(n_channels, 0.0);
  h->wmax.assign(n_channels, 0.0);
  h->heights.assign(n_channels, 1.0);
  h->ch_ptrs.assign(n_channels, nullptr);
  p.xt_0toN = h->xt0N.data();
  p.yt_0toN = h->yt0N.data();
  p.xt_Nto0 = h->xtN0.data();
  p.yt_Nto0 = h->ytN0.data();
  p.heights = h->heights.data();
  p.fit_data = h->ch_ptrs.data();
  return &h->pub;
}

void mpDaoInitialize2DChannel(mpFit* mp, double* rqe, double* variance, double width_min, double width_max,
                              int roi_size, int channel) {
  run([&] {
    MpHandle* h = M(mp);
    if (channel < 0 || channel >= h->pub.n_channels) dao::fail(dao::kErrArg, "channel out of range");
    ensure_engine(h, roi_size);
    if (roi_size != h->roi) dao::fail(dao::kErrArg, "every channel uses the same ROI size");
    h->eng->set_calibration(channel, rqe, variance);
    h->wmin[channel] = width_min;
    h->wmax[channel] = width_max;
    // the channel's own model: its width range clamps its new peaks
    h->ch[channel]->model.wmin = width_min;
    h->ch[channel]->model.wmax = width_max;
    h->cpl_dirty = true;
  });
}

void mpSetTransforms(mpFit* mp, double* xt_0toN, double* yt_0toN, double* xt_Nto0, double* yt_Nto0) {
  MpHandle* h = M(mp);
  const size_t n3 = 3 * (size_t)h->pub.n_channels;
  std::copy(xt_0toN, xt_0toN + n3, h->xt0N.begin());
  std::copy(yt_0toN, yt_0toN + n3, h->yt0N.begin());
  std::copy(xt_Nto0, xt_Nto0 + n3, h->xtN0.begin());
  std::copy(yt_Nto0, yt_Nto0 + n3, h->ytN0.begin());
  h->cpl_dirty = true;
}

void mpSetWeights(mpFit* mp, double* w_bg, double* w_h, double* w_x, double* w_y, double* w_z, int z_size) {
  MpHandle* h = M(mp);
  const size_t n = (size_t)h->pub.n_channels * (z_size > 0 ? z_size : 0);
  h->pub.n_weights = z_size;
  h->w_bg.assign(w_bg, w_bg + n);
  h->w_h.assign(w_h, w_h + n);
  h->w_x.assign(w_x, w_x + n);
  h->w_y.assign(w_y, w_y + n);
  h->w_z.assign(w_z, w_z + n);
  h->heights.assign(h->pub.n_channels, 1.0);
  h->pub.w_bg = h->w_bg.data();
  h->pub.w_h = h->w_h.data();
  h->pub.w_x = h->w_x.data();
  h->pub.w_y = h->w_y.data();
  h->pub.w_z = h->w_z.data();
  h->pub.heights = h->heights.data();
  h->cpl_dirty = true;
}

void mpSetWeightsIndexing(mpFit* mp, double z_offset, double z_scale) {
  MpHandle* h = M(mp);
  h->pub.w_z_offset = z_offset;
  h->pub.w_z_scale = z_scale;
}

void mpDaoNewPeaks(mpFit* mp, double* peak_params, char* p_type, int n_peaks) {
  run([&] {
    MpHandle* h = M(mp);
    require_engine(h);
    const int nch = h->pub.n_channels;
    const bool four = !std::strcmp(p_type, "finder") || !std::strcmp(p_type, "testing");
    const int w = four ? 4 : 7;
    const int start = h->ch[0]->pub.nfit;
    std::vector<double> mapped((size_t)std::max(n_peaks, 1) * w);
    for (int c = 0; c < nch; c++) {
      const double* rows = peak_params;
      if (c > 0) {
        const double* xt = &h->xt0N[3 * c];
        const double* yt = &h->yt0N[3 * c];
        for (int j = 0; j < n_peaks; j++) {
          const double* q = peak_params + (size_t)w * j;
          double* o = mapped.data() + (size_t)w * j;
          std::copy(q, q + w, o);
          o[0] = xt[0] + q[0] * xt[1] + q[1] * xt[2];
          o[1] = yt[0] + q[0] * yt[1] + q[1] * yt[2];
        }
        rows = mapped.data();
      }
      // daoNewPeaks checks each new peak with the channel's daoCheck2D,
      // which clamps its x width to the channel's range
      h->eng->set_model(h->ch[c]->model);
      dao::handle_new_peaks(h->ch[c].get(), rows, p_type, n_peaks);
    }
    // the new groups are the last n_peaks.  The reference checks
    // [old nfit, old nfit + n): after a regrowth that dropped k ERROR peaks
    // that misses the first k new groups (and reads k past the live ones)
    (void)start;
    const int stop = h->ch[0]->pub.nfit;
    sync_errors(h, std::max(0, stop - n_peaks), stop);
  });
}

void mpIterateLM(mpFit* mp) {
  run([&] { (void)fit(M(mp), 0); });
}

int mpDoFit(mpFit* mp, int max_iterations) {
  int s = -1;
  if (dao::guarded([&] { s = fit(M(mp), max_iterations); }) != dao::kOk) return -1;
  return s;
}

void mpCleanup(mpFit* mp) {
  if (mp == nullptr) return;
  delete M(mp);
}

}  // extern "C"
