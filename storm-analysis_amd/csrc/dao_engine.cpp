// Host runtime of the MI355X fitter (see dao_engine.h).
#include "dao_engine.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace dao {

void fatal(const char* file, int line, const char* what, const char* msg) {
  char buf[512];
  const char* base = std::strrchr(file, '/');
  std::snprintf(buf, sizeof(buf), "%s:%d: %s failed: %s", base ? base + 1 : file, line, what, msg);
  throw Error(kErrHip, buf);
}

void fail(int code, const char* msg) { throw Error(code, msg); }

namespace {
thread_local int t_code = kOk;
thread_local std::string t_msg;
}  // namespace

void set_last_error(int code, const char* msg) {
  t_code = code;
  t_msg = msg ? msg : "";
}

int take_last_error(char* msg, int len) {
  const int c = t_code;
  if (msg != nullptr && len > 0) {
    std::snprintf(msg, (size_t)len, "%s", t_msg.c_str());
  }
  t_code = kOk;
  t_msg.clear();
  return c;
}

// Polls before a scheduler wait of the fitters gives up: far longer than
// any wait of a correct schedule (a wait lasts at most one LM turn of an
// earlier overlapping op, or one transition).
constexpr long long kSpinLimit = 1ll << 22;

void roi_table(int roi, std::vector<int>& rx, std::vector<int>& ry) {
  rx.clear();
  ry.clear();
  double c = 0.5 * (double)roi;
  for (int i = 0; i < roi; i++) {
    double dy = (double)i - c + 0.5;
    for (int j = 0; j < roi; j++) {
      double dx = (double)j - c + 0.5;
      if ((dx * dx + dy * dy) <= c * c) {
        rx.push_back(j);
        ry.push_back(i);
      }
    }
  }
}

template <typename T>
static T* dalloc(long long n) {
  void* p = nullptr;
  if (n <= 0) n = 1;
  DAO_CHECK(hipMalloc(&p, sizeof(T) * (size_t)n));
  return (T*)p;
}

// Frees never throw: they also run in destructors, and a failed free (a
// device in an error state) leaves nothing to recover.
template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// A failure part-way through construction releases what was allocated and
// rethrows.
Engine::Engine(const Model& mm, const std::vector<FrameSpec>& specs) : m(mm) {
  try {
    init(specs);
  } catch (...) {
    release();
    throw;
  }
}

void Engine::init(const std::vector<FrameSpec>& specs) {
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1) fail(kErrHip, "no HIP device available");
  int dev = 0;
  DAO_CHECK(hipGetDevice(&dev));
  DAO_CHECK(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
  DAO_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const char* ex = std::getenv("DAO_EXECUTOR");
  use_rounds = ex != nullptr && std::strcmp(ex, "rounds") == 0;

  roi_table(m.roi, roi_x, roi_y);
  m.nroi = (int)roi_x.size();
  if (m.roi < 1 || m.roi > 32) fail(kErrArg, "roi_size must be in [1, 32]");
  std::vector<int> packed(m.nroi);
  for (int j = 0; j < m.nroi; j++) packed[j] = roi_x[j] | (roi_y[j] << 16);
  d_roi = dalloc<int>(m.nroi);
  DAO_CHECK(hipMemcpy(d_roi, packed.data(), sizeof(int) * m.nroi, hipMemcpyHostToDevice));

  F = (int)specs.size();
  fr.resize(F);
  int cell = m.roi + 2 * m.margin;
  long long pix = 0, slots = 0;
  int cells = 0;
  for (int f = 0; f < F; f++) {
    Frame& q = fr[f];
    q.W = specs[f].W;
    q.H = specs[f].H;
    q.pix0 = pix;
    q.peak0 = (int)slots;
    q.cap = specs[f].cap;
    q.nfit = 0;
    q.cell0 = cells;
    q.ncx = (q.W + cell - 1) / cell;
    q.ncy = (q.H + cell - 1) / cell;
    pix += (long long)q.W * q.H;
    slots += q.cap;
    cells += q.ncx * q.ncy;
    pix_cap.push_back((long long)q.W * q.H);
    cell_cap.push_back(q.ncx * q.ncy);
  }
  npix = pix;
  ncells = cells;
  d_fr = dalloc<Frame>(F);
  upload_frames();

  // pixel records start zeroed; every frame gets rqe/scmos from
  // set_calibration and its fit state from new_image
  px.rw = dalloc<PixRW>(npix);
  px.ro = dalloc<PixRO>(npix);
  px.asx = dalloc<double>(npix);
  DAO_CHECK(hipMemsetAsync(px.rw, 0, sizeof(PixRW) * npix, st));
  DAO_CHECK(hipMemsetAsync(px.ro, 0, sizeof(PixRO) * npix, st));
  DAO_CHECK(hipMemsetAsync(px.asx, 0, sizeof(double) * npix, st));

  // scheduling grid (cell_start has one extra entry)
  sc.cell = cell;
  sc.cell_start = dalloc<int>(ncells + 1);
  sc.cell_fill = dalloc<int>(ncells);
  DAO_CHECK(hipMemsetAsync(sc.cell_fill, 0, sizeof(int) * ncells, st));
  sc.n_big = dalloc<int>(1);
  sc.remaining = dalloc<int>(1);
  sc.frame_bad = dalloc<int>(F);
  DAO_CHECK(hipMemsetAsync(sc.frame_bad, 0, sizeof(int) * F, st));
  d_counters = dalloc<int>((long long)F * C_COUNT);
  DAO_CHECK(hipMemsetAsync(d_counters, 0, sizeof(int) * F * C_COUNT, st));
  d_count = dalloc<int>(1);

  cap = 0;
  grow_total_cap(std::max<long long>(slots, 1));
  DAO_CHECK(hipStreamSynchronize(st));
}

Engine::~Engine() { release(); }

void Engine::release() {
  if (st) (void)hipStreamSynchronize(st);
  dfree(pk.rec);
  dfree(pk.bk);
  free_sched();
  dfree(d_fr);
  dfree(px.rw);
  dfree(px.ro);
  dfree(px.asx);
  dfree(sc.cell_start);
  dfree(sc.cell_fill);
  dfree(sc.n_big);
  dfree(sc.remaining);
  dfree(sc.frame_bad);
  dfree(d_counters);
  dfree(d_count);
  dfree(d_roi);
  dfree(d_stage);
  dfree(d_istage);
  dfree(d_sweeps);
  dfree(d_flist);
  dfree(d_err);
  dfree(d_sweep_rw);
  dfree(d_team_ctl);
  dfree(d_team_cur);
  dfree(d_team_tab);
  if (h_pin) (void)hipHostFree(h_pin);
  if (h_fr) (void)hipHostFree(h_fr);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (st) (void)hipStreamDestroy(st);
}

void Engine::set_model(const Model& mm) {
  int nroi = m.nroi;
  m = mm;
  m.nroi = nroi;
}

// Frame descriptors go through a pinned buffer and the copy is waited for,
// so the host copy can change right after and earlier kernels on the stream
// still see the old table.
void Engine::upload_frames() {
  if (h_fr == nullptr) DAO_CHECK(hipHostMalloc((void**)&h_fr, sizeof(Frame) * F, hipHostMallocDefault));
  std::memcpy(h_fr, fr.data(), sizeof(Frame) * F);
  DAO_CHECK(hipMemcpyAsync(d_fr, h_fr, sizeof(Frame) * F, hipMemcpyHostToDevice, st));
  sync();
}

void Engine::alloc_sched(long long c) {
  sc.op_peak = dalloc<int>(c);
  sc.box = dalloc<int4>(c);
  sc.pred_n = dalloc<int>(c);
  sc.pred = dalloc<int>(c * kMaxPred);
  sc.done = dalloc<int>(c);
  sc.cell_ops = dalloc<int>(c);
  sc.big_ops = dalloc<int>(c);
  sc.op_delta = dalloc<int>(c * C_COUNT);
  sc.op_flags = dalloc<int>(c);
  sc.esc_box = dalloc<int4>(c);
  d_override = dalloc<int4>(c);
  d_flag = dalloc<int>(c);
  d_run = dalloc<int>(c);
  d_journal = dalloc<double>(c * journal_words(ppl_for(m.nroi)));
  d_sweep_rec = dalloc<PeakRec>(c);
}

void Engine::free_sched() {
  dfree(sc.op_peak);
  dfree(sc.box);
  dfree(sc.pred_n);
  dfree(sc.pred);
  dfree(sc.done);
  dfree(sc.cell_ops);
  dfree(sc.big_ops);
  dfree(sc.op_delta);
  dfree(sc.op_flags);
  dfree(sc.esc_box);
  dfree(d_override);
  dfree(d_flag);
  dfree(d_run);
  dfree(d_journal);
  dfree(d_sweep_rec);
}

// Grow every per-slot array to c slots, keeping the first `cap` records.
// New slots belong to the frame whose range holds them (the single-frame
// handle owns everything).
void Engine::grow_total_cap(long long c) {
  if (c <= cap) return;
  Peaks np{};
  np.cap = c;
  np.rec = dalloc<PeakRec>(c);
  np.bk = dalloc<PeakRec>(c);
  if (cap > 0) DAO_CHECK(hipMemcpyAsync(np.rec, pk.rec, sizeof(PeakRec) * cap, hipMemcpyDeviceToDevice, st));
  DAO_CHECK(hipStreamSynchronize(st));
  dfree(pk.rec);
  dfree(pk.bk);
  free_sched();
  pk = np;
  alloc_sched(c);
  long long old = cap;
  cap = c;
  for (int f = 0; f < F; f++) {
    long long s0 = std::max<long long>(old, fr[f].peak0);
    long long s1 = std::min<long long>(c, (long long)fr[f].peak0 + fr[f].cap);
    if (F == 1) s1 = c;
    launch_set_owner(pk, s0, s1, f, st);
  }
  DAO_CHECK(hipStreamSynchronize(st));
}

void Engine::ensure_frame_cap(int f, int need) {
  if (F != 1) {
    if (need <= fr[f].cap) return;
    if (!growable) fail(kErrCapacity, "more peaks than the frame's capacity");
    grow_frames(std::max({need, 2 * fr[f].cap, 512}));
    return;
  }
  if (need <= fr[0].cap && need <= cap) return;
  long long c = std::max<long long>({(long long)need, 2 * cap, 512});
  fr[0].cap = (int)c;
  grow_total_cap(c);
  upload_frames();
}

// Every frame gets `newcap` peak slots (frame f at [f newcap, (f + 1)
// newcap)); each frame's live peaks keep their order.
void Engine::grow_frames(int newcap) {
  const long long total = (long long)F * newcap;
  Peaks np{};
  np.cap = total;
  np.rec = dalloc<PeakRec>(total);
  np.bk = dalloc<PeakRec>(total);
  for (int f = 0; f < F; f++)
    if (fr[f].nfit > 0)
      DAO_CHECK(hipMemcpyAsync(np.rec + (long long)f * newcap, pk.rec + fr[f].peak0, sizeof(PeakRec) * fr[f].nfit,
                               hipMemcpyDeviceToDevice, st));
  DAO_CHECK(hipStreamSynchronize(st));
  dfree(pk.rec);
  dfree(pk.bk);
  free_sched();
  pk = np;
  alloc_sched(total);
  cap = total;
  for (int f = 0; f < F; f++) {
    fr[f].peak0 = (int)((long long)f * newcap);
    fr[f].cap = newcap;
    // empty only the slots past the live records copied above
    launch_set_owner(pk, (long long)f * newcap + fr[f].nfit, (long long)(f + 1) * newcap, f, st);
  }
  upload_frames();
}

int* Engine::pin(size_t bytes) {
  if (bytes > h_pin_bytes) {
    if (h_pin) DAO_CHECK(hipHostFree(h_pin));
    size_t b = std::max(bytes, (size_t)4096);
    DAO_CHECK(hipHostMalloc((void**)&h_pin, b, hipHostMallocDefault));
    h_pin_bytes = b;
  }
  return h_pin;
}

double* Engine::stage(size_t bytes) {
  if (bytes > stage_bytes) {
    dfree(d_stage);
    d_stage = dalloc<double>((long long)((bytes + 7) / 8));
    stage_bytes = bytes;
  }
  return d_stage;
}

int* Engine::istage(size_t bytes) {
  if (bytes > istage_bytes) {
    dfree(d_istage);
    d_istage = dalloc<int>((long long)((bytes + 3) / 4));
    istage_bytes = bytes;
  }
  return d_istage;
}

void Engine::sync() { DAO_CHECK(hipStreamSynchronize(st)); }

// ---------------------------------------------------------------------------
// image state
// ---------------------------------------------------------------------------

void Engine::set_calibration(int f, const double* rqe_h, const double* scmos_h) {
  long long n = (long long)fr[f].W * fr[f].H;
  double* d = stage(sizeof(double) * 2 * n);
  DAO_CHECK(hipMemcpyAsync(d, rqe_h, sizeof(double) * n, hipMemcpyHostToDevice, st));
  DAO_CHECK(hipMemcpyAsync(d + n, scmos_h, sizeof(double) * n, hipMemcpyHostToDevice, st));
  launch_set_calibration(px, fr[f].pix0, n, d, d + n, st);
  sync();
}

void Engine::new_image(int f, const double* img) {
  long long n = (long long)fr[f].W * fr[f].H;
  double* d = stage(sizeof(double) * n);
  DAO_CHECK(hipMemcpyAsync(d, img, sizeof(double) * n, hipMemcpyHostToDevice, st));
  new_image_dev(f, d);
  sync();
}

void Engine::new_image_dev(int f, const double* d) {
  long long n = (long long)fr[f].W * fr[f].H;
  launch_new_image(px, fr[f].pix0, n, d, st);
  fr[f].nfit = 0;
  upload_frames();
}

void Engine::new_background(int f, const double* bg) {
  long long n = (long long)fr[f].W * fr[f].H;
  double* d = stage(sizeof(double) * n);
  DAO_CHECK(hipMemcpyAsync(d, bg, sizeof(double) * n, hipMemcpyHostToDevice, st));
  new_background_dev(f, d);
  sync();
}

void Engine::new_background_dev(int f, const double* d) {
  long long n = (long long)fr[f].W * fr[f].H;
  launch_new_background(px, fr[f].pix0, n, d, st);
}

void Engine::anscombe(int f) {
  long long n = (long long)fr[f].W * fr[f].H;
  launch_anscombe(px, fr[f].pix0, n, st);
  sync();
}

// ---------------------------------------------------------------------------
// the ordered executor
// ---------------------------------------------------------------------------

OpArgs Engine::args(int round) {
  OpArgs a;
  a.m = m;
  a.px = px;
  a.pk = pk;
  a.sc = sc;
  a.frames = d_fr;
  a.roi_tab = d_roi;
  a.round = round;
  return a;
}

// Run the ops sc.op_peak[0..nops) (boxes already built) to completion.
void Engine::run_ops(int kind, const double* rows, int fmt, int first_slot) {
  if (sc.nops == 0) return;
  launch_fill_i32(sc.n_big, 1, 0, st);
  launch_binning(sc, d_fr, pk, ncells, st);
  launch_fill_i32(sc.remaining, 1, sc.nops, st);
  int* h = pin(sizeof(int));
  int round = 0;
  int group = 4;
  int last = sc.nops;
  bool timed = timing && kind == OP_LM;
  if (timed && ev0 == nullptr) {
    DAO_CHECK(hipEventCreate(&ev0));
    DAO_CHECK(hipEventCreate(&ev1));
  }
  for (;;) {
    if (timed) DAO_CHECK(hipEventRecord(ev0, st));
    for (int r = 0; r < group; r++) launch_op_round(kind, args(round++), st, rows, fmt, first_slot);
    if (timed) DAO_CHECK(hipEventRecord(ev1, st));
    DAO_CHECK(hipMemcpyAsync(h, sc.remaining, sizeof(int), hipMemcpyDeviceToHost, st));
    sync();
    if (timed) {
      float ms = 0.0f;
      DAO_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
      lm_ms += ms;
      lm_launches += group;
    }
    int rem = h[0];
    if (rem == 0) break;
    if (rem >= last) fail(kErrSched, "the ordered executor made no progress");
    last = rem;
    group = std::min(group * 2, 32);
  }
  rounds_total += round;
}

int Engine::select(int mode) {
  launch_select(pk, d_fr, cap, mode, sc.frame_bad, d_flag, sc.op_peak, d_count, st);
  int* h = pin(sizeof(int));
  DAO_CHECK(hipMemcpyAsync(h, d_count, sizeof(int), hipMemcpyDeviceToHost, st));
  sync();
  return h[0];
}

int Engine::running_count() {
  launch_fill_i32(sc.frame_bad, F, 0, st);
  return select(SEL_RUNNING);
}

// Workgroups of the frame-per-team fitter: frame f gets
// min(kMaxTeam, ceil(nfit / m.team)) of them, the largest teams first (they
// are dispatched first); kTeamSeats seats per workgroup.  tab = block frames,
// block first seats, then the F + 1 seat offsets.  Returns the block count,
// or 0 when no frame gets more than one workgroup (k_fit_frames then).
int Engine::team_blocks(std::vector<int>& tab, int& nseats) const {
  if (m.team <= 0) return 0;
  std::vector<int> order, team(F, 0), s0(F + 1, 0);
  bool multi = false;
  int seats = 0;
  for (int f = 0; f < F; f++) {
    s0[f] = seats;
    if (fr[f].nfit <= 0) continue;
    team[f] = std::max(1, std::min(kMaxTeam, (fr[f].nfit + m.team - 1) / m.team));
    multi = multi || team[f] > 1;
    seats += team[f] * kTeamSeats;
    order.push_back(f);
  }
  s0[F] = seats;
  if (!multi) return 0;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return team[a] > team[b]; });
  std::vector<int> bf, bs;
  for (int f : order)
    for (int k = 0; k < team[f]; k++) {
      bf.push_back(f);
      bs.push_back(s0[f] + k * kTeamSeats);
    }
  tab = bf;
  tab.insert(tab.end(), bs.begin(), bs.end());
  tab.insert(tab.end(), s0.begin(), s0.end());
  nseats = seats;
  return (int)bf.size();
}

FitArgs Engine::fitter_args(int max_more) {
  if (d_sweeps == nullptr) {
    d_sweeps = dalloc<int>(F);
    d_flist = dalloc<int>(F);
    d_err = dalloc<int>(F);
  }
  FitArgs a{};
  a.m = m;
  a.px = px;
  a.pk = pk;
  a.frames = d_fr;
  a.roi_tab = d_roi;
  a.op_peak = sc.op_peak;
  a.run_list = d_run;
  a.journal = d_journal;
  a.box = sc.box;
  a.op_delta = sc.op_delta;
  a.op_done = sc.op_flags;
  a.frame_counters = d_counters;
  a.frame_sweeps = d_sweeps;
  a.frame_list = nullptr;
  a.max_more = max_more;
  a.frame_err = d_err;
  a.spin_limit = spin_limit > 0 ? spin_limit : kSpinLimit;
  return a;
}

int Engine::fit_frames(int max_more) {
  std::vector<int> list;
  for (int f = 0; f < F; f++)
    if (fr[f].nfit > 0 && (fit_mask.empty() || fit_mask[f])) list.push_back(f);
  if (list.empty()) return 0;
  if (d_sweeps == nullptr) {
    d_sweeps = dalloc<int>(F);
    d_flist = dalloc<int>(F);
    d_err = dalloc<int>(F);
  }
  launch_fill_i32(d_err, F, 0, st);
  failed.assign(F, 0);
  failed_count = 0;
  std::vector<int> tab;
  int nseats = 0;
  const int nblocks = team_blocks(tab, nseats);
  int* h = pin(sizeof(int) * (2 * (size_t)F + tab.size()));
  std::memcpy(h, list.data(), sizeof(int) * list.size());
  DAO_CHECK(hipMemcpyAsync(d_flist, h, sizeof(int) * list.size(), hipMemcpyHostToDevice, st));
  launch_fill_i32(d_sweeps, F, 0, st);
  if (nblocks > 0) {
    if (d_team_ctl == nullptr) d_team_ctl = dalloc<TeamCtl>(F);
    if (nseats > team_seat_cap) {
      dfree(d_team_cur);
      team_seat_cap = std::max(nseats, 2 * team_seat_cap);
      d_team_cur = dalloc<int>(team_seat_cap);
    }
    if ((int)tab.size() > team_tab_cap) {
      dfree(d_team_tab);
      team_tab_cap = std::max((int)tab.size(), 2 * team_tab_cap);
      d_team_tab = dalloc<int>(team_tab_cap);
    }
    int* ht = h + 2 * F;
    std::memcpy(ht, tab.data(), sizeof(int) * tab.size());
    DAO_CHECK(hipMemcpyAsync(d_team_tab, ht, sizeof(int) * tab.size(), hipMemcpyHostToDevice, st));
    launch_team_init(d_team_ctl, F, d_team_cur, nseats, st);
  }
  FitArgs a;
  a.m = m;
  // More frames than CUs: four waves per workgroup, so that two frames share
  // a CU and the batch runs in one round (eight waves per frame otherwise).
  if (a.m.waves == 0 && nblocks == 0 && (int)list.size() > num_cus) a.m.waves = 4;
  a.px = px;
  a.pk = pk;
  a.frames = d_fr;
  a.roi_tab = d_roi;
  a.op_peak = sc.op_peak;
  a.run_list = d_run;
  a.journal = d_journal;
  a.box = sc.box;
  a.op_delta = sc.op_delta;
  a.op_done = sc.op_flags;
  a.frame_counters = d_counters;
  a.frame_sweeps = d_sweeps;
  a.frame_list = d_flist;
  a.max_more = max_more;
  a.frame_err = d_err;
  a.spin_limit = spin_limit > 0 ? spin_limit : kSpinLimit;
  a.team_ctl = d_team_ctl;
  a.team_cur = d_team_cur;
  a.team_block_frame = nblocks > 0 ? d_team_tab : nullptr;
  a.team_block_seat = nblocks > 0 ? d_team_tab + nblocks : nullptr;
  a.team_seat0 = nblocks > 0 ? d_team_tab + 2 * nblocks : nullptr;
  if (timing && ev0 == nullptr) {
    DAO_CHECK(hipEventCreate(&ev0));
    DAO_CHECK(hipEventCreate(&ev1));
  }
  if (timing) DAO_CHECK(hipEventRecord(ev0, st));
  if (nblocks > 0) {
    // A team hands a frame's tail (an LM run of at most team_solo ops) to
    // one workgroup: the resume launch, queued behind the team's, finishes
    // those frames and leaves at once for the others.
    a.team_solo = team_solo;
    launch_fit_team(a, nblocks, st);
    if (team_solo > 0) {
      DAO_CHECK(hipGetLastError());
      FitArgs b = a;
      b.team_resume = 1;
      launch_fit_frames(b, (int)list.size(), st);
    }
  } else {
    launch_fit_frames(a, (int)list.size(), st);
  }
  DAO_CHECK(hipGetLastError());
  if (timing) DAO_CHECK(hipEventRecord(ev1, st));
  DAO_CHECK(hipMemcpyAsync(h, d_sweeps, sizeof(int) * F, hipMemcpyDeviceToHost, st));
  DAO_CHECK(hipMemcpyAsync(h + F, d_err, sizeof(int) * F, hipMemcpyDeviceToHost, st));
  sync();
  int most = 0;
  for (int f = 0; f < F; f++) most = std::max(most, h[f]);
  // A frame whose wait gave up stopped part-way: its fit state is not the
  // sequential one, so it is emptied and reported; the others stand.
  bool any = false;
  for (int f = 0; f < F; f++) {
    if (h[F + f] == 0) continue;
    failed[f] = 1;
    failed_count++;
    fr[f].nfit = 0;
    any = true;
  }
  if (any) upload_frames();
  sweeps_total += most;
  if (timing) {
    float ms = 0.0f;
    DAO_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    lm_ms += ms;
    lm_launches += 1;
  }
  return most;
}

int Engine::fit(int max_iterations) {
  if (!use_rounds) return fit_frames(max_iterations);
  if (!fit_mask.empty()) fail(kErrArg, "a frame mask needs the persistent fitter (not the rounds executor)");
  int sweeps = 1;
  sweep();
  int i = 0;
  while (running_count() > 0 && i < max_iterations) {
    sweep();
    sweeps++;
    i++;
  }
  return sweeps;
}

// mFitIterateLM (multi_fit.c:1072-1349) for every frame of the batch.
int Engine::sweep() {
  launch_fill_i32((int*)d_override, cap * 4, 0, st);
  launch_fill_i32(sc.frame_bad, F, 0, st);
  sc.nops = select(SEL_RUNNING);
  const int started = sc.nops;
  if (started == 0) return 0;
  sweeps_total++;
  // A frame whose turn leaves its box returns to its state at the sweep's
  // start (exact) and reruns the sweep with that turn alone in the frame.
  if (d_sweep_rw == nullptr) d_sweep_rw = dalloc<PixRW>(npix);
  DAO_CHECK(hipMemcpyAsync(d_sweep_rw, px.rw, sizeof(PixRW) * (size_t)npix, hipMemcpyDeviceToDevice, st));
  DAO_CHECK(hipMemcpyAsync(d_sweep_rec, pk.rec, sizeof(PeakRec) * (size_t)cap, hipMemcpyDeviceToDevice, st));
  int attempt = 0;
  while (sc.nops > 0) {
    launch_box(OP_LM, sc, pk, d_fr, m, d_override, nullptr, 0, 0, st);
    launch_fill_i32(sc.frame_bad, F, 0, st);
    run_ops(OP_LM, nullptr, 0, 0);
    launch_commit(sc, pk, d_counters, st);
    int* h = pin(sizeof(int) * F);
    DAO_CHECK(hipMemcpyAsync(h, sc.frame_bad, sizeof(int) * F, hipMemcpyDeviceToHost, st));
    sync();
    int bad = 0;
    for (int f = 0; f < F; f++) bad += h[f] ? 1 : 0;
    if (bad == 0) break;
    escapes_total += bad;
    launch_restore_bad_frames(d_fr, F, sc.frame_bad, px.rw, d_sweep_rw, pk.rec, d_sweep_rec, st);
    launch_collect_escapes(sc, pk, d_override, st);
    attempt++;
    max_attempts_seen = std::max(max_attempts_seen, attempt);
    if (attempt > 64) fail(kErrSched, "a sweep did not settle after 64 attempts");
    sc.nops = select(SEL_RUNNING_BAD);
  }
  // mFitRecenterPeaks (multi_fit.c:1644-1690)
  launch_fill_i32(sc.frame_bad, F, 0, st);
  sc.nops = select(SEL_RECENTER);
  if (sc.nops > 0) {
    launch_box(OP_RECENTER, sc, pk, d_fr, m, nullptr, nullptr, 0, 0, st);
    run_ops(OP_RECENTER, nullptr, 0, 0);
    launch_commit(sc, pk, d_counters, st);
  }
  return started;
}

// daoNewPeaks (dao_fit.c:1582-1766) for one frame; storage policy is the
// caller's (the ABI mirrors mFitNewPeaks, multi_fit.c:1419-1499).
void Engine::new_peaks(int f, const double* rows, int fmt, int n) {
  if (n <= 0) return;
  int width = fmt == PF_TEXT ? 7 : 4;
  ensure_frame_cap(f, fr[f].nfit + n);
  double* d_rows = stage(sizeof(double) * (size_t)n * width);
  DAO_CHECK(hipMemcpyAsync(d_rows, rows, sizeof(double) * (size_t)n * width, hipMemcpyHostToDevice, st));
  add_peaks({{d_rows, f, n, fr[f].peak0 + fr[f].nfit, 0}}, fmt);
}

// Batch form: new peaks for several frames in one launch.  rows_dev[k] are
// device pointers; all frames use the same format.
void Engine::new_peaks_multi(const std::vector<int>& frames, const std::vector<const double*>& rows_dev,
                             const std::vector<int>& counts, int fmt) {
  std::vector<AddPeaksJob> jobs;
  for (size_t k = 0; k < frames.size(); k++) {
    const int f = frames[k], n = counts[k];
    if (n == 0) continue;
    if (fr[f].nfit + n > fr[f].cap) fail(kErrCapacity, "more peaks than the frame's capacity");
    jobs.push_back({rows_dev[k], f, n, fr[f].peak0 + fr[f].nfit, 0});
  }
  add_peaks(jobs, fmt);
}

// daoNewPeaks' image work for each job (at most one per frame): k_add_peaks,
// one workgroup per frame, the peaks added in index order where their ROIs
// overlap and then their first errors, as the reference does one frame at a
// time.  The frames' peak counts grow by the jobs' row counts.
void Engine::add_peaks(const std::vector<AddPeaksJob>& jobs, int fmt) {
  if (jobs.empty()) return;
  FitArgs a = fitter_args(0);
  // the caller's wait limit (daoBatchSetWaitLimit) is the fit's; adding
  // peaks keeps the default bound
  a.spin_limit = kSpinLimit;
  const size_t bytes = sizeof(AddPeaksJob) * jobs.size();
  AddPeaksJob* h = reinterpret_cast<AddPeaksJob*>(pin(bytes + sizeof(int) * (size_t)F));
  std::memcpy(h, jobs.data(), bytes);
  AddPeaksJob* d = reinterpret_cast<AddPeaksJob*>(istage(bytes));
  DAO_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
  launch_fill_i32(d_err, F, 0, st);
  // more frames than CUs: four waves per workgroup, two frames per CU
  launch_add_peaks(a, d, (int)jobs.size(), fmt, (int)jobs.size() > num_cus ? 4 : 0, st);
  DAO_CHECK(hipGetLastError());
  int* herr = reinterpret_cast<int*>(reinterpret_cast<char*>(h) + bytes);
  DAO_CHECK(hipMemcpyAsync(herr, d_err, sizeof(int) * F, hipMemcpyDeviceToHost, st));
  sync();
  for (const AddPeaksJob& j : jobs)
    if (herr[j.frame] != 0) fail(kErrSched, "adding peaks: a scheduler wait gave up");
  for (const AddPeaksJob& j : jobs) fr[j.frame].nfit += j.n;
  upload_frames();
}

// Stable compaction of frame f's live peaks to the slots listed in keep.
void Engine::compact(int f, const std::vector<int32_t>& keep) {
  int n = (int)keep.size();
  if (n == fr[f].nfit) {
    bool same = true;
    for (int i = 0; i < n && same; i++) same = keep[i] == i;
    if (same) return;
  }
  if (n > 0) {
    int* d_map = istage(sizeof(int) * n);
    DAO_CHECK(hipMemcpyAsync(d_map, keep.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    // gather into the backup array, then copy the range back
    launch_gather_peaks(pk.bk, pk.rec, d_map, fr[f].peak0, n, st);
    DAO_CHECK(hipMemcpyAsync(pk.rec + fr[f].peak0, pk.bk + fr[f].peak0, sizeof(PeakRec) * n,
                             hipMemcpyDeviceToDevice, st));
  }
  fr[f].nfit = n;
  upload_frames();
  sync();
}

// The copy is waited for: the pinned buffer is reused by the next read-back.
int* Engine::upload_offsets(std::vector<int>& off) {
  off.assign(F + 1, 0);
  for (int f = 0; f < F; f++) off[f + 1] = off[f] + fr[f].nfit;
  int* d = istage(sizeof(int) * (size_t)(F + 1));
  int* h = pin(sizeof(int) * (size_t)(F + 1));
  std::memcpy(h, off.data(), sizeof(int) * (size_t)(F + 1));
  DAO_CHECK(hipMemcpyAsync(d, h, sizeof(int) * (size_t)(F + 1), hipMemcpyHostToDevice, st));
  sync();
  return d;
}

void Engine::set_status_all(const int32_t* status_dev) {
  if (total_nfit() == 0) return;
  std::vector<int> off;
  int* d_off = upload_offsets(off);
  // peaks leaving the fit image: subtracted by the executor, in slot order
  launch_status_flag(pk, d_fr, d_off, cap, status_dev, d_flag, st);
  launch_compact_list(d_flag, cap, sc.op_peak, d_count, st);
  int* h = pin(sizeof(int));
  DAO_CHECK(hipMemcpyAsync(h, d_count, sizeof(int), hipMemcpyDeviceToHost, st));
  sync();
  const int nops = h[0];
  if (nops > 0) {
    sc.nops = nops;
    launch_fill_i32(sc.frame_bad, F, 0, st);
    launch_box(OP_SUBTRACT, sc, pk, d_fr, m, nullptr, nullptr, 0, 0, st);
    run_ops(OP_SUBTRACT, nullptr, 0, 0);
  }
  // d_off may have been reused by the executor: upload again
  d_off = upload_offsets(off);
  launch_apply_status(pk, d_fr, d_off, cap, status_dev, st);
  sync();
}

void Engine::remove_all(int drop) {
  if (total_nfit() == 0) return;
  int* d_counts = istage(sizeof(int) * (size_t)F);
  launch_compact_frames(pk, d_fr, F, drop, d_counts, d_counters, st);
  int* h = pin(sizeof(int) * (size_t)F);
  DAO_CHECK(hipMemcpyAsync(h, d_counts, sizeof(int) * (size_t)F, hipMemcpyDeviceToHost, st));
  sync();
  for (int f = 0; f < F; f++) fr[f].nfit = h[f];
  upload_frames();
  sync();
}

void Engine::significance_all(double* out_dev) {
  const long long total = total_nfit();
  if (total == 0) return;
  std::vector<int> off;
  int* d_off = upload_offsets(off);
  int* d_slots = sc.op_peak;  // a slot list of `cap` entries
  launch_packed_slots(d_fr, F, d_off, d_slots, st);
  launch_peak_sums(m, px, pk, d_fr, d_roi, 0, (int)total, 3, out_dev, st, d_slots);
  sync();
}

// mFitSetPeakStatus (multi_fit.c:1809-1831).
void Engine::set_status(int f, const int32_t* status) {
  int n = fr[f].nfit;
  if (n == 0) return;
  std::vector<PeakRec> rec;
  get_records(f, rec);
  std::vector<int> ops;
  for (int i = 0; i < n; i++)
    if (status[i] == ST_ERROR && rec[i].added > 0) ops.push_back(fr[f].peak0 + i);
  if (!ops.empty()) {
    int* d = istage(sizeof(int) * ops.size());
    DAO_CHECK(hipMemcpyAsync(d, ops.data(), sizeof(int) * ops.size(), hipMemcpyHostToDevice, st));
    DAO_CHECK(hipMemcpyAsync(sc.op_peak, d, sizeof(int) * ops.size(), hipMemcpyDeviceToDevice, st));
    sc.nops = (int)ops.size();
    launch_fill_i32(sc.frame_bad, F, 0, st);
    launch_box(OP_SUBTRACT, sc, pk, d_fr, m, nullptr, nullptr, 0, 0, st);
    run_ops(OP_SUBTRACT, nullptr, 0, 0);
  }
  int* d = istage(sizeof(int) * n);
  DAO_CHECK(hipMemcpyAsync(d, status, sizeof(int) * n, hipMemcpyHostToDevice, st));
  launch_set_status(pk, fr[f].peak0, n, d, st);
  sync();
}

// ---------------------------------------------------------------------------
// read-back
// ---------------------------------------------------------------------------

void Engine::read_counters(std::vector<int>& out, int f) {
  int* h = pin(sizeof(int) * C_COUNT);
  DAO_CHECK(hipMemcpyAsync(h, d_counters + (size_t)f * C_COUNT, sizeof(int) * C_COUNT, hipMemcpyDeviceToHost, st));
  launch_fill_i32(d_counters + (size_t)f * C_COUNT, C_COUNT, 0, st);
  sync();
  out.assign(h, h + C_COUNT);
}

void Engine::read_counters(std::vector<int>& out) {
  size_t n = (size_t)F * C_COUNT;
  int* h = pin(sizeof(int) * n);
  DAO_CHECK(hipMemcpyAsync(h, d_counters, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  launch_fill_i32(d_counters, (long long)n, 0, st);
  sync();
  out.assign(h, h + n);
}

void Engine::get_records(int f, std::vector<PeakRec>& out) {
  int n = fr[f].nfit;
  out.resize(n);
  if (n == 0) return;
  DAO_CHECK(hipMemcpyAsync(out.data(), pk.rec + fr[f].peak0, sizeof(PeakRec) * n, hipMemcpyDeviceToHost, st));
  sync();
}

void Engine::get_status(int f, std::vector<int32_t>& out) {
  std::vector<PeakRec> r;
  get_records(f, r);
  out.resize(r.size());
  for (size_t i = 0; i < r.size(); i++) out[i] = r[i].status;
}

void Engine::get_iterations(int f, std::vector<int32_t>& out) {
  std::vector<PeakRec> r;
  get_records(f, r);
  out.resize(r.size());
  for (size_t i = 0; i < r.size(); i++) out[i] = r[i].iter;
}

void Engine::get_error(int f, std::vector<double>& out) {
  std::vector<PeakRec> r;
  get_records(f, r);
  out.resize(r.size());
  for (size_t i = 0; i < r.size(); i++) out[i] = r[i].err;
}

void Engine::get_param(int f, int q, std::vector<double>& out) {
  std::vector<PeakRec> r;
  get_records(f, r);
  out.resize(r.size());
  for (size_t i = 0; i < r.size(); i++) out[i] = r[i].prm[q];
}

void Engine::get_sums(int f, int what, std::vector<double>& out) {
  int n = fr[f].nfit;
  out.resize(n);
  if (n == 0) return;
  double* d = stage(sizeof(double) * n);
  launch_peak_sums(m, px, pk, d_fr, d_roi, fr[f].peak0, n, what, d, st);
  DAO_CHECK(hipMemcpyAsync(out.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  sync();
}

void Engine::get_jacobian(int f, std::vector<double>& out) {
  int n = fr[f].nfit;
  out.resize((size_t)n * m.jac);
  if (n == 0) return;
  double* d = stage(sizeof(double) * n * m.jac);
  launch_peak_jacobian(m, px, pk, d_fr, d_roi, fr[f].peak0, n, d, st);
  DAO_CHECK(hipMemcpyAsync(out.data(), d, sizeof(double) * n * m.jac, hipMemcpyDeviceToHost, st));
  sync();
}

void Engine::get_image(int f, double* out, int residual) {
  long long n = (long long)fr[f].W * fr[f].H;
  double* d = stage(sizeof(double) * n);
  launch_fit_image(px, fr[f].pix0, n, d, residual, st);
  DAO_CHECK(hipMemcpyAsync(out, d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  sync();
}

int Engine::count_status(int f, int status) {
  std::vector<int32_t> s;
  get_status(f, s);
  int c = 0;
  for (int v : s) c += (v == status) ? 1 : 0;
  return c;
}

long long Engine::total_nfit() const {
  long long n = 0;
  for (int f = 0; f < F; f++) n += fr[f].nfit;
  return n;
}

void Engine::pack_all(double* out, int32_t* status, bool on_device) {
  long long total = total_nfit();
  if (total == 0) return;
  std::vector<int> off(F);
  long long at = 0;
  for (int f = 0; f < F; f++) {
    off[f] = (int)at;
    at += fr[f].nfit;
  }
  size_t rows_bytes = sizeof(double) * 9 * (size_t)total;
  size_t stat_bytes = sizeof(int) * (size_t)total;
  double* d = stage(rows_bytes + stat_bytes + sizeof(int) * F + 64);
  int* d_status = (int*)((char*)d + rows_bytes);
  int* d_off = d_status + total;
  int* h_off = pin(sizeof(int) * F);
  std::memcpy(h_off, off.data(), sizeof(int) * F);
  DAO_CHECK(hipMemcpyAsync(d_off, h_off, sizeof(int) * F, hipMemcpyHostToDevice, st));
  if (on_device) {
    // straight into the caller's device arrays
    launch_pack_peaks(pk, d_fr, d_off, m, cap, out, status, st);
  } else {
    launch_pack_peaks(pk, d_fr, d_off, m, cap, d, d_status, st);
    DAO_CHECK(hipMemcpyAsync(out, d, rows_bytes, hipMemcpyDeviceToHost, st));
    DAO_CHECK(hipMemcpyAsync(status, d_status, stat_bytes, hipMemcpyDeviceToHost, st));
  }
  sync();
}

// Fit images (f / rqe, or the residual x - f) of every frame, concatenated
// in frame order at the frames' current sizes.
void Engine::get_images_all(double* out, int residual, bool on_device) {
  const long long n = frame_pixels_total();
  if (n == 0) return;
  double* d = on_device ? out : stage(sizeof(double) * n);
  if (packed()) {
    launch_fit_image(px, 0, npix, d, residual, st);
  } else {
    long long at = 0;
    for (int f = 0; f < F; f++) {
      const long long nf = (long long)fr[f].W * fr[f].H;
      launch_fit_image(px, fr[f].pix0, nf, d + at, residual, st);
      at += nf;
    }
  }
  if (!on_device) DAO_CHECK(hipMemcpyAsync(out, d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  sync();
}

bool Engine::packed() const {
  for (int f = 0; f < F; f++)
    if ((long long)fr[f].W * fr[f].H != pix_cap[f]) return false;
  return true;
}

// Frames' data concatenated in frame order at their current sizes.
void Engine::new_images_all_dev(const double* img) {
  if (packed()) {
    launch_new_image(px, 0, npix, img, st);
  } else {
    long long at = 0;
    for (int f = 0; f < F; f++) {
      long long n = (long long)fr[f].W * fr[f].H;
      launch_new_image(px, fr[f].pix0, n, img + at, st);
      at += n;
    }
  }
  for (int f = 0; f < F; f++) fr[f].nfit = 0;
  upload_frames();
}

void Engine::new_frames_all_dev(const double* img, const double* bg) {
  if (packed()) {
    launch_new_frame_pixels(px, 0, npix, img, bg, st);
  } else {
    long long at = 0;
    for (int f = 0; f < F; f++) {
      long long n = (long long)fr[f].W * fr[f].H;
      launch_new_frame_pixels(px, fr[f].pix0, n, img + at, bg + at, st);
      at += n;
    }
  }
  for (int f = 0; f < F; f++) fr[f].nfit = 0;
  upload_frames();
}

void Engine::new_backgrounds_all_dev(const double* bg) {
  if (packed()) {
    launch_new_background(px, 0, npix, bg, st);
    return;
  }
  long long at = 0;
  for (int f = 0; f < F; f++) {
    long long n = (long long)fr[f].W * fr[f].H;
    launch_new_background(px, fr[f].pix0, n, bg + at, st);
    at += n;
  }
}

long long Engine::frame_pixels_total() const {
  long long n = 0;
  for (int f = 0; f < F; f++) n += (long long)fr[f].W * fr[f].H;
  return n;
}

// A new size for slot f within the pixel and cell capacity it was created
// with; the slot is emptied.
bool Engine::set_frame_shape(int f, int W, int H, bool upload) {
  const int cell = sc.cell;
  const int ncx = (W + cell - 1) / cell, ncy = (H + cell - 1) / cell;
  if (W < 1 || H < 1 || (long long)W * H > pix_cap[f] || ncx * ncy > cell_cap[f]) return false;
  fr[f].W = W;
  fr[f].H = H;
  fr[f].ncx = ncx;
  fr[f].ncy = ncy;
  fr[f].nfit = 0;
  if (upload) upload_frames();
  return true;
}

void Engine::clear_frame(int f) {
  fr[f].nfit = 0;
  upload_frames();
}

void Engine::set_calibration_const(int f, double rqe, double scmos) {
  launch_set_calibration_const(px, fr[f].pix0, (long long)fr[f].W * fr[f].H, rqe, scmos, st);
}

}  // namespace dao
