// Resident fitter entry points of libdao_fit.so (include/dao_resident_amd.h).
//
// Requests travel through pinned, device-mapped host memory (dao_resident.h):
// the host fills a slot's descriptor, writes the slot into the ticket ring
// and raises the ring's tail with a release store; the kernel's workgroups
// take tickets with a device counter and publish each result with a release
// store of the slot's `done`.  The host starts the kernel when a request
// arrives and none is running, and again whenever it finds the kernel gone
// (its workgroups leave when idle) while requests are waiting.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "dao_batch_handle.h"
#include "dao_resident.h"
#include "../../include/dao_resident_amd.h"

using dao::Engine;

namespace {

enum SlotState { kFree = 0, kQueued = 1, kFinished = 2 };
constexpr int kResidentSpareCUs = 8;

template <class T>
T* pinned(size_t n) {
  void* p = nullptr;
  DAO_CHECK(hipHostMalloc(&p, sizeof(T) * (n > 0 ? n : 1), hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(p, 0, sizeof(T) * (n > 0 ? n : 1));
  return static_cast<T*>(p);
}

template <class T>
T* device_view(T* host) {
  void* d = nullptr;
  DAO_CHECK(hipHostGetDevicePointer(&d, host, 0));
  return static_cast<T*>(d);
}

void unpin(void* p) {
  if (p) (void)hipHostFree(p);
}

}  // namespace

struct daoResident {
  daoBatch* b = nullptr;
  dao::ResQueue* q = nullptr;
  int* ring = nullptr;
  int qcap = 0;
  dao::ResSlot* slots = nullptr;
  std::vector<double*> in_buf;  // per slot (host requests): image | background | rows
  std::vector<double*> out_rows;
  std::vector<int*> out_status;
  std::vector<int> state;
  std::vector<long long> seq;
  std::vector<int> team_of;  // workgroups the slot's request was given
  int members = 0;           // workgroups given to requests not yet collected
  long long* d_head = nullptr;  // [0]: tickets taken, [1]: clock of the last take, [2]: frames in flight
  double* d_rows = nullptr;
  long long* d_pix_cap = nullptr;
  // team requests: per-slot team state, scheduling control and seats
  int team_peaks = 0;  // > 0: a frame of n peaks is fitted by ceil(n / team_peaks) workgroups
  dao::ResTeamSlot* d_team = nullptr;
  dao::TeamCtl* d_team_ctl = nullptr;
  int* d_team_cur = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  bool running = false;
  long long tail = 0;      // ring entries published
  long long next_seq = 0;  // requests published
  long long launches = 0;
  int blocks = 0;
  int max_iterations = 2000;
  long long idle_ticks = 0;
  double tick_hz = 1.0e8;
  std::mutex mu;

  Engine& eng() { return *b->eng; }

  // (Re)start the kernel if it is not running; call with mu held.
  void ensure_running() {
    if (running) {
      const hipError_t e = hipEventQuery(ev);
      if (e == hipErrorNotReady) return;
      if (e != hipSuccess) dao::fatal(__FILE__, __LINE__, "hipEventQuery", hipGetErrorString(e));
      running = false;
    }
    Engine& e = eng();
    dao::FitArgs a = e.fitter_args(max_iterations);
    a.team_ctl = d_team_ctl;
    a.team_cur = d_team_cur;
    dao::ResArgs ra;
    ra.q = device_view(q);
    ra.ring = device_view(ring);
    ra.slots = device_view(slots);
    ra.head = d_head;
    ra.last = d_head + 1;
    ra.busy = d_head + 2;
    ra.team = d_team;
    ra.rows_dev = d_rows;
    ra.pix_cap = d_pix_cap;
    ra.nslots = (int)state.size();
    ra.qcap = qcap;
    ra.idle_ticks = idle_ticks;
    const bool mixed = a.m.prec == dao::PREC_MIXED;
    if (team_peaks > 0) {
      if (mixed) dao::launch_resident_team_mixed(a, ra, blocks, st);
      else dao::launch_resident_team_fp64(a, ra, blocks, st);
    } else {
      if (mixed) dao::launch_resident_mixed(a, ra, blocks, st);
      else dao::launch_resident_fp64(a, ra, blocks, st);
    }
    DAO_CHECK(hipGetLastError());
    DAO_CHECK(hipEventRecord(ev, st));
    running = true;
    launches++;
  }

  void release() {
    if (st) {
      if (q) __atomic_store_n(&q->stop, 1LL, __ATOMIC_RELEASE);
      (void)hipStreamSynchronize(st);
    }
    for (double* p : in_buf) unpin(p);
    for (double* p : out_rows) unpin(p);
    for (int* p : out_status) unpin(p);
    unpin(q);
    unpin(ring);
    unpin(slots);
    if (d_head) (void)hipFree(d_head);
    if (d_rows) (void)hipFree(d_rows);
    if (d_pix_cap) (void)hipFree(d_pix_cap);
    if (d_team) (void)hipFree(d_team);
    if (d_team_ctl) (void)hipFree(d_team_ctl);
    if (d_team_cur) (void)hipFree(d_team_cur);
    if (ev) (void)hipEventDestroy(ev);
    if (st) (void)hipStreamDestroy(st);
  }
};

extern "C" {

daoResident* daoResidentCreate(daoBatch* b, int workgroups, int max_iterations, double idle_seconds, int team_peaks) {
  daoResident* r = nullptr;
  const int rc = dao::guarded([&] {
    if (b == nullptr || max_iterations < 0 || !(idle_seconds > 0.0) || team_peaks < 0)
      dao::fail(dao::kErrArg, "bad resident fitter arguments");
    Engine& e = *b->eng;
    if (e.m.errm != dao::EM_MLE) dao::fail(dao::kErrArg, "the resident fitter supports the MLE error model only");
    if (e.F >= (1 << dao::kResMemberShift)) dao::fail(dao::kErrArg, "too many slots for the ticket encoding");
    r = new daoResident();
    try {
      r->b = b;
      const int F = e.F;
      r->qcap = F * dao::kMaxTeam;  // every slot's request, each with up to kMaxTeam tickets
      r->team_peaks = team_peaks;
      r->q = pinned<dao::ResQueue>(1);
      r->ring = pinned<int>(r->qcap);
      r->slots = pinned<dao::ResSlot>(F);
      r->in_buf.assign(F, nullptr);
      r->out_rows.assign(F, nullptr);
      r->out_status.assign(F, nullptr);
      r->state.assign(F, kFree);
      r->seq.assign(F, 0);
      r->team_of.assign(F, 0);
      for (int f = 0; f < F; f++) {
        r->out_rows[f] = pinned<double>(9 * (size_t)e.fr[f].cap);
        r->out_status[f] = pinned<int>(e.fr[f].cap);
      }
      DAO_CHECK(hipMalloc((void**)&r->d_head, 3 * sizeof(long long)));
      DAO_CHECK(hipMemset(r->d_head, 0, 3 * sizeof(long long)));
      DAO_CHECK(hipMalloc((void**)&r->d_rows, sizeof(double) * 7 * (size_t)std::max(1LL, e.cap)));
      DAO_CHECK(hipMalloc((void**)&r->d_pix_cap, sizeof(long long) * F));
      if (team_peaks > 0) {
        DAO_CHECK(hipMalloc((void**)&r->d_team, sizeof(dao::ResTeamSlot) * F));
        DAO_CHECK(hipMemset(r->d_team, 0, sizeof(dao::ResTeamSlot) * F));
        DAO_CHECK(hipMalloc((void**)&r->d_team_ctl, sizeof(dao::TeamCtl) * F));
        DAO_CHECK(hipMemset(r->d_team_ctl, 0, sizeof(dao::TeamCtl) * F));
        const size_t seats = (size_t)F * dao::kMaxTeam * dao::kTeamSeats;
        DAO_CHECK(hipMalloc((void**)&r->d_team_cur, sizeof(int) * seats));
        DAO_CHECK(hipMemset(r->d_team_cur, 0, sizeof(int) * seats));
      }
      DAO_CHECK(hipMemcpy(r->d_pix_cap, e.pix_cap.data(), sizeof(long long) * F, hipMemcpyHostToDevice));
      DAO_CHECK(hipStreamCreateWithFlags(&r->st, hipStreamNonBlocking));
      DAO_CHECK(hipEventCreateWithFlags(&r->ev, hipEventDisableTiming));
      int dev = 0, khz = 0;
      DAO_CHECK(hipGetDevice(&dev));
      DAO_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
      r->tick_hz = khz > 0 ? 1.0e3 * khz : 1.0e8;
      r->idle_ticks = (long long)(idle_seconds * r->tick_hz);
      // by default every compute unit but kResidentSpareCUs: a resident
      // workgroup holds its CU's registers, and other kernels (the caller's
      // own work) need somewhere to run
      r->blocks = workgroups > 0 ? workgroups : std::max(1, e.num_cus - kResidentSpareCUs);
      r->max_iterations = max_iterations;
      e.sync();
    } catch (...) {
      r->release();
      delete r;
      r = nullptr;
      throw;
    }
  });
  (void)rc;
  return r;
}

void daoResidentDestroy(daoResident* r) {
  if (r == nullptr) return;
  r->release();
  delete r;
}

int daoResidentSubmit(daoResident* r, int slot, int width, int height, const double* image, const double* background,
                      const double* rows, int n_rows, const char* p_type, int on_device, int64_t* ticket) {
  return dao::guarded([&] {
    std::lock_guard<std::mutex> lock(r->mu);
    Engine& e = r->eng();
    if (slot < 0 || slot >= e.F) dao::fail(dao::kErrArg, "slot index out of range");
    if (r->state[slot] != kFree) dao::fail(dao::kErrArg, "the slot holds a request not yet collected");
    const int fmt = dao_peak_format(p_type);
    if (fmt < 0) dao::fail(dao::kErrArg, "unknown peak type");
    const long long npx = (long long)width * height;
    if (width < 1 || height < 1 || npx > e.pix_cap[slot]) dao::fail(dao::kErrArg, "the frame exceeds the slot");
    if (n_rows < 0 || n_rows > e.fr[slot].cap) dao::fail(dao::kErrArg, "more start peaks than the slot holds");
    if (image == nullptr || background == nullptr || (n_rows > 0 && rows == nullptr))
      dao::fail(dao::kErrArg, "missing frame data");
    const int wid = fmt == dao::PF_TEXT ? 7 : 4;
    const double *img = image, *bg = background, *rw = rows;
    if (!on_device) {
      // copied into the slot's pinned buffer, which the kernel reads in place
      if (r->in_buf[slot] == nullptr)
        r->in_buf[slot] = pinned<double>(2 * (size_t)e.pix_cap[slot] + 7 * (size_t)e.fr[slot].cap);
      double* h = r->in_buf[slot];
      std::memcpy(h, image, sizeof(double) * npx);
      std::memcpy(h + npx, background, sizeof(double) * npx);
      if (n_rows > 0) std::memcpy(h + 2 * npx, rows, sizeof(double) * (size_t)n_rows * wid);
      double* d = device_view(h);
      img = d;
      bg = d + npx;
      rw = d + 2 * npx;
    }
    dao::ResSlot& s = r->slots[slot];
    const long long t = ++r->next_seq;
    // a team of ceil(n / team_peaks) workgroups, at most as many as are not
    // already given to requests in flight: large teams while the device has
    // room (latency), one workgroup each when it is busy (throughput)
    int team = 1;
    if (r->team_peaks > 0) {
      const int want = std::min(dao::kMaxTeam, (n_rows + r->team_peaks - 1) / r->team_peaks);
      team = std::max(1, std::min(want, r->blocks - r->members));
    }
    r->team_of[slot] = team;
    r->members += team;
    s.done = 0;
    s.seq = t;
    s.team = team;
    s.img = img;
    s.bg = bg;
    s.rows = n_rows > 0 ? rw : img;
    s.out = device_view(r->out_rows[slot]);
    s.status = device_view(r->out_status[slot]);
    s.W = width;
    s.H = height;
    s.n = n_rows;
    s.fmt = fmt;
    s.rqe = 1.0;
    s.scm = 0.0;
    // one ticket per team member, member 0 (which loads the frame) first
    for (int m = 0; m < team; m++)
      __atomic_store_n(&r->ring[(r->tail + m) % r->qcap], slot | (m << dao::kResMemberShift), __ATOMIC_RELAXED);
    r->tail += team;
    __atomic_store_n(&r->q->tail, r->tail, __ATOMIC_RELEASE);
    r->state[slot] = kQueued;
    r->seq[slot] = t;
    if (ticket) *ticket = t;
    r->ensure_running();
  });
}

int daoResidentWait(daoResident* r, double timeout_seconds, int max_slots, int32_t* slots_out) {
  int found = 0;
  const int rc = dao::guarded([&] {
    using clock = std::chrono::steady_clock;
    const auto t_end = clock::now() + std::chrono::duration<double>(timeout_seconds > 0 ? timeout_seconds : 0.0);
    for (long long spin = 0;; spin++) {
      {
        std::lock_guard<std::mutex> lock(r->mu);
        bool waiting = false;
        for (int f = 0; f < (int)r->state.size() && found < max_slots; f++) {
          if (r->state[f] != kQueued) continue;
          if (__atomic_load_n(&r->slots[f].done, __ATOMIC_ACQUIRE) == r->seq[f]) {
            r->state[f] = kFinished;
            slots_out[found++] = f;
          } else {
            waiting = true;
          }
        }
        if (found > 0) return;
        // the workgroups left (idle) while requests wait: start them again
        if (waiting && (spin & 255) == 0) r->ensure_running();
        if (!waiting) spin = -1;  // nothing outstanding: sleep below
      }
      if (clock::now() >= t_end) return;
      if (spin < 0) {
        std::this_thread::sleep_for(std::chrono::microseconds(200));
        spin = 0;
      } else if (spin > 64) {
        std::this_thread::yield();
      }
    }
  });
  return rc != dao::kOk ? rc : found;
}

int daoResidentResult(daoResident* r, int slot, double* out, int32_t* status, int32_t* counters8, double* info4) {
  int n = 0;
  const int rc = dao::guarded([&] {
    std::lock_guard<std::mutex> lock(r->mu);
    if (slot < 0 || slot >= (int)r->state.size() || r->state[slot] != kFinished)
      dao::fail(dao::kErrArg, "the slot holds no finished request");
    const dao::ResSlot& s = r->slots[slot];
    if (counters8) for (int q = 0; q < dao::C_COUNT; q++) counters8[q] = s.counters[q];
    if (info4) {
      info4[0] = s.sweeps;
      info4[1] = (double)(s.t_done - s.t_take) / r->tick_hz;
      info4[2] = (double)(s.t_load - s.t_take) / r->tick_hz;
      info4[3] = 0.0;
    }
    r->state[slot] = kFree;
    r->seq[slot] = 0;
    r->members -= r->team_of[slot];
    r->team_of[slot] = 0;
    if (s.err) {
      n = dao::kErrFrame;
      dao::set_last_error(dao::kErrFrame,
                          s.err == 2   ? "the request did not fit its slot"
                          : s.err == 3 ? "request refused: a pixel is not finite and positive, or a background or "
                                         "start-peak value is not finite"
                                       : "the frame failed: a scheduler wait gave up");
      return;
    }
    n = s.nfit;
    if (out) std::memcpy(out, r->out_rows[slot], sizeof(double) * 9 * (size_t)n);
    if (status) std::memcpy(status, r->out_status[slot], sizeof(int32_t) * (size_t)n);
  });
  return rc != dao::kOk ? rc : n;
}

void daoResidentStats(daoResident* r, int64_t* out2) {
  std::lock_guard<std::mutex> lock(r->mu);
  out2[0] = r->next_seq;
  out2[1] = r->launches;
}

}  // extern "C"
