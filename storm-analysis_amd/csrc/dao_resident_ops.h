#pragma once
// The resident fitter kernel: workgroups that stay on the device and fit
// frames as their requests arrive (continuous batching without a launch per
// batch).  Each workgroup repeatedly
//   1. takes the next ticket (the device counter `head` below the host's
//      `tail`), or leaves once no ticket has arrived for idle_ticks or the
//      host asked it to stop;
//   2. loads the ticket's frame into its slot: mFitNewImage and
//      mFitNewBackground (multi_fit.c:1360-1408) per pixel, with the slot's
//      uniform calibration;
//   3. adds the start peaks (daoNewPeaks, dao_fit.c:1582-1766: every peak
//      added in index order, then every new peak's first error) and runs
//      MultiFitter.doFit (dao_fit_c.py:332-356) -- the ordered executor and
//      the LM turns of k_fit_frames, so a frame's result is the same as in
//      a batched launch;
//   4. writes the peaks and counters to the slot's mapped host memory and
//      completes the ticket.
// Frames share no state, so a frame's result does not depend on which other
// frames are on the device at the same time.
#include "dao_ops.h"
#include "dao_resident.h"

namespace dao {

// Device wall clock (constant rate, hipDeviceAttributeWallClockRate).
__device__ inline long long wall_ticks() { return (long long)__builtin_amdgcn_s_memrealtime(); }

// Host-written data is read, and results for the host written, with
// system-scope accesses: they bypass the caches that are not coherent with
// the host (and, for inputs in device memory written by another engine,
// with other agents).
template <class T>
__device__ inline T sys_ld(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T>
__device__ inline void sys_st(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Thread 0: the next ticket (slot | member << kResMemberShift), or -1 when
// the workgroup should
// leave: the host asked to stop, or no frame is being fitted and no
// workgroup has taken a ticket for idle_ticks (the whole kernel then leaves
// at about the same time, and the host starts it again with the next
// request; a workgroup that merely found no ticket itself keeps waiting, so
// the pool does not shrink while frames are in flight).
//
// An idle workgroup polls the host's tail every ~14 us and reads the shared
// head only when the tail has moved past what it last saw: hundreds of
// workgroups polling shared words without pause would load the memory
// channels (and the host link) that the fitting workgroups use.  Every
// iteration either takes a ticket, loses one to another workgroup (the head
// grew), or sleeps; the stop flag and the idle bound are checked every 16
// sleeps, so the loop ends.
__device__ inline int res_take(const ResArgs& r) {
  const long long t0 = wall_ticks();
  long long seen = -1;  // every ticket below this was taken when last looked
  for (int it = 0;; it++) {
    const long long tail = __hip_atomic_load(&r.q->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
    if (tail > seen) {
      long long h = __hip_atomic_load(r.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (h < tail) {
        if (__hip_atomic_compare_exchange_strong(r.head, &h, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
          __hip_atomic_fetch_add(r.busy, 1LL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_fetch_max(r.last, wall_ticks(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const int t = sys_ld(&r.ring[h % r.qcap]);
          const int f = t & ((1 << kResMemberShift) - 1);
          if (t >= 0 && f < r.nslots) return t;
          __hip_atomic_fetch_add(r.busy, -1LL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        continue;
      }
      seen = h;
    }
    if ((it & 15) == 15) {
      if (sys_ld(&r.q->stop) != 0) return -1;
      const long long busy = __hip_atomic_load(r.busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const long long last = max(t0, __hip_atomic_load(r.last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      if (busy == 0 && wall_ticks() - last > r.idle_ticks) return -1;
    }
    for (int z = 0; z < 4; z++) __builtin_amdgcn_s_sleep(127);
  }
}

// The request fields of a slot, read once by thread 0.
struct ResReq {
  long long seq;
  const double* img;
  const double* bg;
  const double* rows;
  double* out;
  int* status;
  int W, H, n, fmt, team;
  double rqe, scm;
  long long t_take, t_load;
};

// Thread 0: the request fields of slot rs.
__device__ inline void res_read(const ResSlot* rs, ResReq& q) {
  q.seq = sys_ld(&rs->seq);
  q.img = sys_ld(&rs->img);
  q.bg = sys_ld(&rs->bg);
  q.rows = sys_ld(&rs->rows);
  q.out = sys_ld(&rs->out);
  q.status = sys_ld(&rs->status);
  q.W = sys_ld(&rs->W);
  q.H = sys_ld(&rs->H);
  q.n = sys_ld(&rs->n);
  q.fmt = sys_ld(&rs->fmt);
  q.team = sys_ld(&rs->team);
  q.rqe = sys_ld(&rs->rqe);
  q.scm = sys_ld(&rs->scm);
  q.t_take = wall_ticks();
  q.t_load = q.t_take;
}

// The n peaks of frame fr as k_pack_peaks' rows into the request's mapped
// host memory (records read through M: device-coherent after a team fit).
template <class M>
__device__ void res_pack(const FitArgs& a, const Frame& fr, int n, const ResReq& q) {
  const Model& m = a.m;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const PeakRec* rec = a.pk.rec + fr.peak0 + k;
    double* o = q.out + 9LL * k;
    sys_st(o + 0, M::ld(&rec->prm[P_XC]) + m.xoff);
    sys_st(o + 1, M::ld(&rec->prm[P_YC]) + m.xoff);
    sys_st(o + 2, M::ld(&rec->prm[P_ZC]) + m.zoff);
    sys_st(o + 3, M::ld(&rec->prm[P_HEIGHT]));
    sys_st(o + 4, M::ld(&rec->prm[P_BG]));
    sys_st(o + 5, sqrt(1.0 / (2.0 * M::ld(&rec->prm[P_XW]))));
    sys_st(o + 6, sqrt(1.0 / (2.0 * M::ld(&rec->prm[P_YW]))));
    sys_st(o + 7, M::ld(&rec->err));
    sys_st(o + 8, (double)M::ld(&rec->iter));
    sys_st(q.status + k, M::ld(&rec->status));
  }
}

// Completes the request of slot f: counters (cnt[C_COUNT], in LDS or read
// through M), peak count, sweeps, error code, then `done` = seq last.
template <class M>
__device__ void res_publish(const ResArgs& r, int f, const ResReq& q, const int* cnt, int nfit, int sweeps, int err) {
  ResSlot* const rs = r.slots + f;
  if (threadIdx.x < C_COUNT) sys_st(&rs->counters[threadIdx.x], M::ld(&cnt[threadIdx.x]));
  // every thread's result stores complete before thread 0 publishes
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
  __syncthreads();
  if (threadIdx.x == 0) {
    sys_st(&rs->nfit, nfit);
    sys_st(&rs->sweeps, sweeps);
    sys_st(&rs->err, err);
    sys_st(&rs->t_take, q.t_take);
    sys_st(&rs->t_load, q.t_load);
    sys_st(&rs->t_done, wall_ticks());
    __hip_atomic_store(&rs->done, q.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Loads the ticket's request into q and the frame into slot f, adds its
// start peaks and their first errors; returns the slot's frame record.  Out
// of line: it runs once per frame, and its registers stay out of the LM
// loop's allocation.
template <int PPL, int SHAPE, int NW, bool MP, int EM>
__device__ __noinline__ Frame res_load(const FitArgs& a, const ResArgs& r, int f, FpwShared<NW>& sh, ResReq& q,
                                       double* slice) {
  ResSlot* const rs = r.slots + f;
  if (threadIdx.x == 0) {
    res_read(rs, q);
    sh.failed = 0;
    a.frame_err[f] = 0;
  }
  if (threadIdx.x < C_COUNT) sh.cnt[threadIdx.x] = 0;
  __syncthreads();
  Frame fr = a.frames[f];
  // the host checked the request against the slot; checked again here, so
  // that no request can address memory outside its slot
  const bool fits = q.W >= 1 && q.H >= 1 && (long long)q.W * q.H <= r.pix_cap[f] && q.n >= 0 && q.n <= fr.cap &&
                    (q.fmt == PF_TEXT || q.fmt == PF_FINDER || q.fmt == PF_TESTING);
  fr.W = fits ? q.W : 0;
  fr.H = fits ? q.H : 0;
  fr.nfit = 0;
  if (!fits) {
    if (threadIdx.x == 0) sh.failed = 2;
    return fr;
  }
  const int n = q.n;
  const int wid = q.fmt == PF_TEXT ? 7 : 4;

  // mFitNewImage + mFitNewBackground (k_new_image, k_new_background); a
  // frame with a pixel that is not finite and positive (the MLE deviance
  // takes its logarithm), or a non-finite background or start-peak value,
  // is refused
  const long long npx = (long long)fr.W * fr.H;
  const double rqe = q.rqe, scm = q.scm;
  bool bad = false;
  for (long long i = threadIdx.x; i < npx; i += blockDim.x) {
    const double im = sys_ld(q.img + i);
    const double b = sys_ld(q.bg + i);
    bad |= !(im > 0.0) || !isfinite(im) || !isfinite(b);
    PixRO ro;
    ro.x = im * rqe + scm;
    ro.rqe = rqe;
    ro.scm = scm;
    ro.bgest = b * rqe;
    a.px.ro[fr.pix0 + i] = ro;
    PixRW w;
    w.f = 0.0;
    w.bg = 0.0;
    w.tfi = 0.0;
    w.cnt = 0;
    w.stale = 1;
    a.px.rw[fr.pix0 + i] = w;
  }
  // start rows into device memory, op k = new peak k in slot peak0 + k, its
  // box the ROI at the floor of its start position (k_box_newpeak)
  double* const rows = r.rows_dev + 7LL * fr.peak0;
  for (int i = threadIdx.x; i < n * wid; i += blockDim.x) {
    rows[i] = sys_ld(q.rows + i);
    bad |= !isfinite(rows[i]);
  }
  if (__any(bad) && lane_id() == 0) sh.failed = 3;
  __syncthreads();
  if (sh.failed) return fr;
  add_peaks<PPL, NW>(a, f, fr, n, rows, q.fmt, fr.peak0, sh, slice);
  fr.nfit = n;
  if (threadIdx.x == 0) q.t_load = wall_ticks();
  return fr;
}

// A request fitted by this workgroup alone (the frame-per-workgroup fitter).
template <int PPL, int SHAPE, int NW, bool MP, int EM>
__device__ void res_frame(const FitArgs& a, const ResArgs& r, int f, FpwShared<NW>& sh, ResReq& q, double* slice) {
  const Frame fr = res_load<PPL, SHAPE, NW, MP, EM>(a, r, f, sh, q, slice);
  __syncthreads();
  const int n = fr.nfit;
  const int sweeps = sh.failed ? 0 : fit_frame<PPL, SHAPE, NW, MP, EM>(a, f, fr, sh, slice);
  const int err = sh.failed == 0 ? 0 : (sh.failed >= 2 ? sh.failed : 1);
  if (err == 0) res_pack<MemL1>(a, fr, n, q);
  res_publish<MemL1>(r, f, q, sh.cnt, err == 0 ? n : 0, sweeps, err);
  if (threadIdx.x == 0) __hip_atomic_fetch_add(r.busy, -1LL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One member workgroup of a team request (ResSlot::team > 1).  Member 0
// loads the frame as res_frame does, resets the team's scheduling state and
// then lets the other members in (ResTeamSlot::ready); every member runs the
// team fitter (team_fit: the frame's ops on any CU, in the reference's order
// where ROIs overlap); the member that leaves last writes the result.  A
// member never waits for one that has not started, except for member 0's
// load, whose ticket precedes the others' in the ring.
template <int PPL, int SHAPE, int NW, bool MP, int EM>
__device__ void res_team(const FitArgs& a, const ResArgs& r, int f, int member, FpwShared<NW>& fsh,
                         TeamShared& tsh, ResReq& q, double* slice, int* flag) {
  ResTeamSlot* const ts = r.team + f;
  if (member == 0) {
    const Frame fr0 = res_load<PPL, SHAPE, NW, MP, EM>(a, r, f, fsh, q, slice);
    __syncthreads();
    const int s0 = f * kMaxTeam * kTeamSeats;
    const int T = max(1, min(kMaxTeam, q.team));
    for (int s = threadIdx.x; s < T * kTeamSeats; s += blockDim.x) MemDev::st(&a.team_cur[s0 + s], kIdle);
    if (threadIdx.x < C_COUNT) MemDev::st(&ts->cnt[threadIdx.x], fsh.cnt[threadIdx.x]);
    if (threadIdx.x == 0) {
      TeamCtl* const ctl = a.team_ctl + f;
      MemDev::st(&ctl->phase, 0);
      MemDev::st(&ctl->lead, 0);
      MemDev::st(&ctl->kind, 0);
      MemDev::st(&ctl->n, 0);
      MemDev::st(&ctl->next, 0);
      MemDev::st(&ctl->done, 0);
      MemDev::st(&ctl->esc, kNoEscape);
      MemDev::st(&ctl->it, 0);
      MemDev::st(&ctl->sweeps, 0);
      MemDev::st(&ctl->nrun, 0);
      MemDev::st(&a.frame_err[f], 0);
      MemDev::st(&a.frame_sweeps[f], 0);
      MemDev::st(&ts->left, 0);
      MemDev::st(&ts->err, fsh.failed == 0 ? 0 : (fsh.failed >= 2 ? fsh.failed : 1));
      MemDev::st(&ts->W, fr0.W);
      MemDev::st(&ts->H, fr0.H);
      MemDev::st(&ts->nfit, fr0.nfit);
    }
    dev_order();
    __syncthreads();
    if (threadIdx.x == 0) MemDev::st(&ts->ready, q.seq);
  } else if (threadIdx.x == 0) {
    res_read(r.slots + f, q);
    // member 0 took its ticket first and is loading the frame; the wait is
    // bounded all the same (a member that gives up fits nothing)
    q.W = 0;
    for (long long t0 = wall_ticks();;) {
      if (MemDev::ld(&ts->ready) == q.seq) {
        q.W = 1;
        break;
      }
      if (wall_ticks() - t0 > 20 * r.idle_ticks) break;
      __builtin_amdgcn_s_sleep(2);
    }
    dev_order();
  }
  __syncthreads();
  const int T = max(1, min(kMaxTeam, q.team));
  if (member == 0 || q.W == 1) {
    Frame fr = a.frames[f];
    fr.W = MemDev::ld(&ts->W);
    fr.H = MemDev::ld(&ts->H);
    fr.nfit = MemDev::ld(&ts->nfit);
    if (uniform_i(MemDev::ld(&ts->err)) == 0) {
      if (threadIdx.x < C_COUNT) tsh.cnt[threadIdx.x] = 0;
      __syncthreads();
      const int s0 = f * kMaxTeam * kTeamSeats;
      team_fit<PPL, SHAPE, NW, MP, EM>(a, f, fr, s0, s0 + T * kTeamSeats, s0 + member * kTeamSeats, tsh, slice);
      if (threadIdx.x < C_COUNT && tsh.cnt[threadIdx.x] != 0) MemDev::add(&ts->cnt[threadIdx.x], tsh.cnt[threadIdx.x]);
    }
  }
  dev_order();
  __syncthreads();
  if (threadIdx.x == 0) *flag = MemDev::add(&ts->left, 1) == T - 1;
  __syncthreads();
  if (*flag) {
    // the last member: every other member's stores were ordered before its
    // increment of `left`
    dev_order();
    Frame fr = a.frames[f];
    fr.W = MemDev::ld(&ts->W);
    fr.H = MemDev::ld(&ts->H);
    fr.nfit = MemDev::ld(&ts->nfit);
    int err = uniform_i(MemDev::ld(&ts->err));
    if (err == 0 && uniform_i(MemDev::ld(&a.frame_err[f])) != 0) err = 1;
    if (err == 0) res_pack<MemDev>(a, fr, fr.nfit, q);
    res_publish<MemDev>(r, f, q, ts->cnt, err == 0 ? fr.nfit : 0, uniform_i(MemDev::ld(&a.frame_sweeps[f])), err);
  }
  if (threadIdx.x == 0) __hip_atomic_fetch_add(r.busy, -1LL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// TEAM = false: every request is fitted by one workgroup (res_frame); TEAM =
// true: every request by a team of ResSlot::team workgroups (res_team, also
// for a team of one).  Two kernels rather than one with both paths: the
// register allocation of either alone keeps the LM loop free of spills.
template <int PPL, int SHAPE, int NW, bool MP, int EM, bool TEAM>
__global__ __launch_bounds__(NW * 64) void k_resident(FitArgs a, ResArgs r) {
  constexpr int kSlice = fit_frames_slice<PPL, MP>() + fit_frames_stage<PPL>();
  // new-peak and first-error ops use the FP64 ROI cache over the same slice
  static_assert(fit_frames_slice<PPL, MP>() >= RoiLds<PPL>::kDoubles, "slice holds the FP64 ROI cache");
  __shared__ FpwShared<NW> sh;
  __shared__ TeamShared tsh;
  __shared__ ResReq q;
  __shared__ int take, flag;
  extern __shared__ double lds_dyn[];
  double* const slice = lds_dyn + (threadIdx.x >> 6) * kSlice;
  for (;;) {
    if (threadIdx.x == 0) take = res_take(r);
    __syncthreads();
    const int t = take;
    __syncthreads();
    if (t < 0) break;
    const int f = t & ((1 << kResMemberShift) - 1), member = t >> kResMemberShift;
    if constexpr (TEAM) {
      res_team<PPL, SHAPE, NW, MP, EM>(a, r, f, member, sh, tsh, q, slice, &flag);
    } else {
      res_frame<PPL, SHAPE, NW, MP, EM>(a, r, f, sh, q, slice);
    }
    __syncthreads();
  }
}

template <int PPL, int SHP, int NW, bool MP, int EM, bool TEAM>
void launch_resident_t(const FitArgs& a, const ResArgs& r, int blocks, hipStream_t st) {
  const size_t bytes = sizeof(double) * (size_t)NW * (fit_frames_slice<PPL, MP>() + fit_frames_stage<PPL>());
  static size_t allowed = 64 * 1024;
  if (bytes > allowed) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_resident<PPL, SHP, NW, MP, EM, TEAM>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      return;  // the launch below fails; the caller's error check reports it
    allowed = bytes;
  }
  k_resident<PPL, SHP, NW, MP, EM, TEAM><<<blocks, NW * 64, bytes, st>>>(a, r);
}

}  // namespace dao
