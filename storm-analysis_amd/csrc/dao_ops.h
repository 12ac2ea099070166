#pragma once
// Device code of the MI355X fitter shared by the kernel translation units:
// the wave helpers, the ROI caches, the LM turn and the other image
// operations, and the frame-per-workgroup fitter (templates and inline
// functions only; the non-template kernels live in dao_kernels.hip).
//
// The reference visits peaks in index order and each peak read-modify-writes
// the fit image over its ROI (storm_analysis/sa_library/multi_fit.c,
// mFitIterateLM:1072-1349).  Here every image operation ("op": one LM turn,
// one recentering, one new peak, ...) reserves a box of pixels, and op j
// starts only after every earlier op i < j of its frame whose box overlaps;
// all other ops run concurrently.  Each op is one 64-lane wavefront: pixel j
// of the ROI lives in lane j % 64, slot j / 64, its fit values in registers
// and its read-only data (x, rqe, scmos) in the wave's LDS slice, so a turn
// reads and writes HBM once.  Sums over the ROI are wavefront DPP reductions; the
// 4..6 parameter damped normal equations are solved in FP64 registers by
// every lane (uniform control flow).
//
// Two drivers run ops:
//  - k_fit_frames: one persistent workgroup per frame runs whole sweeps
//    (mFitIterateLM, or all of doFit) without returning to the host;
//  - the multi-kernel executor (k_bin_*, k_preds, k_round): one launch per
//    round, for image operations outside the sweep loop (new peaks,
//    subtraction) and as an alternative sweep driver.
// An LM turn may move its ROI (multi_fit.c:1933-1946).  If a candidate ROI
// leaves the reserved box the turn puts the frame back as it found it; the
// driver then re-runs it in order (see fpw_run and Engine::sweep).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

#include "dao_device.h"

namespace dao {

// ---------------------------------------------------------------------------
// wavefront helpers
// ---------------------------------------------------------------------------

__device__ inline int lane_id() { return threadIdx.x & (kWave - 1); }

// A value every lane of the wave holds, in a scalar register: branches on it
// are uniform.  Scheduler data (op index, peak index, box) are per op, hence
// per wave; saying so keeps the LM turn's control flow uniform.
__device__ inline int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int4 uniform_i4(int4 v) {
  return make_int4(uniform_i(v.x), uniform_i(v.y), uniform_i(v.z), uniform_i(v.w));
}

// v of the lane selected by the DPP control CTRL (quad permutes and row
// mirrors: every lane has a source inside its 16-lane row, so no lane keeps
// an old value and none needs to be supplied).
template <int CTRL>
__device__ inline double dpp_mov(double v) {
  int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
  int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ inline double lane_value(double v, int lane) {
  int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Lanes of one wave exchange values through LDS (the shape table): stores
// before this point are seen by every lane's loads after it.  The wave
// barrier alone touches no memory, so the compiler could move LDS accesses
// across it; the wavefront-scope fences forbid that (they cost no
// instructions: a wave's LDS accesses execute in order).
__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over the wave, the same bits in every lane: pairs, quads, 8- and
// 16-lane rows by DPP (each step adds two equal-order partial sums, so every
// lane of a row holds the identical row sum), then (r0 + r1) + (r2 + r3).
__device__ inline double wsum(double v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// Sums over the wave of K values per lane (v[k] of every lane), each summed
// in wsum's order, into out[0, K) (LDS, read back as uniform values).  A
// transposed butterfly: at the pair level (lane bit 0) each lane keeps half
// of the values and adds its partner's copy of that half; at the quad level
// (bit 1) half again; the K/4 partial sums of each lane then go to `red`
// (64 * kRedSlots doubles of the wave's LDS slice), where lane k < K adds
// value k's quads and rows in wsum's order.  Every sum is formed from the
// same operands in the same tree as wsum's, so the results are wsum's bit for
// bit, with about a quarter of the cross-lane moves.
constexpr int kRedSlots = 7;  // ceil(ceil(K / 2) / 2) for K <= 28

template <class T, int K>
__device__ inline T pick(const T (&v)[K], int k) {
  return k < K ? v[k] : T(0);
}

template <int K>
struct Fold {
  static constexpr int K1 = (K + 1) & ~1;
  static constexpr int H1 = K1 / 2;          // values per lane after the pair level
  static constexpr int H2 = (H1 + 1) / 2;    // after the quad level
};

template <int CTRL, int K, int H>
__device__ inline void fold_level(const double (&v)[K], double (&out)[H], bool upper) {
#pragma unroll
  for (int i = 0; i < H; i++) {
    const double lo = pick(v, i), hi = pick(v, i + H);
    const double keep = upper ? hi : lo, send = upper ? lo : hi;
    out[i] = keep + dpp_mov<CTRL>(send);
  }
}

template <int K>
__device__ void wsum_many(const double (&v)[K], double* red, double* out) {
  using F = Fold<K>;
  static_assert(F::H2 <= kRedSlots, "reduction slots");
  const int l = lane_id();
  double a[F::H1];
  fold_level<0xB1>(v, a, (l & 1) != 0);  // quad_perm [1,0,3,2]: lane bit 0
  double b[F::H2];
  fold_level<0x4E>(a, b, (l & 2) != 0);  // quad_perm [2,3,0,1]: lane bit 1
#pragma unroll
  for (int j = 0; j < F::H2; j++) red[j * kWave + l] = b[j];
  wave_sync_lds();
  if (l < K) {
    const int i = l % F::H1;
    const int lane0 = ((i / F::H2) << 1) | (l / F::H1);  // holder of value l in quad 0 of row 0
    const double* q = red + (i % F::H2) * kWave + lane0;
    double r[4];
#pragma unroll
    for (int row = 0; row < 4; row++) {
      const double* p = q + row * 16;
      r[row] = (p[0] + p[4]) + (p[8] + p[12]);
    }
    out[l] = (r[0] + r[1]) + (r[2] + r[3]);
  }
  wave_sync_lds();
}

// ---------------------------------------------------------------------------
// access to the mutable shared arrays
// ---------------------------------------------------------------------------
//
// The pixel fit state, the peak records and the op lists change during a
// launch.  MemL1: ordinary loads and stores, coherent among the waves of one
// workgroup (they share one CU's vector L1, which writes through).  MemDev:
// agent-scope relaxed atomics, which go around the caches that are not
// coherent between CUs and between XCDs, for data that workgroups on
// different CUs share inside one launch; dev_order() then orders them.

// All of the fitter's arrays are in global memory.  Pointers read out of the
// kernel's argument struct are generic to the compiler once that struct's
// address reaches a function that is not inlined, and generic accesses
// compile to flat instructions, whose completion every LDS wait (lgkmcnt)
// then also waits for.  gptr() states the address space.
template <class T>
using GlobalT = __attribute__((address_space(1))) T;
template <class T>
__device__ inline const GlobalT<T>* gptr(const T* p) { return (const GlobalT<T>*)p; }
template <class T>
__device__ inline GlobalT<T>* gptr(T* p) { return (GlobalT<T>*)p; }

struct MemL1 {
  template <class T>
  __device__ static T ld(const T* p) { return *p; }
  template <class T>
  __device__ static void st(T* p, const T& v) { *p = v; }
  // stores of data that is rarely read back (turn journals): streamed
  template <class T>
  __device__ static void st_stream(T* p, const T& v) { __builtin_nontemporal_store(v, p); }
};

struct MemDev {
  template <class T>
  __device__ static T ld(const T* p) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "scalar access");
    return __hip_atomic_load(gptr(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <class T>
  __device__ static void st(T* p, const T& v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "scalar access");
    __hip_atomic_store(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <class T>
  __device__ static void st_stream(T* p, const T& v) { st(p, v); }
  // atomic read-modify-writes of the team's control words
  __device__ static int add(int* p, int v) {
    return __hip_atomic_fetch_add(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static void min(int* p, int v) {
    __hip_atomic_fetch_min(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static int cas(int* p, int expected, int desired) {
    __hip_atomic_compare_exchange_strong(gptr(p), &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return expected;
  }
  __device__ static int4 ld(const int4* p) {
    const long long* q = reinterpret_cast<const long long*>(p);
    const long long a = ld(q), b = ld(q + 1);
    return make_int4((int)a, (int)(a >> 32), (int)b, (int)(b >> 32));
  }
  __device__ static void st(int4* p, const int4& v) {
    long long* q = reinterpret_cast<long long*>(p);
    st(q, (long long)(unsigned)v.x | ((long long)v.y << 32));
    st(q + 1, (long long)(unsigned)v.z | ((long long)v.w << 32));
  }
  __device__ static PixRW ld(const PixRW* p) {
    PixRW w;
    w.f = ld(&p->f);
    w.bg = ld(&p->bg);
    w.tfi = ld(&p->tfi);
    const long long cs = ld(reinterpret_cast<const long long*>(&p->cnt));
    w.cnt = (int)cs;
    w.stale = (int)(cs >> 32);
    return w;
  }
  __device__ static void st(PixRW* p, const PixRW& w) {
    st(&p->f, w.f);
    st(&p->bg, w.bg);
    st(&p->tfi, w.tfi);
    st(reinterpret_cast<long long*>(&p->cnt), (long long)(unsigned)w.cnt | ((long long)w.stale << 32));
  }
};

// An agent-scope acquire-release fence: every memory access of this wave
// before it happens before every access after it, for every workgroup of the
// launch that synchronises with it.  Each publish of the cross-workgroup
// protocols (op done, seat idle, generation, escape) is a relaxed MemDev
// store after a dev_order(); each observation is a relaxed MemDev load
// followed by a dev_order() -- release and acquire fences around relaxed
// atomics, so the ordering follows from the HIP (C++) memory model alone.
// On gfx950 it compiles to buffer_wbl2 sc1, s_waitcnt vmcnt(0) and
// buffer_inv sc1: the writeback and invalidate cover cached (non-MemDev)
// accesses too.
__device__ inline void dev_order() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent"); }

// ---------------------------------------------------------------------------
// per-peak uniform state (every lane holds the same copy)
// ---------------------------------------------------------------------------

struct PeakState {
  double prm[kNFit];
  double lam;
  double err;
  int xi, yi;
  int status;
  int iter;
  int added;
};

template <class M = MemL1>
__device__ inline void load_state(const PeakRec* r, PeakState& s) {
#pragma unroll
  for (int q = 0; q < kNFit; q++) s.prm[q] = M::ld(&r->prm[q]);
  s.lam = M::ld(&r->lam);
  s.err = M::ld(&r->err);
  s.xi = M::ld(&r->xi);
  s.yi = M::ld(&r->yi);
  s.status = M::ld(&r->status);
  s.iter = M::ld(&r->iter);
  s.added = M::ld(&r->added);
}

// The device-coherent forms move the 96-byte record as twelve 8-byte words,
// one per lane, in one instruction (each coherent access is its own request
// to the memory side), and broadcast them with readlane.
constexpr int kRecWords = (int)(sizeof(PeakRec) / 8);
static_assert(sizeof(PeakRec) == 96 && offsetof(PeakRec, lam) == 56 && offsetof(PeakRec, err) == 64 &&
                  offsetof(PeakRec, xi) == 72 && offsetof(PeakRec, yi) == 76 && offsetof(PeakRec, status) == 80 &&
                  offsetof(PeakRec, iter) == 84 && offsetof(PeakRec, added) == 88,
              "PeakRec word layout of load_state / store_state<MemDev>");

__device__ inline long long lane_word(long long v, int lane) {
  const int lo = __builtin_amdgcn_readlane((int)v, lane);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

template <>
__device__ inline void load_state<MemDev>(const PeakRec* r, PeakState& s) {
  const int l = lane_id();
  const long long v = l < kRecWords ? MemDev::ld(reinterpret_cast<const long long*>(r) + l) : 0;
#pragma unroll
  for (int q = 0; q < kNFit; q++) s.prm[q] = __longlong_as_double(lane_word(v, q));
  s.lam = __longlong_as_double(lane_word(v, 7));
  s.err = __longlong_as_double(lane_word(v, 8));
  const long long w9 = lane_word(v, 9), w10 = lane_word(v, 10), w11 = lane_word(v, 11);
  s.xi = (int)w9;
  s.yi = (int)(w9 >> 32);
  s.status = (int)w10;
  s.iter = (int)(w10 >> 32);
  s.added = (int)w11;
}

template <class M = MemL1>
__device__ inline void store_state(PeakRec* r, const PeakState& s) {
  if (lane_id() == 0) {
#pragma unroll
    for (int q = 0; q < kNFit; q++) M::st(&r->prm[q], s.prm[q]);
    M::st(&r->lam, s.lam);
    M::st(&r->err, s.err);
    M::st(&r->xi, s.xi);
    M::st(&r->yi, s.yi);
    M::st(&r->status, s.status);
    M::st(&r->iter, s.iter);
    M::st(&r->added, s.added);
  }
}

template <>
__device__ inline void store_state<MemDev>(PeakRec* r, const PeakState& s) {
  const int l = lane_id();
  long long v = 0;
#pragma unroll
  for (int q = 0; q < kNFit; q++)
    if (l == q) v = __double_as_longlong(s.prm[q]);
  if (l == 7) v = __double_as_longlong(s.lam);
  if (l == 8) v = __double_as_longlong(s.err);
  if (l == 9) v = (long long)(((unsigned long long)(unsigned)s.yi << 32) | (unsigned)s.xi);
  if (l == 10) v = (long long)(((unsigned long long)(unsigned)s.iter << 32) | (unsigned)s.status);
  // word 11 holds `added` and the slot's owner frame, which the fit never changes
  if (l < 10) {
    MemDev::st(reinterpret_cast<long long*>(r) + l, v);
  } else if (l == 10) {
    MemDev::st(reinterpret_cast<long long*>(r) + 10, v);
    MemDev::st(&r->added, s.added);
  }
}

// ---------------------------------------------------------------------------
// Gaussian model (dao_fit.c)
// ---------------------------------------------------------------------------

// daoCalcWidthsFromZ (dao_fit.c:1009-1035); wxz[0] holds wo^2.
__device__ inline void widths_from_z(const Model& m, PeakState& s) {
  double u = (s.prm[P_ZC] - m.wxz[1]) / m.wxz[2];
  double u2 = u * u;
  double t = 1.0 + u2 + m.wxz[3] * u2 * u + m.wxz[4] * u2 * u2;
  s.prm[P_XW] = 2.0 / (m.wxz[0] * t);
  u = (s.prm[P_ZC] - m.wyz[1]) / m.wyz[2];
  u2 = u * u;
  t = 1.0 + u2 + m.wyz[3] * u2 * u + m.wyz[4] * u2 * u2;
  s.prm[P_YW] = 2.0 / (m.wyz[0] * t);
}

// Chain factors d(width)/dz of daoCalcJHZ (dao_fit.c:904-916), without the
// 1/d factor, as there.
__device__ inline void z_grad(const Model& m, const PeakState& s, double& gx, double& gy) {
  double u = (s.prm[P_ZC] - m.wxz[1]) / m.wxz[2];
  double u2 = u * u;
  double t = 1.0 + u2 + m.wxz[3] * u2 * u + m.wxz[4] * u2 * u2;
  double zt = 2.0 * u + 3.0 * m.wxz[3] * u2 + 4.0 * m.wxz[4] * u2 * u;
  gx = -2.0 * zt / (m.wxz[0] * (t * t));
  u = (s.prm[P_ZC] - m.wyz[1]) / m.wyz[2];
  u2 = u * u;
  t = 1.0 + u2 + m.wyz[3] * u2 * u + m.wyz[4] * u2 * u2;
  zt = 2.0 * u + 3.0 * m.wyz[3] * u2 + 4.0 * m.wyz[4] * u2 * u;
  gy = -2.0 * zt / (m.wyz[0] * (t * t));
}

// One separable shape sample per lane: lanes 0..31 hold column offsets
// xt = i - (X - xi) - xoff and ext = exp(-xt^2 XW); lanes 32..63 the rows
// (dao_fit.c:968-998; the row offset also uses xoff, dao_fit.c:979).
__device__ inline void lane_shape(const Model& m, const PeakState& s, double& t, double& e) {
  int l = lane_id();
  int i = l & 31;
  bool row = l >= 32;
  double o = row ? -(s.prm[P_YC] - (double)s.yi) - m.xoff : -(s.prm[P_XC] - (double)s.xi) - m.xoff;
  double w = row ? s.prm[P_YW] : s.prm[P_XW];
  t = (double)i + o;
  e = exp(-t * t * w);
}

// ---------------------------------------------------------------------------
// turn journal (layout at kJournalSnaps)
// ---------------------------------------------------------------------------

// Journals slot s of this lane.
template <int PPL, class M>
__device__ inline void journal_pixel(double* snap, int s, double f, double bg, double tfi) {
  const int l = lane_id();
  M::st_stream(snap + s * kWave + l, f);
  M::st_stream(snap + (PPL + s) * kWave + l, bg);
  M::st_stream(snap + (2 * PPL + s) * kWave + l, tfi);
}

template <int PPL, class M>
__device__ inline void journal_header(double* snap, unsigned stale, int xi, int yi) {
  long long* h = reinterpret_cast<long long*>(snap) + 3 * PPL * kWave + lane_id();
  M::st_stream(h, (long long)stale | ((long long)(xi | (yi << 16)) << 32));
}

// Writes snapshot snap's foreground, background sums, cached values and
// stale flags back to the pixels it was taken from (c supplies the ROI's
// pixel offsets).  M accesses the pixels, JM the snapshot.
template <int PPL, class M, class JM, class R>
__device__ void journal_restore(const Pix& px, const Frame& fr, const R& c, const double* snap) {
  const int l = lane_id();
  const long long h = JM::ld(reinterpret_cast<const long long*>(snap) + 3 * PPL * kWave + l);
  const unsigned stale = (unsigned)h;
  const int org = (int)(h >> 32);
  const long long base = fr.pix0 + (long long)(org >> 16) * fr.W + (org & 0xffff);
#pragma unroll
  for (int s = 0; s < PPL; s++) {
    if (c.ok(s)) {
      PixRW* q = px.rw + base + c.koff[s];
      M::st(&q->f, JM::ld(snap + s * kWave + l));
      M::st(&q->bg, JM::ld(snap + (PPL + s) * kWave + l));
      M::st(&q->tfi, JM::ld(snap + (2 * PPL + s) * kWave + l));
      M::st(&q->stale, (int)((stale >> s) & 1u));
    }
  }
}

// Puts back, newest first, the nsnap snapshots of journal jr (snapshot 0 in
// the wave's LDS `stage` instead when given): every pixel ends with the
// state its first snapshot holds (counts aside).
template <int PPL, class M, class R>
__device__ void journal_unwind(const Pix& px, const Frame& fr, const R& c, const double* jr, int nsnap,
                               const double* stage = nullptr) {
  for (int v = nsnap - 1; v >= 0; v--) {
    dev_order();  // the previous writes of these pixels have completed
    const double* snap = (v == 0 && stage) ? stage : jr + v * journal_snap_words(PPL);
    if constexpr (std::is_same<M, MemL1>::value) {
      journal_restore<PPL, M, M>(px, fr, c, snap);
    } else if (v == 0 && stage) {
      journal_restore<PPL, M, MemL1>(px, fr, c, snap);
    } else {
      journal_restore<PPL, M, M>(px, fr, c, snap);
    }
  }
}

// One pixel of mFitAddPeak / mFitSubtractPeak (multi_fit.c:50-55,
// 1916-1920): f +/-= (h psf) rqe and bg +/-= (b rqe) + scmos, each operation
// rounded on its own in the reference's order.  Contraction is off here:
// left to the compiler, the product and the sum fuse into an FMA, the
// subtraction then yields the rounding error of the product instead of
// cancelling the addition, and a peak taken out of the fit leaves a residue
// where the reference leaves exact zeros.  (The __d*_rn intrinsics do not
// prevent that: without OCML's rounded operations they are plain operators.)
__device__ inline void stamp_pixel(double& f, double& bg, double h, double psf, double b, double r, double sc,
                                   int sign) {
#pragma clang fp contract(off)
  const double a = (h * psf) * r;
  const double t = b * r + sc;
  if (sign > 0) {
    f = f + a;
    bg = bg + t;
  } else {
    f = f - a;
    bg = bg - t;
  }
}

// ---------------------------------------------------------------------------
// ROI cache of one wave
// ---------------------------------------------------------------------------

// The wave's LDS slice (RoiLds<PPL>::kDoubles doubles): read-only data of the
// ROI pixels -- for slot s of lane l, x at [s*64+l], rqe at [(PPL+s)*64+l],
// scmos at [(2PPL+s)*64+l], each lane touching only its own entries -- then
// the shape table (64 (offset, exp) pairs, written by lane_shape) and the
// reduced J and packed H of the turn (uniform values).
// The lane's ROI geometry in a frame: slot s holds ROI pixel l + 64 s of the
// ROI table (column | row << 8 in rxy, row * W + column in koff); bit s of
// okm is set when that pixel exists.  The same for every turn of a frame, so
// the frame fitter reads the table once per run instead of once per turn.
template <int PPL>
struct RoiGeo {
  int rxy[PPL];
  int koff[PPL];
  unsigned okm;

  __device__ void make(const Model& m, const int* roi_tab, const Frame& fr) {
    const int l = lane_id();
    okm = 0;
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      const int j = l + s * kWave;
      const bool o = j < m.nroi;
      const int v = o ? *gptr(roi_tab + j) : 0;
      const int cx = v & 0xffff, cy = v >> 16;
      rxy[s] = cx | (cy << 8);
      koff[s] = cy * fr.W + cx;
      if (o) okm |= 1u << s;
    }
  }
};

template <int PPL>
struct RoiLds {
  static constexpr int kTab = 3 * PPL * kWave;
  static constexpr int kHJ = kTab + 2 * kWave;
  static constexpr int kRed = kHJ + 32;  // wsum_many's partial sums
  static constexpr int kDoubles = kRed + kRedSlots * kWave;
};

template <int PPL>
struct Roi {
  int rxy[PPL];   // column | row << 8
  int koff[PPL];  // row * W + column
  double f[PPL], bg[PPL], tfi[PPL], psf[PPL];
  int cnt[PPL];
  unsigned stale;  // bit s: slot s stale
  unsigned okm;    // bit s: slot s is a pixel of the ROI
  long long base;  // pixel index of the ROI origin held
  int at_x, at_y;
  bool valid;
  double* lds;

  __device__ bool ok(int s) const { return (okm >> s) & 1u; }
  __device__ int rx(int s) const { return rxy[s] & 0xff; }
  __device__ int ry(int s) const { return rxy[s] >> 8; }
  __device__ long long k(int s) const { return base + koff[s]; }
  __device__ double& x(int s) { return lds[s * kWave + lane_id()]; }
  __device__ double& rqe(int s) { return lds[(PPL + s) * kWave + lane_id()]; }
  __device__ double& scm(int s) { return lds[(2 * PPL + s) * kWave + lane_id()]; }

  __device__ void init(const RoiGeo<PPL>& g, double* slice) {
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      rxy[s] = g.rxy[s];
      koff[s] = g.koff[s];
    }
    okm = g.okm;
    lds = slice;
    valid = false;
  }
  __device__ void init(const Model& m, const int* roi_tab, const Frame& fr, double* slice) {
    RoiGeo<PPL> g;
    g.make(m, roi_tab, fr);
    init(g, slice);
  }

  // Loads the ROI at (xi, yi) (pixels through M); with snap, also journals
  // the pixels' fit state there (see kJournalSnaps; the snapshot through JM).
  template <class M = MemL1, class JM = M>
  __device__ void load(const Pix& px, const Frame& fr, int xi, int yi, double* snap = nullptr) {
    base = fr.pix0 + (long long)yi * fr.W + xi;
    stale = 0;
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      if (ok(s)) {
        const PixRW w = M::ld(px.rw + k(s));
        const GlobalT<PixRO>* ro = gptr(px.ro + k(s));
        const PixRO r = {ro->x, ro->rqe, ro->scm, 0.0};
        f[s] = w.f;
        bg[s] = w.bg;
        tfi[s] = w.tfi;
        cnt[s] = w.cnt;
        if (w.stale) stale |= 1u << s;
        x(s) = r.x;
        rqe(s) = r.rqe;
        scm(s) = r.scm;
      }
    }
    // journal after all loads are issued: a store between them would hold
    // back the later loads (the compiler cannot tell the arrays apart)
    if (snap) {
#pragma unroll
      for (int s = 0; s < PPL; s++)
        if (ok(s)) journal_pixel<PPL, JM>(snap, s, f[s], bg[s], tfi[s]);
      journal_header<PPL, JM>(snap, stale, xi, yi);
    }
    at_x = xi;
    at_y = yi;
    valid = true;
  }

  template <class M = MemL1>
  __device__ void store(const Pix& px) {
    if (!valid) return;
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      if (ok(s)) {
        PixRW w;
        w.f = f[s];
        w.bg = bg[s];
        w.tfi = tfi[s];
        w.cnt = cnt[s];
        w.stale = (stale >> s) & 1u;
        M::st(px.rw + k(s), w);
      }
    }
  }

  template <class M = MemL1>
  __device__ void move_to(const Pix& px, const Frame& fr, int xi, int yi, double* snap = nullptr) {
    if (valid && at_x == xi && at_y == yi) return;
    store<M>(px);
    load<M>(px, fr, xi, yi, snap);
  }

  __device__ double2* tab() { return reinterpret_cast<double2*>(lds + RoiLds<PPL>::kTab); }
  __device__ double* hj() { return lds + RoiLds<PPL>::kHJ; }
  __device__ double* red() { return lds + RoiLds<PPL>::kRed; }

  // Shape table of state st in LDS: entry c < 32 = column c, 32 + r = row r.
  __device__ void table(const Model& m, const PeakState& st) {
    double t, e;
    lane_shape(m, st, t, e);
    wave_sync_lds();
    tab()[lane_id()] = make_double2(t, e);
    wave_sync_lds();
  }

  // psf at the held pixels for state st (cache at st.xi, st.yi); leaves the
  // shape table of st in LDS.
  __device__ void shape(const Model& m, const PeakState& st) {
    table(m, st);
    const double2* tb = tab();
#pragma unroll
    for (int s = 0; s < PPL; s++) psf[s] = tb[32 + ry(s)].y * tb[rx(s)].y;
  }

  // mFitAddPeak / mFitSubtractPeak (multi_fit.c:28-57, 1894-1922).
  __device__ void stamp(const PeakState& st, int sign) {
    double h = st.prm[P_HEIGHT];
    double b = st.prm[P_BG];
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      if (ok(s)) {
        double r = rqe(s);
        double sc = scm(s);
        stamp_pixel(f[s], bg[s], h, psf[s], b, r, sc, sign);
        cnt[s] += sign > 0 ? 1 : -1;
      }
    }
    stale = okm;
  }
};

__device__ inline double anscombe(double v) {
  if (v < (-0.375 + 1.0e-6)) return 0.0;
  return 2.0 * sqrt(v + 0.375);
}

// mFitCalcErr{,ALS,LS,DWLS,FWLS} (multi_fit.c:111-458) over the held ROI.
// Returns 1 when MLE/FWLS meet fi <= 0: pixels before the first such pixel
// (ROI order) are refreshed, the rest keep their cached fi, and the error is
// not updated.  Adds n_neg_fi into neg_fi.
//
// EM >= 0 fixes the error model at compile time (the kernels' MLE
// instances); EM < 0 reads it from m.
template <int PPL, int EM = -1>
__device__ int eval_error(const Model& m, Roi<PPL>& c, const double* asx, double& err, int& neg_fi) {
  const int l = lane_id();
  const int errm = EM >= 0 ? EM : m.errm;
  const bool stop_on_neg = (errm == EM_MLE || errm == EM_FWLS);
  double sum = 0.0;
  bool stopped = false;  // uniform: a slot held the ROI's first fi <= 0
#pragma unroll
  for (int s = 0; s < PPL; s++) {
    if (!stopped) {
      const bool o = c.ok(s);
      double t = o ? c.f[s] + c.bg[s] / ((double)c.cnt[s]) : 1.0;
      const unsigned long long bad = __ballot(o && stop_on_neg && t <= 0.0);
      // pixels of this slot before the first bad one (ROI order j = s*64 + l)
      const int lim = bad ? __builtin_ctzll(bad) : kWave;
      bool zero = false;
      if (o && l < lim) {
        double xk = c.x(s);
        double e;
        if (errm == EM_MLE) {
          e = 2.0 * ((t - xk) - xk * log(t / xk));
        } else if (errm == EM_ALS) {
          double a = anscombe(t);
          zero = (a == 0.0) && ((c.stale >> s) & 1u);
          t = a;
          double d = a - asx[c.k(s)];
          e = d * d;
        } else if (errm == EM_LS) {
          double d = t - xk;
          e = d * d;
        } else if (errm == EM_DWLS) {
          double d = t - xk;
          e = d * d / xk;
        } else {
          double d = t - xk;
          e = d * d / t;
        }
        c.tfi[s] = t;
        c.stale &= ~(1u << s);
        sum += e;
      }
      if (errm == EM_ALS) neg_fi += __popcll(__ballot(zero));
      if (bad) stopped = true;
    }
  }
  if (stopped) {
    neg_fi += 1;
    return 1;
  }
  err = wsum(sum);
  return 0;
}

// daoCalcJH* (dao_fit.c:48-958): J and the upper triangle of H, packed row by
// row, reduced over the wave into the slice's J/H area (J first).  Reads the
// cached transformed fit tfi and the shape table, which must hold st's.
template <int PPL, int SHAPE, int EM = -1>
__device__ void calc_jh(const Model& m, Roi<PPL>& c, const PeakState& st, const double* asx) {
  constexpr int N = SHAPE == SH_2DFIXED ? 4 : (SHAPE == SH_3D ? 6 : 5);
  const int errm = EM >= 0 ? EM : m.errm;
  double a1 = st.prm[P_HEIGHT];
  double wx = st.prm[P_XW];
  double wy = (SHAPE == SH_3D || SHAPE == SH_Z) ? st.prm[P_YW] : wx;  // 2D/2DFixed: XWIDTH twice
  double gx = 0.0, gy = 0.0;
  if (SHAPE == SH_Z) z_grad(m, st, gx, gy);
  const double2* tb = c.tab();

  double j_acc[N];
  double h_acc[N * (N + 1) / 2];
#pragma unroll
  for (int a = 0; a < N; a++) j_acc[a] = 0.0;
#pragma unroll
  for (int a = 0; a < N * (N + 1) / 2; a++) h_acc[a] = 0.0;

#pragma unroll
  for (int s = 0; s < PPL; s++) {
    if (c.ok(s)) {
      const double2 cx = tb[c.rx(s)];
      const double2 cy = tb[32 + c.ry(s)];
      const double xt = cx.x, ex = cx.y, yt = cy.x, ey = cy.y;
      double fi = c.tfi[s];
      double xk = c.x(s);
      double rq = c.rqe(s);
      double et = ex * ey;
      double g[N];
      if (SHAPE == SH_2DFIXED) {
        g[0] = et;
        g[1] = 2.0 * a1 * wx * xt * et;
        g[2] = 2.0 * a1 * wy * yt * et;
        g[3] = 1.0;
      } else if (SHAPE == SH_2D) {
        g[0] = et;
        g[1] = 2.0 * a1 * wx * xt * et;
        g[2] = 2.0 * a1 * wy * yt * et;
        g[3] = -a1 * xt * xt * et - a1 * yt * yt * et;
        g[4] = 1.0;
      } else if (SHAPE == SH_3D) {
        g[0] = et;
        g[1] = 2.0 * a1 * wx * xt * et;
        g[2] = -a1 * xt * xt * et;
        g[3] = 2.0 * a1 * wy * yt * et;
        g[4] = -a1 * yt * yt * et;
        g[5] = 1.0;
      } else {
        g[0] = et;
        g[1] = 2.0 * a1 * wx * xt * et;
        g[2] = 2.0 * a1 * wy * yt * et;
        g[3] = -a1 * xt * xt * gx * et - a1 * yt * yt * gy * et;
        g[4] = 1.0;
      }
      const double rf = 1.0 / fi;
      double sc = (errm == EM_ALS) ? rq * 2.0 * rf : rq;
#pragma unroll
      for (int a = 0; a < N; a++) g[a] *= sc;
      double wj, wh;
      if (errm == EM_MLE) {
        const double xr = xk * rf;
        wj = 2.0 * (1.0 - xr);
        wh = 2.0 * xr * rf;
      } else if (errm == EM_ALS) {
        wj = fi - asx[c.k(s)];
        wh = 1.0;
      } else if (errm == EM_LS) {
        wj = fi - xk;
        wh = 1.0;
      } else if (errm == EM_DWLS) {
        const double rx = 1.0 / xk;
        wj = (fi - xk) * rx;
        wh = rx;
      } else {
        double t1 = (fi - xk) * rf;
        wj = 2.0 * t1 - t1 * t1;
        wh = 2.0 * rf * (xk * rf - t1 + t1 * t1);
      }
      int q = 0;
#pragma unroll
      for (int a = 0; a < N; a++) {
        j_acc[a] += wj * g[a];
#pragma unroll
        for (int b = a; b < N; b++) h_acc[q++] += wh * g[a] * g[b];
      }
    }
  }
  double v[N + N * (N + 1) / 2];  // J, then H packed row by row
#pragma unroll
  for (int a = 0; a < N; a++) v[a] = j_acc[a];
#pragma unroll
  for (int a = 0; a < N * (N + 1) / 2; a++) v[N + a] = h_acc[a];
  wsum_many(v, c.red(), c.hj());
}

// ---------------------------------------------------------------------------
// mixed-precision ROI (MLE)
// ---------------------------------------------------------------------------
//
// The deviance terms, the Jacobian terms and the Hessian in FP32 (the Hessian
// summed over the ROI by FP32 wave reductions); the peak shape, the fit image
// and background accumulators, fi - x, the deviance and Jacobian sums, the
// damped normal equations and the accept/reject test in FP64.  An FP32 shape (or fi - x
// rounded from FP32 values) would put ~1e-4 photons of rounding noise into
// each pixel's model, comparable to the deviance change at which the 1e-6
// tolerance decides convergence, and change the iteration at which peaks stop.

// 32-bit DPP move (as dpp_mov).
template <int CTRL>
__device__ inline float dpp_movf(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}

__device__ inline float lane_valuef(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// FP32 wave sum (same order as wsum), returned in FP64.
__device__ inline double wsumf(float v) {
  v += dpp_movf<0xB1>(v);
  v += dpp_movf<0x4E>(v);
  v += dpp_movf<0x141>(v);
  v += dpp_movf<0x140>(v);
  return ((double)lane_valuef(v, 0) + (double)lane_valuef(v, 16)) +
         ((double)lane_valuef(v, 32) + (double)lane_valuef(v, 48));
}

template <int CTRL, int K, int H>
__device__ inline void fold_levelf(const float (&v)[K], float (&out)[H], bool upper) {
#pragma unroll
  for (int i = 0; i < H; i++) {
    const float lo = pick(v, i), hi = pick(v, i + H);
    const float keep = upper ? hi : lo, send = upper ? lo : hi;
    out[i] = keep + dpp_movf<CTRL>(send);
  }
}

// wsumf of each of K values per lane into out[0, K) (LDS), as wsum_many:
// FP32 sums inside each 16-lane row, the four rows combined in FP64, each
// bit for bit wsumf's.
template <int K>
__device__ void wsumf_many(const float (&v)[K], double* red, double* out) {
  using F = Fold<K>;
  static_assert(F::H2 <= kRedSlots, "reduction slots");
  float* const rf = reinterpret_cast<float*>(red);
  const int l = lane_id();
  float a[F::H1];
  fold_levelf<0xB1>(v, a, (l & 1) != 0);
  float b[F::H2];
  fold_levelf<0x4E>(a, b, (l & 2) != 0);
#pragma unroll
  for (int j = 0; j < F::H2; j++) rf[j * kWave + l] = b[j];
  wave_sync_lds();
  if (l < K) {
    const int i = l % F::H1;
    const int lane0 = ((i / F::H2) << 1) | (l / F::H1);
    const float* q = rf + (i % F::H2) * kWave + lane0;
    double r[4];
#pragma unroll
    for (int row = 0; row < 4; row++) {
      const float* p = q + row * 16;
      r[row] = (double)((p[0] + p[4]) + (p[8] + p[12]));
    }
    out[l] = (r[0] + r[1]) + (r[2] + r[3]);
  }
  wave_sync_lds();
}

// h(u) = u - ln(1 + u) >= 0 for u > -1; the MLE deviance term is
// 2 x h((fi - x) / x).  With s = u / (2 + u), ln(1 + u) = 2 atanh(s) and
// u = 2s / (1 - s), so h = 2 s^2 (1/(1 - s) - s P(s^2)) with
// P = 1/3 + s^2/5 + ... + s^14/17 (truncation < 1e-8 for |s| <= 0.35): no
// cancellation near u = 0, where the direct form loses digits.  Farther out
// the direct form is used.
__device__ inline float mle_h(float u) {
  const float s = u * __builtin_amdgcn_rcpf(2.0f + u);
  const float s2 = s * s;
  float p = 1.0f / 17.0f;
  p = fmaf(p, s2, 1.0f / 15.0f);
  p = fmaf(p, s2, 1.0f / 13.0f);
  p = fmaf(p, s2, 1.0f / 11.0f);
  p = fmaf(p, s2, 1.0f / 9.0f);
  p = fmaf(p, s2, 1.0f / 7.0f);
  p = fmaf(p, s2, 1.0f / 5.0f);
  p = fmaf(p, s2, 1.0f / 3.0f);
  const float hs = 2.0f * s2 * (__builtin_amdgcn_rcpf(1.0f - s) - s * p);
  const float hl = u - __logf(1.0f + u);
  return fabsf(s) <= 0.35f ? hs : hl;
}

// LDS slice of a mixed-precision wave, in doubles: x (FP64), rqe and scmos
// (FP32), the shape table (64 FP64 (offset, exp) pairs), J/H (FP64), and the
// turn-start fi - x of every ROI pixel (FP64, for eval_delta).
template <int PPL>
struct RoiLdsM {
  static constexpr int kRqe = PPL * kWave;
  static constexpr int kScm = kRqe + PPL * kWave / 2;
  static constexpr int kTab = kScm + PPL * kWave / 2;
  static constexpr int kHJ = kTab + 2 * kWave;
  static constexpr int kD0 = kHJ + 32;
  static constexpr int kRed = kD0 + PPL * kWave;  // wsum_many's partial sums
  static constexpr int kDoubles = kRed + kRedSlots * kWave;
};

template <int PPL>
struct RoiM {
  int rxy[PPL];
  int koff[PPL];
  double f[PPL], bg[PPL];
  double psf[PPL];
  float dfi[PPL];  // cached fi - x (the FP64 record keeps fi)
  int cnt[PPL];
  unsigned stale;
  unsigned okm;
  long long base;
  int at_x, at_y;
  bool valid;
  // the turn-start fi - x (d0) is held for the ROI at (d0x, d0y)
  int d0x, d0y;
  bool d0ok;
  double* lds;

  __device__ bool ok(int s) const { return (okm >> s) & 1u; }
  __device__ int rx(int s) const { return rxy[s] & 0xff; }
  __device__ int ry(int s) const { return rxy[s] >> 8; }
  __device__ long long k(int s) const { return base + koff[s]; }
  __device__ double& x(int s) { return lds[s * kWave + lane_id()]; }
  __device__ float& rqe(int s) { return reinterpret_cast<float*>(lds + RoiLdsM<PPL>::kRqe)[s * kWave + lane_id()]; }
  __device__ float& scm(int s) { return reinterpret_cast<float*>(lds + RoiLdsM<PPL>::kScm)[s * kWave + lane_id()]; }
  __device__ double2* tab() { return reinterpret_cast<double2*>(lds + RoiLdsM<PPL>::kTab); }
  __device__ double* hj() { return lds + RoiLdsM<PPL>::kHJ; }
  __device__ double* red() { return lds + RoiLdsM<PPL>::kRed; }
  __device__ double& d0(int s) { return lds[RoiLdsM<PPL>::kD0 + s * kWave + lane_id()]; }

  __device__ void init(const RoiGeo<PPL>& g, double* slice) {
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      rxy[s] = g.rxy[s];
      koff[s] = g.koff[s];
    }
    okm = g.okm;
    d0ok = false;
    lds = slice;
    valid = false;
  }
  __device__ void init(const Model& m, const int* roi_tab, const Frame& fr, double* slice) {
    RoiGeo<PPL> g;
    g.make(m, roi_tab, fr);
    init(g, slice);
  }

  template <class M = MemL1, class JM = M>
  __device__ void load(const Pix& px, const Frame& fr, int xi, int yi, double* snap = nullptr) {
    base = fr.pix0 + (long long)yi * fr.W + xi;
    stale = 0;
    double tfi[PPL];  // the cached values in FP64, for the journal
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      tfi[s] = 0.0;
      if (ok(s)) {
        const PixRW w = M::ld(px.rw + k(s));
        const GlobalT<PixRO>* ro = gptr(px.ro + k(s));
        const PixRO r = {ro->x, ro->rqe, ro->scm, 0.0};
        f[s] = w.f;
        bg[s] = w.bg;
        dfi[s] = (float)(w.tfi - r.x);
        tfi[s] = w.tfi;
        cnt[s] = w.cnt;
        if (w.stale) stale |= 1u << s;
        x(s) = r.x;
        rqe(s) = (float)r.rqe;
        scm(s) = (float)r.scm;
      }
    }
    // journal after all loads are issued (see Roi::load)
    if (snap) {
#pragma unroll
      for (int s = 0; s < PPL; s++)
        if (ok(s)) journal_pixel<PPL, JM>(snap, s, f[s], bg[s], tfi[s]);
      journal_header<PPL, JM>(snap, stale, xi, yi);
    }
    at_x = xi;
    at_y = yi;
    valid = true;
  }

  template <class M = MemL1>
  __device__ void store(const Pix& px) {
    if (!valid) return;
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      if (ok(s)) {
        PixRW w;
        w.f = f[s];
        w.bg = bg[s];
        w.tfi = x(s) + (double)dfi[s];
        w.cnt = cnt[s];
        w.stale = (stale >> s) & 1u;
        M::st(px.rw + k(s), w);
      }
    }
  }

  template <class M = MemL1>
  __device__ void move_to(const Pix& px, const Frame& fr, int xi, int yi, double* snap = nullptr) {
    if (valid && at_x == xi && at_y == yi) return;
    store<M>(px);
    load<M>(px, fr, xi, yi, snap);
  }

  __device__ void table(const Model& m, const PeakState& st) {
    double t, e;
    lane_shape(m, st, t, e);
    wave_sync_lds();
    tab()[lane_id()] = make_double2(t, e);
    wave_sync_lds();
  }

  __device__ void shape(const Model& m, const PeakState& st) {
    table(m, st);
    const double2* tb = tab();
#pragma unroll
    for (int s = 0; s < PPL; s++) psf[s] = tb[32 + ry(s)].y * tb[rx(s)].y;
  }

  __device__ void stamp(const PeakState& st, int sign) {
    const double h = st.prm[P_HEIGHT];
    const double b = st.prm[P_BG];
#pragma unroll
    for (int s = 0; s < PPL; s++) {
      if (ok(s)) {
        const double r = (double)rqe(s);
        const double sc = (double)scm(s);
        stamp_pixel(f[s], bg[s], h, psf[s], b, r, sc, sign);
        cnt[s] += sign > 0 ? 1 : -1;
      }
    }
    stale = okm;
  }
};

// MLE deviance over the held ROI (mFitCalcErr, multi_fit.c:111-195), mixed
// precision; same early stop on fi <= 0 as eval_error.
template <int PPL, int EM = -1>
__device__ int eval_error(const Model& m, RoiM<PPL>& c, const double* asx, double& err, int& neg_fi) {
  const int l = lane_id();
  double sum = 0.0;
  bool stopped = false;
#pragma unroll
  for (int s = 0; s < PPL; s++) {
    if (!stopped) {
      const bool o = c.ok(s);
      const double xk = o ? c.x(s) : 1.0;
      // fi - x = (f - x) + bg / cnt in FP64; 1/cnt by rcp and one Newton step
      const double cn = (double)c.cnt[s];
      double rc = __builtin_amdgcn_rcp(cn);
      rc = fma(fma(-cn, rc, 1.0), rc, rc);
      const float dd = (float)fma(c.bg[s], rc, c.f[s] - xk);
      const float xf = (float)xk;
      const float fi = xf + dd;  // only for the fi <= 0 test
      const unsigned long long bad = __ballot(o && fi <= 0.0f);
      const int lim = bad ? __builtin_ctzll(bad) : kWave;
      if (o && l < lim) {
        const float h = mle_h(dd * __builtin_amdgcn_rcpf(xf));
        sum += (double)(2.0f * xf * h);
        c.dfi[s] = dd;
        c.stale &= ~(1u << s);
      }
      if (bad) stopped = true;
    }
  }
  if (stopped) {
    neg_fi += 1;
    return 1;
  }
  err = wsum(sum);
  return 0;
}

// fi - x of every held pixel in FP64, as eval_error forms it before
// rounding to FP32.
template <int PPL>
__device__ inline double fi_minus_x(RoiM<PPL>& c, int s) {
  const double cn = (double)c.cnt[s];
  double rc = __builtin_amdgcn_rcp(cn);
  rc = fma(fma(-cn, rc, 1.0), rc, rc);
  return fma(c.bg[s], rc, c.f[s] - c.x(s));
}

// Keep the held pixels' fi - x as the turn's start (after a successful
// eval_error of the start state).
template <int PPL>
__device__ void keep_start(RoiM<PPL>& c) {
#pragma unroll
  for (int s = 0; s < PPL; s++)
    if (c.ok(s)) c.d0(s) = fi_minus_x(c, s);
  c.d0x = c.at_x;
  c.d0y = c.at_y;
  c.d0ok = true;
  wave_sync_lds();
}

// The change of the MLE deviance since keep_start(), over the same ROI: with
// e = fi1 - fi0 (FP64) and t = e / fi0 (FP32), each pixel's change
// 2 (fi1 - fi0) - 2 x ln(fi1 / fi0) is 2 (e (fi0 - x) / fi0 + x h(t)), h(t) =
// t - ln(1 + t).  Summed, it carries the accuracy of FP32 relative to the
// change itself -- not to the two totals, whose FP32 rounding is larger
// than the 1e-6 relative tolerance the turn compares the change with.
// Same fi <= 0 stop and cached values as eval_error.
template <int PPL>
__device__ int eval_delta(const Model& m, RoiM<PPL>& c, double& delta, int& neg_fi) {
  const int l = lane_id();
  double sum = 0.0;
  bool stopped = false;
#pragma unroll
  for (int s = 0; s < PPL; s++) {
    if (!stopped) {
      const bool o = c.ok(s);
      const double xk = o ? c.x(s) : 1.0;
      const double dd1 = o ? fi_minus_x(c, s) : 0.0;
      const float dd = (float)dd1;
      const float xf = (float)xk;
      const float fi = xf + dd;  // only for the fi <= 0 test
      const unsigned long long bad = __ballot(o && fi <= 0.0f);
      const int lim = bad ? __builtin_ctzll(bad) : kWave;
      if (o && l < lim) {
        const double d0 = c.d0(s);
        const float e = (float)(dd1 - d0);
        const float f0 = (float)(xk + d0);
        const float r0 = __builtin_amdgcn_rcpf(f0);
        const float t = e * r0;
        sum += (double)(2.0f * (e * ((float)d0 * r0) + xf * mle_h(t)));
        c.dfi[s] = dd;
        c.stale &= ~(1u << s);
      }
      if (bad) stopped = true;
    }
  }
  if (stopped) {
    neg_fi += 1;
    return 1;
  }
  delta = wsum(sum);
  return 0;
}

// daoCalcJH* for MLE (dao_fit.c:48-958) in FP32; results to the slice's J/H
// area in FP64.  The shape table must hold st's.
template <int PPL, int SHAPE, int EM = -1>
__device__ void calc_jh(const Model& m, RoiM<PPL>& c, const PeakState& st, const double*) {
  constexpr int N = SHAPE == SH_2DFIXED ? 4 : (SHAPE == SH_3D ? 6 : 5);
  const float a1 = (float)st.prm[P_HEIGHT];
  const float wx = (float)st.prm[P_XW];
  const float wy = (SHAPE == SH_3D || SHAPE == SH_Z) ? (float)st.prm[P_YW] : wx;
  double gxd = 0.0, gyd = 0.0;
  if (SHAPE == SH_Z) z_grad(m, st, gxd, gyd);
  const float gx = (float)gxd, gy = (float)gyd;
  const double2* tb = c.tab();

  double j_acc[N];  // J sums in FP64: the fit stops where J ~ 0
  float h_acc[N * (N + 1) / 2];
#pragma unroll
  for (int a = 0; a < N; a++) j_acc[a] = 0.0;
#pragma unroll
  for (int a = 0; a < N * (N + 1) / 2; a++) h_acc[a] = 0.0f;

#pragma unroll
  for (int s = 0; s < PPL; s++) {
    if (c.ok(s)) {
      const double2 cx = tb[c.rx(s)];
      const double2 cy = tb[32 + c.ry(s)];
      const float xt = (float)cx.x, ex = (float)cx.y, yt = (float)cy.x, ey = (float)cy.y;
      const float dd = c.dfi[s];  // fi - x
      const float xk = (float)c.x(s);
      const float fi = xk + dd;
      const float rq = c.rqe(s);
      const float et = ex * ey;
      float g[N];
      if (SHAPE == SH_2DFIXED) {
        g[0] = et;
        g[1] = 2.0f * a1 * wx * xt * et;
        g[2] = 2.0f * a1 * wy * yt * et;
        g[3] = 1.0f;
      } else if (SHAPE == SH_2D) {
        g[0] = et;
        g[1] = 2.0f * a1 * wx * xt * et;
        g[2] = 2.0f * a1 * wy * yt * et;
        g[3] = -a1 * xt * xt * et - a1 * yt * yt * et;
        g[4] = 1.0f;
      } else if (SHAPE == SH_3D) {
        g[0] = et;
        g[1] = 2.0f * a1 * wx * xt * et;
        g[2] = -a1 * xt * xt * et;
        g[3] = 2.0f * a1 * wy * yt * et;
        g[4] = -a1 * yt * yt * et;
        g[5] = 1.0f;
      } else {
        g[0] = et;
        g[1] = 2.0f * a1 * wx * xt * et;
        g[2] = 2.0f * a1 * wy * yt * et;
        g[3] = -a1 * xt * xt * gx * et - a1 * yt * yt * gy * et;
        g[4] = 1.0f;
      }
#pragma unroll
      for (int a = 0; a < N; a++) g[a] *= rq;
      // 2 (1 - x/fi) as 2 (fi - x)/fi: no cancellation where fi ~ x, which
      // is where J -> 0 and its rounding would steer the steps
      const float rf = __builtin_amdgcn_rcpf(fi);
      const float wj = 2.0f * dd * rf;
      const float wh = 2.0f * xk * rf * rf;
      int q = 0;
#pragma unroll
      for (int a = 0; a < N; a++) {
        j_acc[a] += (double)(wj * g[a]);
#pragma unroll
        for (int b = a; b < N; b++) h_acc[q++] += wh * g[a] * g[b];
      }
    }
  }
  wsum_many(j_acc, c.red(), c.hj());
  wsumf_many(h_acc, c.red(), c.hj() + N);
}

template <int N>
__device__ constexpr int upk(int r, int c) {  // packed index of (r, c), r <= c
  return r * N - r * (r - 1) / 2 + (c - r);
}

// mFitSolve (multi_fit.c:1839-1883): LAPACK dposv "Lower" on row-major
// storage = Cholesky of the symmetric matrix given by the upper triangle,
// whose diagonal is scaled by (1 + lambda) (multi_fit.c:1174-1185).
// Non-zero on a pivot that is not > 0 (or NaN) or a NaN in the solution.
// Branch-free: H (in LDS) is read into registers at once, and a failed pivot
// only sets the result (the columns after it compute values nobody uses);
// per-column early exits made every LDS read wait for the branch before it.
template <int N>
__device__ inline int chol_solve(const double* H, double lam, double* b) {
  constexpr int NH = N * (N + 1) / 2;
  double h[NH];
#pragma unroll
  for (int q = 0; q < NH; q++) h[q] = H[q];
  double L[N][N];
  double inv[N];  // 1 / L[j][j]
  bool ok = true;
#pragma unroll
  for (int j = 0; j < N; j++) {
    double s = (1.0 + lam) * h[upk<N>(j, j)];
#pragma unroll
    for (int k = 0; k < j; k++) s -= L[j][k] * L[j][k];
    ok = ok && (s > 0.0);
    const double d = sqrt(s);
    inv[j] = 1.0 / d;
#pragma unroll
    for (int i = j + 1; i < N; i++) {
      double t = h[upk<N>(j, i)];
#pragma unroll
      for (int k = 0; k < j; k++) t -= L[i][k] * L[j][k];
      L[i][j] = t * inv[j];
    }
  }
  if (!ok) return 1;
#pragma unroll
  for (int i = 0; i < N; i++) {
    double t = b[i];
#pragma unroll
    for (int k = 0; k < i; k++) t -= L[i][k] * b[k];
    b[i] = t * inv[i];
  }
#pragma unroll
  for (int i = N - 1; i >= 0; i--) {
    double t = b[i];
#pragma unroll
    for (int k = i + 1; k < N; k++) t -= L[k][i] * b[k];
    b[i] = t * inv[i];
  }
#pragma unroll
  for (int i = 0; i < N; i++)
    if (isnan(b[i])) return 1;
  return 0;
}

// mFitUpdate (multi_fit.c:1933-1946).
__device__ inline void roi_follow(PeakState& s) {
  double dx = s.prm[P_XC] - s.xi;
  if ((dx < -0.5) || (dx > 1.5)) s.xi = (int)floor(s.prm[P_XC]);
  double dy = s.prm[P_YC] - s.yi;
  if ((dy < -0.5) || (dy > 1.5)) s.yi = (int)floor(s.prm[P_YC]);
}

// daoUpdate2DFixed/2D/3D/Z (dao_fit.c:1777-1880).
template <int SHAPE>
__device__ inline void apply_update(const Model& m, PeakState& s, const double* d) {
  if (SHAPE == SH_2DFIXED) {
    s.prm[P_HEIGHT] -= d[0];
    s.prm[P_XC] -= d[1];
    s.prm[P_YC] -= d[2];
    s.prm[P_BG] -= d[3];
    roi_follow(s);
  } else if (SHAPE == SH_2D) {
    s.prm[P_HEIGHT] -= d[0];
    s.prm[P_XC] -= d[1];
    s.prm[P_YC] -= d[2];
    s.prm[P_XW] -= d[3];
    s.prm[P_BG] -= d[4];
    s.prm[P_YW] = s.prm[P_XW];
    roi_follow(s);
  } else if (SHAPE == SH_3D) {
    s.prm[P_HEIGHT] -= d[0];
    s.prm[P_XC] -= d[1];
    s.prm[P_XW] -= d[2];
    s.prm[P_YC] -= d[3];
    s.prm[P_YW] -= d[4];
    s.prm[P_BG] -= d[5];
    roi_follow(s);
  } else {
    s.prm[P_HEIGHT] -= d[0];
    s.prm[P_XC] -= d[1];
    s.prm[P_YC] -= d[2];
    s.prm[P_ZC] -= d[3];
    s.prm[P_BG] -= d[4];
    if (s.prm[P_ZC] < m.zmin) s.prm[P_ZC] = m.zmin;
    if (s.prm[P_ZC] > m.zmax) s.prm[P_ZC] = m.zmax;
    roi_follow(s);
    widths_from_z(m, s);
  }
}

// ROI origin inside the frame (every pixel of the ROI addressable).  The
// reference never checks this before touching pixels; a peak whose origin is
// off-image (e.g. an ERROR peak set back to RUNNING by the caller) would read
// outside its arrays there, and is turned into an ERROR peak here instead.
__device__ inline bool origin_ok(const Model& m, const Frame& fr, int xi, int yi) {
  return xi >= 0 && yi >= 0 && xi + m.roi <= fr.W && yi + m.roi <= fr.H;
}

// mFitCheck (multi_fit.c:470-502).
__device__ inline int margin_height_check(const Model& m, const Frame& fr, const PeakState& s, int* d) {
  if ((s.xi < 0) || (s.xi >= (fr.W - m.roi)) || (s.yi < 0) || (s.yi >= (fr.H - m.roi))) {
    d[C_MARGIN]++;
    return 1;
  }
  if (s.prm[P_HEIGHT] < 0.0) {
    d[C_NEG_HEIGHT]++;
    return 1;
  }
  return 0;
}

// daoCheck2DFixed/2D/3D/Z (dao_fit.c:1047-1133): 2D and 3D clamp the widths.
template <int SHAPE>
__device__ inline int peak_check(const Model& m, const Frame& fr, PeakState& s, int* d) {
  if (margin_height_check(m, fr, s, d)) return 1;
  if (SHAPE == SH_2D || SHAPE == SH_3D) {
    if (s.prm[P_XW] < m.wmin) s.prm[P_XW] = m.wmin;
    else if (s.prm[P_XW] > m.wmax) s.prm[P_XW] = m.wmax;
  }
  if (SHAPE == SH_3D) {
    if (s.prm[P_YW] < m.wmin) s.prm[P_YW] = m.wmin;
    else if (s.prm[P_YW] > m.wmax) s.prm[P_YW] = m.wmax;
  }
  return 0;
}

// mFitResetPeak (multi_fit.c:1788-1798): back to the turn's start (still in
// the peak's record), keeping `added`, lambda times 4.
template <class M = MemL1>
__device__ inline void reset_from(PeakState& w, const PeakRec* r) {
  int added = w.added;
  double lam = w.lam;
  load_state<M>(r, w);
  w.added = added;
  w.lam = lam * 4.0;
}

// ---------------------------------------------------------------------------
// the operations
// ---------------------------------------------------------------------------

// One Levenberg-Marquardt turn of peak p: the body of the loop at
// multi_fit.c:1089-1345.  The turn may touch only pixels inside `box`.
// Returns 1 when the turn completed, 2 when a candidate ROI left the box (or
// the journal is full): the frame is then as before the turn (restored from
// the journal) and the peak record is untouched.  d[] receives the turn's
// counter increments.  With jr, the turn journals every ROI position it
// loads into jr (nsnap snapshots, see kJournalSnaps; the first one into
// `stage` instead when given), so that it can be undone exactly; a turn that
// cannot be undone or leave its box (its box is the whole frame) needs none.
template <int PPL, int SHAPE, class R = Roi<PPL>, class M = MemL1, int EM = -1>
__device__ int lm_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr, const int* roi_tab,
                       double* slice, int p, int4 box, int* d, double* jr, int& nsnap, double* stage = nullptr,
                       const RoiGeo<PPL>* geo = nullptr) {
  constexpr int N = SHAPE == SH_2DFIXED ? 4 : (SHAPE == SH_3D ? 6 : 5);
  PeakRec* rec = pk.rec + p;
  nsnap = 0;
  // The record keeps the turn's starting state until the turn ends, so
  // resets read it back from there instead of holding a second copy.
  PeakState w;
  load_state<M>(rec, w);
  if (!origin_ok(m, fr, w.xi, w.yi)) {
    w.status = ST_ERROR;
    store_state<M>(rec, w);
    return 1;
  }

  R c;
  if (geo)
    c.init(*geo, slice);
  else
    c.init(m, roi_tab, fr, slice);
  if constexpr (std::is_same<M, MemL1>::value) {
    c.template load<M>(px, fr, w.xi, w.yi, (jr && stage) ? stage : jr);
  } else if (jr && stage) {
    c.template load<M, MemL1>(px, fr, w.xi, w.yi, stage);  // the first snapshot into LDS
  } else {
    c.template load<M>(px, fr, w.xi, w.yi, jr);
  }
  if (jr) nsnap = 1;
  c.shape(m, w);

  // mixed precision judges a candidate at the start's ROI origin on the
  // deviance change (eval_delta) instead of two rounded totals
  constexpr bool kDelta = std::is_same<R, RoiM<PPL>>::value;
  double wJ[N];
  double e;
  if (eval_error<PPL, EM>(m, c, px.asx, e, d[C_NEG_FI]) == 0) {
    w.err = e;
    if constexpr (kDelta) keep_start<PPL>(c);
  }
  const double start = w.err;
  calc_jh<PPL, SHAPE, EM>(m, c, w, px.asx);
  const double* J = c.hj();
  const double* H = c.hj() + N;
  c.stamp(w, -1);
  w.added--;

  bool escaped = false;
  // Every retry multiplies lambda by 4 from >= 7.5e-4, so more than 60
  // cycles cannot happen; the cap only guarantees termination.
  for (int guard = 0;; guard++) {
    if (w.lam > 1.0e20 || guard > 200) {
      d[C_LOST]++;
      w.status = ST_ERROR;
      break;
    }
    w.iter++;
    d[C_ITER]++;
    w.status = ST_RUNNING;

#pragma unroll
    for (int r = 0; r < N; r++) wJ[r] = J[r];
    if (chol_solve<N>(H, w.lam, wJ) != 0) {
      d[C_DPOSV]++;
      w.status = ST_ERROR;
      w.lam = w.lam * 4.0;
      continue;
    }

    apply_update<SHAPE>(m, w, wJ);
    if (peak_check<SHAPE>(m, fr, w, d)) {
      reset_from<M>(w, rec);
      w.status = ST_ERROR;
      continue;
    }

    if (w.xi < box.x || w.yi < box.y || w.xi + m.roi > box.z || w.yi + m.roi > box.w) {
      escaped = true;
      break;
    }
    double* snap = nullptr;
    if (jr && (w.xi != c.at_x || w.yi != c.at_y)) {
      if (nsnap == kJournalSnaps) {
        escaped = true;
        break;
      }
      snap = jr + nsnap * journal_snap_words(PPL);
      nsnap++;
    }

    c.template move_to<M>(px, fr, w.xi, w.yi, snap);
    c.shape(m, w);
    c.stamp(w, +1);
    w.added++;

    // dl = candidate error - start error (for FP64 exactly that difference)
    double dl = 0.0;
    int bad_e;
    if constexpr (kDelta) {
      if (c.d0ok && c.at_x == c.d0x && c.at_y == c.d0y) {
        bad_e = eval_delta<PPL>(m, c, dl, d[C_NEG_FI]);
        e = start + dl;
      } else {
        bad_e = eval_error<PPL, EM>(m, c, px.asx, e, d[C_NEG_FI]);
        dl = e - start;
      }
    } else {
      bad_e = eval_error<PPL, EM>(m, c, px.asx, e, d[C_NEG_FI]);
      dl = e - start;
    }
    if (bad_e != 0) {
      c.stamp(w, -1);
      w.added--;
      reset_from<M>(w, rec);
      w.status = ST_ERROR;
      continue;
    }
    w.err = e;

    if (dl > 0.0) {
      if ((dl / start) < m.tol) {
        w.status = ST_CONVERGED;
        break;
      }
      d[C_NON_DECR]++;
      c.stamp(w, -1);
      w.added--;
      reset_from<M>(w, rec);
    } else {
      if (((-dl) / start) < m.tol) {
        w.status = ST_CONVERGED;
        break;
      }
      if (w.lam > 1.0e-3) w.lam = 0.75 * w.lam;
      break;
    }
  }

  if (escaped) {
    // Put the frame back: re-add the turn's starting contribution (the
    // escape is detected before the candidate touches any pixel), then
    // restore the exact state of every pixel the turn touched from its
    // journal.
    load_state<M>(rec, w);
    c.template move_to<M>(px, fr, w.xi, w.yi);
    c.shape(m, w);
    c.stamp(w, +1);
    c.template store<M>(px);
    if (jr) journal_unwind<PPL, M>(px, fr, c, jr, nsnap, stage);
#pragma unroll
    for (int q = 0; q < C_COUNT; q++) d[q] = 0;
    nsnap = 0;
    return 2;
  }

  c.template store<M>(px);
  store_state<M>(rec, w);
  return 1;
}

// mFitRecenterPeaks body (multi_fit.c:1656-1688) for peak p, whose centre
// left [-0.25, 1.25] of its ROI origin.
template <int PPL, class R = Roi<PPL>, class M = MemL1>
__device__ void recenter_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr,
                              const int* roi_tab, double* slice, int p, int* d) {
  PeakRec* rec = pk.rec + p;
  PeakState w;
  load_state<M>(rec, w);
  if (!origin_ok(m, fr, w.xi, w.yi)) {
    w.status = ST_ERROR;
    store_state<M>(rec, w);
    return;
  }
  R c;
  c.init(m, roi_tab, fr, slice);
  c.template load<M>(px, fr, w.xi, w.yi);
  c.shape(m, w);
  c.stamp(w, -1);
  w.added--;
  w.xi = (int)floor(w.prm[P_XC]);
  w.yi = (int)floor(w.prm[P_YC]);
  if (margin_height_check(m, fr, w, d)) {
    d[C_LOST] += 1;
    w.status = ST_ERROR;
  } else {
    c.template move_to<M>(px, fr, w.xi, w.yi);
    c.shape(m, w);
    c.stamp(w, +1);
    w.added++;
  }
  c.template store<M>(px);
  store_state<M>(rec, w);
}

// Add (sign +1) or remove (-1) `added` copies of state s's contribution with
// atomic adds (undos; the callers order overlapping ones).  With counts_only
// just the covering counts change (the caller restores the rest from the
// turn's journal).
template <int PPL, class M = MemL1>
__device__ void undo_stamp(const Model& m, const Pix& px, const Frame& fr, const Roi<PPL>& c,
                           const PeakState& s, int sign, bool counts_only) {
  const int n = s.added * sign;
  if (n == 0 || !origin_ok(m, fr, s.xi, s.yi)) return;
  double t, e;
  lane_shape(m, s, t, e);
  const long long base = fr.pix0 + (long long)s.yi * fr.W + s.xi;
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    double ex = __shfl(e, c.rx(q), kWave);
    double ey = __shfl(e, 32 + c.ry(q), kWave);
    if (c.ok(q)) {
      const long long k = base + c.koff[q];
      atomicAdd(&px.rw[k].cnt, n);
      if (!counts_only) {
        const PixRO r = px.ro[k];
        atomicAdd(&px.rw[k].f, (double)n * (s.prm[P_HEIGHT] * (ey * ex) * r.rqe));
        atomicAdd(&px.rw[k].bg, (double)n * (s.prm[P_BG] * r.rqe + r.scm));
        M::st(&px.rw[k].stale, 1);
      }
    }
  }
}

// Undo a completed LM turn of peak p: the image loses the peak's current
// contribution and regains the one it had at the sweep's start (backup
// record), and the peak returns to the backup.  With the turn's journal (jr,
// nsnap snapshots) the pixels the turn touched return exactly to their state
// before it (the counts by arithmetic); without, the contributions are
// subtracted and added back and the touched pixels marked stale.
template <int PPL, class M = MemL1>
__device__ void undo_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr, const int* roi_tab,
                          double* slice, int p, const double* jr, int nsnap) {
  PeakState w, s0;
  load_state<M>(pk.rec + p, w);
  load_state<M>(pk.bk + p, s0);
  Roi<PPL> c;
  c.init(m, roi_tab, fr, slice);
  const bool journaled = jr != nullptr && nsnap > 0;
  undo_stamp<PPL, M>(m, px, fr, c, w, -1, journaled);
  undo_stamp<PPL, M>(m, px, fr, c, s0, +1, journaled);
  store_state<M>(pk.rec + p, s0);
  if (journaled) journal_unwind<PPL, M>(px, fr, c, jr, nsnap);
}

// daoNewPeaks body (dao_fit.c:1605-1751) for one new peak from a row of 4
// ({x,y,z,sigma}) or 7 ({x,y,z,bg,h,sx,sy}) doubles.
template <int PPL>
__device__ void newpeak_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr, const int* roi_tab,
                             double* slice, int p, const double* q, int fmt, int row, int* d) {
  PeakState w;
#pragma unroll
  for (int k = 0; k < kNFit; k++) w.prm[k] = 0.0;
  w.added = 0;
  w.iter = 0;
  w.status = ST_RUNNING;
  w.err = 0.0;
  w.lam = 1.0;
  w.prm[P_XC] = q[0] - m.xoff;
  w.prm[P_YC] = q[1] - m.xoff;
  w.prm[P_ZC] = q[2] - m.zoff;
  if (fmt == PF_TEXT) {
    w.prm[P_BG] = q[3];
    w.prm[P_HEIGHT] = q[4];
    w.prm[P_XW] = 1.0 / (2.0 * q[5] * q[5]);
    w.prm[P_YW] = 1.0 / (2.0 * q[6] * q[6]);
    if (m.shape == SH_Z) widths_from_z(m, w);
  } else {
    if (m.shape == SH_Z) {
      widths_from_z(m, w);
    } else {
      double width = 1.0 / (2.0 * q[3] * q[3]);
      w.prm[P_XW] = width;
      w.prm[P_YW] = width;
    }
  }
  w.xi = (int)floor(w.prm[P_XC]);
  w.yi = (int)floor(w.prm[P_YC]);
  if (fmt != PF_TEXT) {
    w.prm[P_BG] = 1.0;
    w.prm[P_HEIGHT] = 1.0;
  }

  int bad;
  switch (m.shape) {
    case SH_2D: bad = peak_check<SH_2D>(m, fr, w, d); break;
    case SH_3D: bad = peak_check<SH_3D>(m, fr, w, d); break;
    default: bad = peak_check<SH_2DFIXED>(m, fr, w, d); break;  // no width clamp
  }
  bool good = bad == 0;
  Roi<PPL> c;
  c.init(m, roi_tab, fr, slice);
  if (good) {
    c.load(px, fr, w.xi, w.yi);
    c.shape(m, w);
    if (fmt != PF_TEXT) {
      // mFitEstimatePeakBackground (multi_fit.c:638-662)
      double sb = 0.0;
#pragma unroll
      for (int s = 0; s < PPL; s++)
        if (c.ok(s)) sb += px.ro[c.k(s)].bgest / c.rqe(s);
      w.prm[P_BG] = wsum(sb) / ((double)m.nroi);
      if (fmt == PF_FINDER) {
        // mFitEstimatePeakHeight (multi_fit.c:678-700)
        double sp = 0.0, sx = 0.0;
#pragma unroll
        for (int s = 0; s < PPL; s++) {
          if (c.ok(s)) {
            double t = c.psf[s] * c.rqe(s);
            sp += t * t * t;
            sx += t * t * (c.x(s) - c.f[s] - px.ro[c.k(s)].bgest);
          }
        }
        w.prm[P_HEIGHT] = wsum(sx) / wsum(sp);
        if (w.prm[P_HEIGHT] < m.min_height) w.prm[P_HEIGHT] = m.min_height;
        if (w.prm[P_HEIGHT] <= 0.0) {
          if (lane_id() == 0) printf("Warning peak %d has negative estimated height!\n", row);
          good = false;
        }
      }
    }
  }
  if (good) {
    c.stamp(w, +1);
    w.added++;
    c.store(px);
  } else {
    w.status = ST_ERROR;
  }
  store_state(pk.rec + p, w);
}

// Initial error of a new peak (dao_fit.c:1758-1763).
template <int PPL>
__device__ void peakerror_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr,
                               const int* roi_tab, double* slice, int p, int* d) {
  PeakState w;
  load_state(pk.rec + p, w);
  if (w.status == ST_RUNNING && origin_ok(m, fr, w.xi, w.yi)) {
    Roi<PPL> c;
    c.init(m, roi_tab, fr, slice);
    c.load(px, fr, w.xi, w.yi);
    double e;
    if (eval_error<PPL>(m, c, px.asx, e, d[C_NEG_FI]) == 0) w.err = e;
    c.store(px);
    store_state(pk.rec + p, w);
  }
}

// mFitSetPeakStatus (multi_fit.c:1809-1831): subtract a peak marked ERROR.
template <int PPL>
__device__ void subtract_turn(const Model& m, const Pix& px, const Peaks& pk, const Frame& fr,
                              const int* roi_tab, double* slice, int p) {
  PeakState w;
  load_state(pk.rec + p, w);
  if (origin_ok(m, fr, w.xi, w.yi)) {
    Roi<PPL> c;
    c.init(m, roi_tab, fr, slice);
    c.load(px, fr, w.xi, w.yi);
    c.shape(m, w);
    c.stamp(w, -1);
    w.added--;
    c.store(px);
    store_state(pk.rec + p, w);
  }
}

__device__ inline bool boxes_overlap(int4 a, int4 b) {
  return (a.x < b.z) && (b.x < a.z) && (a.y < b.w) && (b.y < a.w);
}

// ---------------------------------------------------------------------------
// frame-per-workgroup fitter
// ---------------------------------------------------------------------------
//
// One workgroup of nw <= NW waves (nw = blockDim.x / 64, chosen at launch)
// runs whole sweeps of one frame (one
// mFitIterateLM, or all of doFit, dao_fit_c.py:332-356) without returning to
// the host.  The running peaks of a sweep are the ops 0..n-1 in index order.
// Ops are handed out in index order, so every op below the one a wave just
// took is finished or held by another wave; op j starts once no other wave
// holds an op i < j whose box overlaps j's.  The wave with the smallest held
// op can always start, so the frame makes progress, and overlapping ops run
// in index order -- the reference's update order (multi_fit.c:1089).

constexpr int kNoEscape = 0x7fffffff;
constexpr int kIdle = 0x7fffffff;  // wave holds no op
constexpr int kPending = -1;       // wave is taking an op (index not known yet)

// Launched with nw * 64 threads (nw <= NW) and nw * kSlice doubles of
// dynamic LDS (fit_frames_lds): fewer waves per frame leave room for more
// workgroups per CU.
template <int PPL, bool MP>
constexpr int fit_frames_slice() {
  return MP ? RoiLdsM<PPL>::kDoubles : RoiLds<PPL>::kDoubles;
}

// k_fit_frames stages each turn's first journal snapshot in the wave's LDS
// (after its slice) when that fits beside eight waves' slices: the snapshot
// reaches HBM only if the turn can still be undone when it ends.
template <int PPL>
constexpr int fit_frames_stage() {
  return PPL <= 4 ? (int)journal_snap_words(PPL) : 0;
}

template <int NW>
struct FpwShared {
  int next;         // next op to hand out
  int esc;          // smallest op whose LM turn left its box
  int cur[NW];      // op held by each wave (kIdle / kPending / index)
  int4 curbox[NW];  // its box
  int wsum[NW];
  int cnt[C_COUNT];
  int failed;       // a wait of this frame gave up: the frame stops
};

// Order-preserving list of the frame's peaks that satisfy `mode`
// (SEL_RUNNING or SEL_RECENTER) into dst[peak0 + k]; returns the count.
// The candidates are src[peak0 + k], k < nsrc (a list in index order), or,
// with src == nullptr, every peak of the frame.  dst may be src: entry k is
// read before any entry at or above the count of earlier matches is written.
// Within doFit a peak that is not RUNNING never becomes RUNNING again, so a
// sweep's RUNNING list is the previous one's filtered, and a sweep costs
// O(RUNNING peaks), not O(all peaks), in the long tail of a fit.
template <class M = MemL1>
__device__ int fpw_list(const FitArgs& a, const Frame& fr, int mode, int* wsums, const int* src, int nsrc,
                        int* dst) {
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const int total = src ? nsrc : fr.nfit;
  int count = 0;  // every thread keeps the same running total
  const int nw = blockDim.x >> 6;
  for (int base = 0; base < total; base += nw * kWave) {
    int t = base + threadIdx.x;
    bool flag = false;
    int p = 0;
    if (t < total) {
      p = src ? M::ld(&src[fr.peak0 + t]) : fr.peak0 + t;
      const PeakRec* r = a.pk.rec + p;
      if (M::ld(&r->status) == ST_RUNNING) {
        if (mode == SEL_RUNNING) {
          flag = true;
        } else {
          double dx = M::ld(&r->prm[P_XC]) - M::ld(&r->xi);
          double dy = M::ld(&r->prm[P_YC]) - M::ld(&r->yi);
          flag = (dx < -0.25) || (dx > 1.25) || (dy < -0.25) || (dy > 1.25);
        }
      }
    }
    unsigned long long mask = __ballot(flag);
    int pre = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wsums[wv] = __popcll(mask);
    __syncthreads();
    int off = count;
    int tot = 0;
    for (int u = 0; u < nw; u++) {
      if (u < wv) off += wsums[u];
      tot += wsums[u];
    }
    if (flag) M::st(&dst[fr.peak0 + off + pre], p);
    count += tot;
    __syncthreads();
  }
  return count;
}

// Journal of op slot `slot`.
template <int PPL>
__device__ inline double* op_journal(const FitArgs& a, int slot) {
  return a.journal + (long long)slot * journal_words(PPL);
}

// Journal for an LM op with box b, or none: an op whose box is the whole
// frame starts when every earlier op has finished and cannot leave its box,
// so it is never undone.
template <int PPL>
__device__ inline double* op_journal(const FitArgs& a, const Frame& fr, int slot, int4 b) {
  const bool whole = b.x == 0 && b.y == 0 && b.z == fr.W && b.w == fr.H;
  return (a.journal == nullptr || whole) ? nullptr : op_journal<PPL>(a, slot);
}

// Wait until op j may start.  Returns false when an earlier op escaped (op j
// must not run now) or the wait bound was hit; then frame f is marked failed
// (frame_err[f] for the host, sh.failed for the workgroup) and every later
// wait of the frame gives up at once.
template <int NW>
__device__ bool fpw_wait(const FitArgs& a, int f, int j, int4 b, FpwShared<NW>& sh) {
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (long long spin = 0;; spin++) {
    bool blocked = false;
    if (lane < (int)(blockDim.x >> 6) && lane != wv) {
      int c = __hip_atomic_load(&sh.cur[lane], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (c == kPending) {
        blocked = true;
      } else if (c < j) {
        blocked = boxes_overlap(b, sh.curbox[lane]);
      }
    }
    if (!__any(blocked)) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      return true;
    }
    if (__hip_atomic_load(&sh.esc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < j) return false;
    if (spin > a.spin_limit || __hip_atomic_load(&sh.failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
      if (lane == 0) {
        __hip_atomic_store(&sh.failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        atomicExch(&a.frame_err[f], 1);
      }
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__device__ inline void fpw_set_cur(int* cur, int v) {
  if (lane_id() == 0) __hip_atomic_store(cur, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Run ops 0..n-1 (kind OP_LM or OP_RECENTER; op k is peak list[peak0 + k]) of the frame.  If an LM turn
// leaves its box, every op below it has finished (ops are handed out in order
// and never wait on later ones): the finished ops above it are undone newest
// first, the turn gets the whole frame as its box (it then runs alone and
// cannot escape again), and the run resumes from it.
//
// OP_NEWPEAK ops (daoNewPeaks) take op k's start row from rows (fmt's
// layout, row k), OP_PEAKERROR ops compute a new peak's first error; neither
// moves out of its box.
template <int PPL, int SHAPE, int NW, bool MP, int EM, int kind>
__device__ void fpw_run(const FitArgs& a, int f, const Frame& fr, int n, const int* list, FpwShared<NW>& sh,
                        double* slice, const double* rows = nullptr, int fmt = 0) {
  using R = typename std::conditional<MP, RoiM<PPL>, Roi<PPL>>::type;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (int k = threadIdx.x; k < n; k += blockDim.x) a.op_done[fr.peak0 + k] = 0;
  RoiGeo<PPL> geo;
  if constexpr (kind == OP_LM) geo.make(a.m, a.roi_tab, fr);
  int from = 0;
  for (;;) {
    if (threadIdx.x < (blockDim.x >> 6)) sh.cur[threadIdx.x] = kIdle;
    if (threadIdx.x == 0) {
      sh.esc = kNoEscape;
      sh.next = from;
    }
    __syncthreads();
    for (;;) {
      if (__hip_atomic_load(&sh.esc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != kNoEscape) break;
      fpw_set_cur(&sh.cur[wv], kPending);
      int j = 0;
      if (lane == 0) j = atomicAdd(&sh.next, 1);
      j = uniform_i(j);
      if (j >= n) {
        fpw_set_cur(&sh.cur[wv], kIdle);
        break;
      }
      const int slot = fr.peak0 + j;
      // box and peak index in one round trip
      const int4 b = uniform_i4(a.box[slot]);
      const int p = uniform_i(list[slot]);
      if (lane == 0) sh.curbox[wv] = b;
      fpw_set_cur(&sh.cur[wv], j);
      if (!fpw_wait<NW>(a, f, j, b, sh)) {
        fpw_set_cur(&sh.cur[wv], kIdle);
        break;
      }
      int d[C_COUNT];
#pragma unroll
      for (int q = 0; q < C_COUNT; q++) d[q] = 0;
      int r = 1, nsnap = 0;
      double* const jr = kind == OP_LM ? op_journal<PPL>(a, fr, slot, b) : nullptr;
      double* const stage = fit_frames_stage<PPL>() > 0 ? slice + fit_frames_slice<PPL, MP>() : nullptr;
      if constexpr (kind == OP_LM) {
        r = lm_turn<PPL, SHAPE, R, MemL1, EM>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, b, d, jr, nsnap, stage, &geo);
      } else if constexpr (kind == OP_RECENTER) {
        recenter_turn<PPL, R>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, d);
      } else if constexpr (kind == OP_NEWPEAK) {
        newpeak_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, rows + (long long)j * (fmt == PF_TEXT ? 7 : 4),
                          fmt, j, d);
      } else {
        peakerror_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, d);
      }
      if (r == 2) {
        if (lane == 0) atomicMin(&sh.esc, j);
        fpw_set_cur(&sh.cur[wv], kIdle);
        break;
      }
      int done = 1 + nsnap;
      bool copy = false;
      if (jr && stage) {
        // The staged snapshot goes to the op's journal if an earlier op is
        // still running or has left its box (the op may then be undone);
        // otherwise every earlier op has finished for good and this one
        // stays.  An earlier op that escapes sets esc before it leaves its
        // seat, so a seat read as free means esc is already visible.
        bool open = false;
        if (lane < (int)(blockDim.x >> 6) && lane != wv) {
          const int c = __hip_atomic_load(&sh.cur[lane], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
          open = c >= 0 && c < j;
        }
        copy = __any(open) || __hip_atomic_load(&sh.esc, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) < j;
        if (!copy) done = 1;  // never undone: no journal
      }
      if (lane == 0) {
#pragma unroll
        for (int q = 0; q < C_COUNT; q++) {
          a.op_delta[(long long)slot * C_COUNT + q] = d[q];
          if (d[q]) atomicAdd(&sh.cnt[q], d[q]);
        }
        a.op_done[slot] = done;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      fpw_set_cur(&sh.cur[wv], kIdle);
      // The journal is read only by a rollback, after the workgroup's
      // barrier: the copy need not hold back the ops waiting for this one.
      if (copy)
        for (int q = lane; q < (int)journal_snap_words(PPL); q += kWave) MemL1::st_stream(jr + q, stage[q]);
    }
    __syncthreads();
    const int e = sh.esc;
    if (e == kNoEscape) break;
    if (wv == 0) {
      for (int k = n - 1; k > e; k--) {
        const int slot = fr.peak0 + k;
        const int done = a.op_done[slot];
        if (done) {
          undo_turn<PPL>(a.m, a.px, a.pk, fr, a.roi_tab, slice, list[slot], op_journal<PPL>(a, slot), done - 1);
          if (lane == 0) {
            for (int q = 0; q < C_COUNT; q++) sh.cnt[q] -= a.op_delta[(long long)slot * C_COUNT + q];
            a.op_done[slot] = 0;
          }
        }
      }
      if (lane == 0) a.box[fr.peak0 + e] = make_int4(0, 0, fr.W, fr.H);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    }
    __syncthreads();
    from = e;
  }
}

// MultiFitter.doFit (dao_fit_c.py:332-356) of frame f by the calling
// workgroup: sweeps (mFitIterateLM, multi_fit.c:1072-1349) and recentering
// passes until no peak is RUNNING or a.max_more sweeps followed the first
// (a.max_more = 0: exactly one sweep).  Counter increments accumulate in
// sh.cnt; sh.failed is set if a scheduler wait gave up.  Returns the sweeps
// run.
//
// With n0 >= 0 the fit resumes where a team handed it off: iteration it0,
// whose list of n0 RUNNING peaks is already in run_list (built and checked
// against max_more as the loop below would).
template <int PPL, int SHAPE, int NW, bool MP, int EM = -1>
__device__ int fit_frame(const FitArgs& a, int f, const Frame& fr, FpwShared<NW>& sh, double* slice, int it0 = 0,
                         int n0 = -1) {
  const Model& m = a.m;
  int sweeps = 0;
  int n = n0 > 0 ? n0 : 0;  // RUNNING peaks at the sweep's start: a.run_list[peak0 .. peak0 + n)
  for (int it = it0;; it++) {
    // a wait gave up (every wave passed a barrier since): the frame's state
    // is not the sequential one, so it stops and the host reports it
    if (sh.failed) break;
    if (n0 < 0 || it > it0) {
      n = it == 0 ? fpw_list(a, fr, SEL_RUNNING, sh.wsum, nullptr, 0, a.run_list)
                  : fpw_list(a, fr, SEL_RUNNING, sh.wsum, a.run_list, n, a.run_list);
      if (it > 0 && (n == 0 || it > a.max_more)) break;
    }
    if (n == 0) continue;  // the unconditional first sweep found nothing
    sweeps++;
    // reserved boxes (ROI at the sweep's start + margin) and backups
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const int slot = fr.peak0 + k;
      const int p = a.run_list[slot];
      const PeakRec r = a.pk.rec[p];
      int4 b;
      b.x = max(0, r.xi - m.margin);
      b.y = max(0, r.yi - m.margin);
      b.z = min(fr.W, r.xi + m.roi + m.margin);
      b.w = min(fr.H, r.yi + m.roi + m.margin);
      a.box[slot] = b;
      a.pk.bk[p] = r;
    }
    __syncthreads();
    fpw_run<PPL, SHAPE, NW, MP, EM, OP_LM>(a, f, fr, n, a.run_list, sh, slice);
    // mFitRecenterPeaks (multi_fit.c:1644-1690)
    const int nr = fpw_list(a, fr, SEL_RECENTER, sh.wsum, a.run_list, n, a.op_peak);
    if (nr > 0) {
      for (int k = threadIdx.x; k < nr; k += blockDim.x) {
        const int slot = fr.peak0 + k;
        const PeakRec r = a.pk.rec[a.op_peak[slot]];
        int nx = (int)floor(r.prm[P_XC]);
        int ny = (int)floor(r.prm[P_YC]);
        int4 b;
        b.x = max(0, min(fr.W, min(r.xi, nx)));
        b.y = max(0, min(fr.H, min(r.yi, ny)));
        b.z = max(b.x, min(fr.W, max(r.xi, nx) + m.roi));
        b.w = max(b.y, min(fr.H, max(r.yi, ny) + m.roi));
        a.box[slot] = b;
      }
      __syncthreads();
      fpw_run<PPL, SHAPE, NW, MP, EM, OP_RECENTER>(a, f, fr, nr, a.op_peak, sh, slice);
    }
  }
  __syncthreads();
  return sweeps;
}

// RESUME: the launch follows k_fit_team on the stream and continues only
// the frames their team handed off (a separate instance, so that the
// ordinary one compiles as if resuming did not exist).
template <int PPL, int SHAPE, int NW, bool MP, int EM = -1, bool RESUME = false>
__global__ __launch_bounds__(NW * 64) void k_fit_frames(FitArgs a) {
  constexpr int kSlice = fit_frames_slice<PPL, MP>() + fit_frames_stage<PPL>();
  __shared__ FpwShared<NW> sh;
  extern __shared__ double lds_dyn[];
  double* const slice = lds_dyn + (threadIdx.x >> 6) * kSlice;
  const int f = a.frame_list[blockIdx.x];
  const Frame fr = a.frames[f];
  int it0 = 0, n0 = -1, sweeps0 = 0;
  if constexpr (RESUME) {
    const TeamCtl* ctl = a.team_ctl + f;
    if (ctl->kind != kTeamHandoff) return;
    it0 = ctl->it;
    n0 = ctl->nrun;
    sweeps0 = ctl->sweeps;
  }
  if (threadIdx.x < C_COUNT) sh.cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) sh.failed = 0;
  __syncthreads();
  const int sweeps = fit_frame<PPL, SHAPE, NW, MP, EM>(a, f, fr, sh, slice, it0, n0);
  if (threadIdx.x < C_COUNT && sh.cnt[threadIdx.x] != 0)
    atomicAdd(&a.frame_counters[f * C_COUNT + threadIdx.x], sh.cnt[threadIdx.x]);
  if (threadIdx.x == 0) a.frame_sweeps[f] = sweeps0 + sweeps;
}

// The ROI box at origin (x0, y0), clipped to the frame.
__device__ inline int4 box_at(int x0, int y0, int roi, const Frame& fr) {
  int4 b;
  b.x = max(0, min(fr.W, x0));
  b.y = max(0, min(fr.H, y0));
  b.z = max(b.x, min(fr.W, x0 + roi));
  b.w = max(b.y, min(fr.H, y0 + roi));
  return b;
}

// daoNewPeaks' image work (dao_fit.c:1582-1766) for n start rows of frame f
// by the calling workgroup: op k (box: the ROI at the floor of row k's
// position) adds peak list[peak0 + k] from row k, in index order where ROIs
// overlap (the ordered executor of the frame fitter), then every new peak's
// first error is computed (dao_fit.c:1758-1763; box: its ROI).  Counter
// increments accumulate in sh.cnt.
template <int PPL, int NW>
__device__ void add_peaks(const FitArgs& a, int f, const Frame& fr, int n, const double* rows, int fmt, int first,
                          FpwShared<NW>& sh, double* slice) {
  const Model& m = a.m;
  const int wid = fmt == PF_TEXT ? 7 : 4;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int slot = fr.peak0 + k;
    a.op_peak[slot] = first + k;
    const double* row = rows + (long long)k * wid;
    a.box[slot] = box_at((int)floor(row[0] - m.xoff), (int)floor(row[1] - m.xoff), m.roi, fr);
  }
  __syncthreads();
  fpw_run<PPL, SH_2DFIXED, NW, false, -1, OP_NEWPEAK>(a, f, fr, n, a.op_peak, sh, slice, rows, fmt);
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const PeakRec* p = a.pk.rec + first + k;
    a.box[fr.peak0 + k] = box_at(p->xi, p->yi, m.roi, fr);
  }
  __syncthreads();
  fpw_run<PPL, SH_2DFIXED, NW, false, -1, OP_PEAKERROR>(a, f, fr, n, a.op_peak, sh, slice);
}

// add_peaks for a batch of frames, one persistent workgroup per job (one
// launch instead of an executor round per dependency level); dynamic LDS as
// k_fit_frames (FP64 slices).
template <int PPL, int NW>
__global__ __launch_bounds__(NW * 64) void k_add_peaks(FitArgs a, const AddPeaksJob* jobs, int fmt) {
  __shared__ FpwShared<NW> sh;
  extern __shared__ double lds_dyn[];
  double* const slice = lds_dyn + (threadIdx.x >> 6) * fit_frames_slice<PPL, false>();
  const AddPeaksJob job = jobs[blockIdx.x];
  const int f = job.frame;
  const Frame fr = a.frames[f];
  if (threadIdx.x < C_COUNT) sh.cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) sh.failed = 0;
  __syncthreads();
  add_peaks<PPL, NW>(a, f, fr, job.n, job.rows, fmt, job.first, sh, slice);
  __syncthreads();
  if (threadIdx.x < C_COUNT && sh.cnt[threadIdx.x] != 0)
    atomicAdd(&a.frame_counters[f * C_COUNT + threadIdx.x], sh.cnt[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// frame-per-team fitter
// ---------------------------------------------------------------------------
//
// A team of workgroups, on any CUs, runs all of doFit for one frame, like
// k_fit_frames does with one workgroup.  The frame's scheduling state is a
// TeamCtl, the run's per-op boxes and finished flags, and one seat per wave
// (inside the run or not) in global memory; every access to it and to the
// frame's pixel fit state and peak records is a MemDev access, ordered by
// dev_order().  The order rule is k_fit_frames': ops are handed out in
// index order (an atomic counter) and op j starts once no earlier op that
// has not finished overlaps j's box, so overlapping ops keep the
// reference's update order.
//
// Between runs (an LM sweep, a recentering pass) one workgroup makes the
// transition: it claims it (compare-and-swap of `lead` from the current
// generation g to g + 1), waits until every seat is idle, builds the next
// run (or rolls back after an escape, or ends the fit) and publishes
// generation g + 1.  A wave marks its seat active before it checks, ahead
// of each op it takes, that no transition out of its generation has been
// claimed; a claimant marks `lead` first and then reads the seats.  Each
// side's store completes before its load is issued, so at least one of them
// sees the other: no op is handed out across a transition.  No workgroup
// ever waits for a team mate that has not started (teams need not be
// co-resident), and the smallest unfinished op can always run, so every
// run ends.

constexpr int kActive = 1;                  // seat of a wave inside team_ops (else kIdle)

struct TeamShared {
  int wsum[16];
  int cnt[C_COUNT];
  int gen;   // generation of the current action
  int act;   // kActRun, kActLead or kActStop
  int kind;  // the run's kind (OP_LM, OP_RECENTER, kTeamDone, kTeamHandoff)
  int n;     // the run's op count
};
constexpr int kActRun = 0, kActLead = 1, kActStop = 2;

// The scheduler's cold paths (team_await, team_quiet, team_transition) are
// inlined into k_fit_team.  Measured on one MI355X with dev_order() as the
// agent-scope fence above: a 2048^2 frame of 20,000 emitters with teams of
// 500 peaks per workgroup fits in 499 ms inlined against 534 ms out of line
// (-DDAO_TEAM_COLD=__noinline__), a 512^2 Z frame of 1000 emitters with
// teams of 64 in 23.9 ms against 26.4 ms; both builds give identical fits.
#ifndef DAO_TEAM_COLD
#define DAO_TEAM_COLD inline
#endif

// Every wait of a team polls at most a.spin_limit times.  A wait that gives
// up marks frame f failed (frame_err[f]); every wait of the frame's team
// then gives up at its next poll and every workgroup of the team stops, so
// the launch ends and the host reports the frame.
__device__ inline bool team_failed(const FitArgs& a, int f) { return uniform_i(MemDev::ld(&a.frame_err[f])) != 0; }
__device__ inline void team_fail(const FitArgs& a, int f) {
  if (lane_id() == 0) MemDev::st(&a.frame_err[f], 1);
}

// Wave 0: wait for a generation other than g, or claim the transition out
// of g once it is due (g == 0: the frame's first run; then: the run's ops
// are all done, or one escaped).  Returns the action; sh.gen its generation.
__device__ DAO_TEAM_COLD int team_await(const FitArgs& a, int f, TeamCtl* ctl, int g, TeamShared& sh) {
  for (long long spin = 0;; spin++) {
    if (team_failed(a, f)) return kActStop;
    const int ph = uniform_i(MemDev::ld(&ctl->phase));
    if (ph != g) {
      dev_order();
      if (lane_id() == 0) sh.gen = ph;
      return kActRun;
    }
    bool due = g == 0;
    if (!due) {
      const int esc = uniform_i(MemDev::ld(&ctl->esc));
      const int done = uniform_i(MemDev::ld(&ctl->done));
      const int n = uniform_i(MemDev::ld(&ctl->n));
      due = esc != kNoEscape || done == n;
    }
    if (due && uniform_i(MemDev::ld(&ctl->lead)) == g) {
      int won = 0;
      if (lane_id() == 0) won = MemDev::cas(&ctl->lead, g, g + 1) == g;
      if (uniform_i(won)) {
        dev_order();
        if (lane_id() == 0) sh.gen = g;
        return kActLead;
      }
    }
    if (spin > a.spin_limit) {
      team_fail(a, f);
      return kActStop;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Wave 0 of the claimant: wait until no seat of the frame holds or is
// taking an op.  False if the frame failed.
__device__ DAO_TEAM_COLD bool team_quiet(const FitArgs& a, int f, int s0, int s1) {
  for (long long spin = 0;; spin++) {
    if (team_failed(a, f)) return false;
    bool busy = false;
    for (int s = s0 + lane_id(); s < s1; s += kWave) busy |= MemDev::ld(&a.team_cur[s]) != kIdle;
    if (!__any(busy)) {
      dev_order();
      return true;
    }
    if (spin > a.spin_limit) {
      team_fail(a, f);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Wait until op j (box b) may start: no earlier op of the run that has not
// finished overlaps b.  Every op below j has been handed out (the counter
// hands them out in order), so the unfinished ones below j are exactly the
// held or waiting ones.  `lo` is this wave's bound below which every op has
// finished; it only grows during a run.  False when an earlier op escaped
// (op j must not run now) or the wait gave up.
__device__ bool team_wait(const FitArgs& a, int f, TeamCtl* ctl, const Frame& fr, int j, int4 b, int& lo) {
  for (long long spin = 0;; spin++) {
    while (lo < j) {
      const int i = lo + lane_id();
      const bool open = i < j && MemDev::ld(&a.op_done[fr.peak0 + i]) == 0;
      const unsigned long long m = __ballot(open);
      if (m) {
        lo += __builtin_ctzll(m);
        break;
      }
      lo = min(lo + kWave, j);
    }
    bool blocked = false;
    for (int i = lo + lane_id(); i < j; i += kWave) {
      const int slot = fr.peak0 + i;
      if (MemDev::ld(&a.op_done[slot]) == 0) blocked |= boxes_overlap(b, MemDev::ld(&a.box[slot]));
    }
    if (!__any(blocked)) {
      dev_order();
      return true;
    }
    if (uniform_i(MemDev::ld(&ctl->esc)) < j) return false;
    if (spin > a.spin_limit || team_failed(a, f)) {
      team_fail(a, f);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// One wave takes and runs ops of generation g until none is left, a
// transition is claimed or an op escapes.  Its seat reads kActive meanwhile.
template <int PPL, int SHAPE, bool MP, int EM = -1>
__device__ void team_ops(const FitArgs& a, int f, const Frame& fr, TeamCtl* ctl, int g, int kind, int n, int seat,
                         TeamShared& sh, double* slice) {
  using R = typename std::conditional<MP, RoiM<PPL>, Roi<PPL>>::type;
  const int lane = lane_id();
  if (lane == 0) MemDev::st(&a.team_cur[seat], kActive);
  dev_order();
  int lo = 0;
  for (;;) {
    if (uniform_i(MemDev::ld(&ctl->lead)) != g) break;
    if (uniform_i(MemDev::ld(&ctl->esc)) != kNoEscape) break;
    if (team_failed(a, f)) break;
    int j = 0;
    if (lane == 0) j = MemDev::add(&ctl->next, 1);
    j = uniform_i(j);
    if (j >= n) break;
    const int slot = fr.peak0 + j;
    const int4 b = uniform_i4(MemDev::ld(&a.box[slot]));
    if (!team_wait(a, f, ctl, fr, j, b, lo)) break;
    const int p = uniform_i(MemDev::ld(&(kind == OP_LM ? a.run_list : a.op_peak)[slot]));
    int d[C_COUNT];
#pragma unroll
    for (int q = 0; q < C_COUNT; q++) d[q] = 0;
    int r = 1, nsnap = 0;
    double* const jr = kind == OP_LM ? op_journal<PPL>(a, fr, slot, b) : nullptr;
    double* const stage = fit_frames_stage<PPL>() > 0 ? slice + fit_frames_slice<PPL, MP>() : nullptr;
    if (kind == OP_LM) {
      r = lm_turn<PPL, SHAPE, R, MemDev, EM>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, b, d, jr, nsnap, stage);
    } else {
      recenter_turn<PPL, R, MemDev>(a.m, a.px, a.pk, fr, a.roi_tab, slice, p, d);
    }
    if (r == 2) {
      if (lane == 0) MemDev::min(&ctl->esc, j);
      break;
    }
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < C_COUNT; q++) {
        MemDev::st(&a.op_delta[(long long)slot * C_COUNT + q], d[q]);
        if (d[q]) atomicAdd(&sh.cnt[q], d[q]);
      }
    }
    dev_order();  // the turn's stores have completed
    if (lane == 0) {
      MemDev::st(&a.op_done[slot], 1 + nsnap);
      MemDev::add(&ctl->done, 1);
    }
    // The staged first snapshot goes to the journal after the op is
    // published: only a rollback reads it, once every seat is idle (a seat
    // is left after dev_order).
    if (jr && stage)
      for (int q = lane; q < (int)journal_snap_words(PPL); q += kWave) MemDev::st(jr + q, stage[q]);
  }
  dev_order();
  if (lane == 0) MemDev::st(&a.team_cur[seat], kIdle);
  dev_order();
}

__device__ inline void team_sync() {
  dev_order();
  __syncthreads();
}

// Reserved boxes (ROI at the sweep's start + margin) and backups of the LM
// run's ops, as in k_fit_frames.
__device__ void team_boxes_lm(const FitArgs& a, const Frame& fr, int n) {
  const Model& m = a.m;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int slot = fr.peak0 + k;
    const int p = MemDev::ld(&a.run_list[slot]);
    const PeakRec* r = a.pk.rec + p;
    const int xi = MemDev::ld(&r->xi), yi = MemDev::ld(&r->yi);
    int4 b;
    b.x = max(0, xi - m.margin);
    b.y = max(0, yi - m.margin);
    b.z = min(fr.W, xi + m.roi + m.margin);
    b.w = min(fr.H, yi + m.roi + m.margin);
    MemDev::st(&a.box[slot], b);
    const long long* src = reinterpret_cast<const long long*>(r);
    long long* dst = reinterpret_cast<long long*>(a.pk.bk + p);
#pragma unroll
    for (int q = 0; q < (int)(sizeof(PeakRec) / 8); q++) MemDev::st(dst + q, MemDev::ld(src + q));
    MemDev::st(&a.op_done[slot], 0);
  }
}

// Boxes of the recentering ops (old and new ROI), as in k_fit_frames.
__device__ void team_boxes_recenter(const FitArgs& a, const Frame& fr, int n) {
  const Model& m = a.m;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int slot = fr.peak0 + k;
    const PeakRec* r = a.pk.rec + MemDev::ld(&a.op_peak[slot]);
    const int xi = MemDev::ld(&r->xi), yi = MemDev::ld(&r->yi);
    const int nx = (int)floor(MemDev::ld(&r->prm[P_XC]));
    const int ny = (int)floor(MemDev::ld(&r->prm[P_YC]));
    int4 b;
    b.x = max(0, min(fr.W, min(xi, nx)));
    b.y = max(0, min(fr.H, min(yi, ny)));
    b.z = max(b.x, min(fr.W, max(xi, nx) + m.roi));
    b.w = max(b.y, min(fr.H, max(yi, ny) + m.roi));
    MemDev::st(&a.box[slot], b);
    MemDev::st(&a.op_done[slot], 0);
  }
}

// The transition out of generation g, by the whole claimant workgroup once
// the frame is quiet: roll back after an escape (as fpw_run), or start the
// next run of doFit's loop (as k_fit_frames), or end the fit; then publish
// generation g + 1.
template <int PPL>
__device__ DAO_TEAM_COLD void team_transition(const FitArgs& a, const Frame& fr, TeamCtl* ctl, int f, int g, TeamShared& sh,
                                double* slice) {
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const int kind = MemDev::ld(&ctl->kind);
  const int n = MemDev::ld(&ctl->n);
  const int esc = MemDev::ld(&ctl->esc);
  int it = MemDev::ld(&ctl->it);
  int sweeps = MemDev::ld(&ctl->sweeps);
  int nrun = MemDev::ld(&ctl->nrun);
  int nk = kind, nn = n, nnext = 0, ndone = 0;
  if (g > 0 && esc != kNoEscape) {
    // every op below esc has finished (ops are handed out in order, and only
    // ops above an escape give theirs up); undo the finished ones above it,
    // newest first, and re-run from esc with the whole frame as its box
    if (wv == 0) {
      for (int k = n - 1; k > esc; k--) {
        const int slot = fr.peak0 + k;
        const int done = uniform_i(MemDev::ld(&a.op_done[slot]));
        if (done) {
          undo_turn<PPL, MemDev>(a.m, a.px, a.pk, fr, a.roi_tab, slice, uniform_i(MemDev::ld(&a.run_list[slot])),
                                 op_journal<PPL>(a, slot), done - 1);
          dev_order();
          if (lane == 0) {
            for (int q = 0; q < C_COUNT; q++) sh.cnt[q] -= MemDev::ld(&a.op_delta[(long long)slot * C_COUNT + q]);
            MemDev::st(&a.op_done[slot], 0);
          }
        }
      }
      if (lane == 0) MemDev::st(&a.box[fr.peak0 + esc], make_int4(0, 0, fr.W, fr.H));
    }
    nnext = esc;
    ndone = esc;
  } else {
    bool lm_next = true;
    if (g > 0 && kind == OP_LM) {
      // mFitRecenterPeaks (multi_fit.c:1644-1690) after the sweep
      const int nr = fpw_list<MemDev>(a, fr, SEL_RECENTER, sh.wsum, a.run_list, nrun, a.op_peak);
      if (nr > 0) {
        team_boxes_recenter(a, fr, nr);
        nk = OP_RECENTER;
        nn = nr;
        lm_next = false;
      }
    }
    if (lm_next) {
      if (g > 0) it++;
      const int nl = g == 0 ? fpw_list<MemDev>(a, fr, SEL_RUNNING, sh.wsum, nullptr, 0, a.run_list)
                            : fpw_list<MemDev>(a, fr, SEL_RUNNING, sh.wsum, a.run_list, nrun, a.run_list);
      nrun = nl;
      if (nl == 0 || it > a.max_more) {
        nk = kTeamDone;
        nn = 0;
      } else if (nl <= a.team_solo) {
        // In a fit's long tail a few peaks stay RUNNING for hundreds of
        // sweeps; through the team each such sweep pays two transitions
        // (claim, quiet, lists, publish, every team mate's poll), about
        // twice what one workgroup's sweep costs.  The team stops here and
        // k_fit_frames resumes the frame from this run (team_resume).
        nk = kTeamHandoff;
        nn = 0;
      } else {
        team_boxes_lm(a, fr, nl);
        nk = OP_LM;
        nn = nl;
        sweeps++;
      }
    }
  }
  team_sync();
  if (threadIdx.x == 0) {
    MemDev::st(&ctl->kind, nk);
    MemDev::st(&ctl->n, nn);
    MemDev::st(&ctl->next, nnext);
    MemDev::st(&ctl->done, ndone);
    MemDev::st(&ctl->esc, kNoEscape);
    MemDev::st(&ctl->it, it);
    MemDev::st(&ctl->sweeps, sweeps);
    MemDev::st(&ctl->nrun, nrun);
    if (nk == kTeamDone) MemDev::st(&a.frame_sweeps[f], sweeps);
    dev_order();
    MemDev::st(&ctl->phase, g + 1);
    dev_order();
  }
  __syncthreads();
}

// All of doFit for frame f by one workgroup of its team: the frame's seats
// are [s0, s1), this workgroup's waves use seats seat0 + wave.  Counter
// increments accumulate in sh.cnt.
template <int PPL, int SHAPE, int NW, bool MP, int EM = -1>
__device__ void team_fit(const FitArgs& a, int f, const Frame& fr, int s0, int s1, int seat0, TeamShared& sh,
                         double* slice) {
  static_assert(NW <= kTeamSeats, "seats per workgroup");
  TeamCtl* const ctl = a.team_ctl + f;
  const int seat = seat0 + (threadIdx.x >> 6);
  int g = 0;  // the last generation this workgroup ran (0: none)
  for (;;) {
    if (threadIdx.x < kWave) {
      const int act = team_await(a, f, ctl, g, sh);
      if (threadIdx.x == 0) {
        sh.act = act;
        if (act == kActRun) {
          sh.kind = MemDev::ld(&ctl->kind);
          sh.n = MemDev::ld(&ctl->n);
        }
      }
    }
    __syncthreads();
    const int act = sh.act, gen = sh.gen, kind = sh.kind, n = sh.n;
    __syncthreads();
    if (act == kActStop) break;
    if (act == kActLead) {
      if (threadIdx.x < kWave) {
        const bool quiet = team_quiet(a, f, s0, s1);
        if (threadIdx.x == 0) sh.act = quiet ? kActLead : kActStop;
      }
      __syncthreads();
      if (sh.act == kActStop) break;
      team_transition<PPL>(a, fr, ctl, f, gen, sh, slice);
      continue;  // g unchanged: the next await sees the new generation
    }
    g = gen;
    if (kind == kTeamDone || kind == kTeamHandoff) break;
    team_ops<PPL, SHAPE, MP, EM>(a, f, fr, ctl, g, kind, n, seat, sh, slice);
    __syncthreads();
  }
  __syncthreads();
}

// Workgroup b of the launch belongs to frame team_block_frame[b]; its waves
// use seats team_block_seat[b] + wave.  Dynamic LDS as k_fit_frames.
template <int PPL, int SHAPE, int NW, bool MP, int EM = -1>
__global__ __launch_bounds__(NW * 64) void k_fit_team(FitArgs a) {
  constexpr int kSlice = fit_frames_slice<PPL, MP>() + fit_frames_stage<PPL>();
  __shared__ TeamShared sh;
  extern __shared__ double lds_dyn[];
  double* const slice = lds_dyn + (threadIdx.x >> 6) * kSlice;
  const int f = a.team_block_frame[blockIdx.x];
  const Frame fr = a.frames[f];
  if (threadIdx.x < C_COUNT) sh.cnt[threadIdx.x] = 0;
  __syncthreads();
  team_fit<PPL, SHAPE, NW, MP, EM>(a, f, fr, a.team_seat0[f], a.team_seat0[f + 1], a.team_block_seat[blockIdx.x], sh,
                                   slice);
  if (threadIdx.x < C_COUNT && sh.cnt[threadIdx.x] != 0)
    atomicAdd(&a.frame_counters[f * C_COUNT + threadIdx.x], sh.cnt[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------


#define DAO_PPL_SWITCH(nroi, KCALL)                    \
  switch (ppl_for(nroi)) {                             \
    case 2: { constexpr int PPL = 2; KCALL; } break;   \
    case 4: { constexpr int PPL = 4; KCALL; } break;   \
    case 6: { constexpr int PPL = 6; KCALL; } break;   \
    case 8: { constexpr int PPL = 8; KCALL; } break;   \
    case 12: { constexpr int PPL = 12; KCALL; } break; \
    default: { constexpr int PPL = 16; KCALL; } break; \
  }

#define DAO_SHAPE_SWITCH(shape, KCALL)                                 \
  switch (shape) {                                                     \
    case SH_2DFIXED: { constexpr int SHP = SH_2DFIXED; KCALL; } break; \
    case SH_2D: { constexpr int SHP = SH_2D; KCALL; } break;           \
    case SH_3D: { constexpr int SHP = SH_3D; KCALL; } break;           \
    default: { constexpr int SHP = SH_Z; KCALL; } break;               \
  }

static inline dim3 waves_grid(int n) { return dim3((n + kWavesPerBlock - 1) / kWavesPerBlock); }
static inline dim3 flat_grid(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// Launch k_fit_frames<PPL, SHP, NW, MP> for nframes frames with a.m.waves
// waves per workgroup (0: NW; at most NW) and their LDS slices.
template <int PPL, int SHP, int NW, bool MP, int EM, bool RESUME>
void launch_fit_frames_k(const FitArgs& a, int nframes, hipStream_t st) {
  const int nw = a.m.waves > 0 && a.m.waves < NW ? a.m.waves : NW;
  const size_t bytes = sizeof(double) * (size_t)nw * (fit_frames_slice<PPL, MP>() + fit_frames_stage<PPL>());
  static size_t allowed = 64 * 1024;  // larger dynamic LDS must be enabled per kernel
  if (bytes > allowed) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fit_frames<PPL, SHP, NW, MP, EM, RESUME>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      return;  // the launch below would fail; the caller's error check reports it
    allowed = bytes;
  }
  k_fit_frames<PPL, SHP, NW, MP, EM, RESUME><<<nframes, nw * 64, bytes, st>>>(a);
}
template <int PPL, int SHP, int NW, bool MP, int EM = -1>
void launch_fit_frames_t(const FitArgs& a, int nframes, hipStream_t st) {
  if (a.team_resume)
    launch_fit_frames_k<PPL, SHP, NW, MP, EM, true>(a, nframes, st);
  else
    launch_fit_frames_k<PPL, SHP, NW, MP, EM, false>(a, nframes, st);
}

// Launch k_add_peaks<PPL, NW> over njobs workgroups of nw waves.
template <int PPL, int NW>
void launch_add_peaks_t(const FitArgs& a, const AddPeaksJob* jobs, int njobs, int fmt, int nw, hipStream_t st) {
  nw = nw > 0 && nw < NW ? nw : NW;
  const size_t bytes = sizeof(double) * (size_t)nw * fit_frames_slice<PPL, false>();
  static size_t allowed = 64 * 1024;
  if (bytes > allowed) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_add_peaks<PPL, NW>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      return;
    allowed = bytes;
  }
  k_add_peaks<PPL, NW><<<njobs, nw * 64, bytes, st>>>(a, jobs, fmt);
}

// Launch k_fit_team<PPL, SHP, NW, MP> over nblocks workgroups (the team
// tables of `a` map them to frames and seats), waves as k_fit_frames.
template <int PPL, int SHP, int NW, bool MP, int EM = -1>
void launch_fit_team_t(const FitArgs& a, int nblocks, hipStream_t st) {
  const int nw = a.m.waves > 0 && a.m.waves < NW ? a.m.waves : NW;
  const size_t bytes = sizeof(double) * (size_t)nw * (fit_frames_slice<PPL, MP>() + fit_frames_stage<PPL>());
  static size_t allowed = 64 * 1024;
  if (bytes > allowed) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fit_team<PPL, SHP, NW, MP, EM>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      return;
    allowed = bytes;
  }
  k_fit_team<PPL, SHP, NW, MP, EM><<<nblocks, nw * 64, bytes, st>>>(a);
}

// k_fit_frames / k_fit_team launchers (dao_fit_fp64.hip, dao_fit_mixed.hip,
// dao_team_*.hip); FP64 with the MLE error model has its own instances
// (dao_fit_fp64_mle.hip, dao_team_fp64_mle.hip)
void launch_fit_frames_fp64(const FitArgs& a, int nframes, hipStream_t st);
void launch_fit_frames_fp64_mle(const FitArgs& a, int nframes, hipStream_t st);
void launch_fit_team_fp64_mle(const FitArgs& a, int nblocks, hipStream_t st);
void launch_fit_frames_mixed(const FitArgs& a, int nframes, hipStream_t st);
void launch_fit_team_fp64(const FitArgs& a, int nblocks, hipStream_t st);
void launch_fit_team_mixed(const FitArgs& a, int nblocks, hipStream_t st);

}  // namespace dao
