// Device-side data layout shared by the HIP kernels and the host engine.
//
// One "batch" holds F frames.  Pixel records of all frames share two arrays
// (frame f occupies [pix0, pix0 + W*H)), row-major with x the fast axis,
// pixel index k = (yi + ry) * W + (xi + rx) (multi_fit.c:50).  Peak records
// share one array; frame f owns slots [peak0, peak0 + cap) and its live
// peaks are the first nfit of them, in the reference's index order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dao {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kMaxPred = 128;  // stored predecessors per op (more -> cell rescan)
constexpr int kNFit = 7;

// parameter slots (multi_fit.h:20-26)
constexpr int P_HEIGHT = 0, P_XC = 1, P_XW = 2, P_YC = 3, P_YW = 4, P_BG = 5, P_ZC = 6;
// status (multi_fit.h:33-35)
constexpr int ST_RUNNING = 0, ST_CONVERGED = 1, ST_ERROR = 2;
// models
enum Shape { SH_2DFIXED = 0, SH_2D = 1, SH_3D = 2, SH_Z = 3 };
enum ErrModel { EM_MLE = 0, EM_ALS = 1, EM_LS = 2, EM_DWLS = 3, EM_FWLS = 4 };
// counters, in the order of the fitData prefix (multi_fit.h:87-94)
enum Counter {
  C_DPOSV = 0, C_ITER, C_LOST, C_MARGIN, C_NEG_FI, C_NEG_HEIGHT, C_NON_CONV, C_NON_DECR, C_COUNT
};
// kinds of image operation run through the ordered executor
enum OpKind { OP_LM = 0, OP_RECENTER = 1, OP_NEWPEAK = 2, OP_SUBTRACT = 3, OP_PEAKERROR = 4 };
// selection predicates over peak slots (k_flag)
enum SelectMode { SEL_RUNNING = 0, SEL_RECENTER = 1, SEL_RUNNING_BAD = 2 };
// arithmetic of the LM turns in k_fit_frames: FP64 throughout, or FP32
// per-pixel work with FP64 accumulators, solve and accept/reject (MLE only)
enum Precision { PREC_FP64 = 0, PREC_MIXED = 1 };
// start-peak formats of daoNewPeaks (dao_fit.c:1605, 1695)
enum PeakFmt { PF_FINDER = 0, PF_TESTING = 1, PF_TEXT = 2 };

struct Model {
  int shape;
  int errm;
  int roi;       // fit_size_x == fit_size_y
  int nroi;      // roi_n_index
  int jac;       // jac_size
  int margin;    // reserved-box margin around the ROI, pixels
  int prec;      // Precision
  int waves;     // waves per workgroup of k_fit_frames (0: the kernel's maximum)
  int team;      // peaks per workgroup of the frame-per-team fitter (0: one workgroup per frame)
  double xoff;   // == yoff
  double zoff;
  double tol;
  double min_height;
  double wmin, wmax;
  double zmin, zmax;
  double wxz[5], wyz[5];  // wxz[0], wyz[0] hold wo^2 (dao_fit.c:1566-1567)
};

struct Frame {
  int W, H;
  long long pix0;
  int peak0;
  int cap;
  int nfit;
  int cell0;  // first cell of this frame's scheduling grid
  int ncx, ncy;
};

// Pixel state, one 32-byte record per pixel.  rw: what the fit changes;
// ro: the frame's data and calibration.
struct PixRW {
  double f;    // fit foreground (rqe applied)
  double bg;   // sum of covering peaks' bg*rqe + scmos
  double tfi;  // cached transformed fit value (multi_fit.c:160)
  int cnt;     // number of peaks covering the pixel
  int stale;   // fit changed since the error was last evaluated here
};
struct PixRO {
  double x;      // image * rqe + scmos
  double rqe;
  double scm;    // variance / gain^2
  double bgest;  // background estimate * rqe
};
struct Pix {
  PixRW* rw;
  PixRO* ro;
  double* asx;  // Anscombe transform of x (ALS only)
};

// Peak state, one 96-byte record per peak slot (peakData, multi_fit.h:58-74).
struct PeakRec {
  double prm[kNFit];
  double lam;
  double err;
  int xi, yi;
  int status;
  int iter;
  int added;
  int frame;  // owner frame of the slot
};
struct Peaks {
  PeakRec* rec;
  PeakRec* bk;  // copy at the start of a sweep (rollback)
  long long cap;
};

// The ordered executor: ops i < j of one frame whose boxes overlap run in
// index order; the others run concurrently.
struct Sched {
  int nops;
  int* op_peak;     // peak slot of op i
  int4* box;        // [x0, y0, x1, y1) per op
  int* pred_n;      // number of predecessors (> kMaxPred: rescan cells)
  int* pred;        // [nops][kMaxPred]
  int* done;        // round in which the op finished, -1 pending
  int* cell_start;  // per cell, into cell_ops (size ncells + 1)
  int* cell_ops;    // op ids sorted by cell
  int* cell_fill;
  int* big_ops;     // ops whose box is wider than the cell
  int* n_big;
  int cell;         // cell edge, pixels
  int* op_delta;    // [nops][C_COUNT] counter deltas of the op
  int* op_flags;    // bit0 completed, bit1 escaped
  int4* esc_box;    // box needed by an escaped op
  int* frame_bad;   // per frame: an op escaped its box in this attempt
  int* remaining;   // ops not yet done
};

struct OpArgs {
  Model m;
  Pix px;
  Peaks pk;
  Sched sc;
  const Frame* frames;
  const int* roi_tab;
  int round;
};

// Scheduling state of one frame fitted by a team of workgroups
// (k_fit_team); only device-coherent accesses touch it during the launch.
struct TeamCtl {
  int phase;   // generation: raised by each transition (a new run, a rollback, the end)
  int lead;    // the generation whose transition a workgroup has claimed
  int kind;    // OP_LM, OP_RECENTER, kTeamDone or kTeamHandoff
  int n;       // ops of the current run
  int next;    // next op to hand out
  int done;    // ops of the run finished
  int esc;     // smallest op whose LM turn left its box
  int it;      // doFit iteration (mFitIterateLM call)
  int sweeps;  // LM sweeps run
  int nrun;    // RUNNING peaks at the current sweep's start (the LM run's list)
  int pad[6];
};
constexpr int kTeamDone = -1;
// The team stopped at an LM run of at most FitArgs::team_solo ops (run_list,
// nrun, it and sweeps as the transition left them): k_fit_frames with
// FitArgs::team_resume finishes the frame with one workgroup.
constexpr int kTeamHandoff = -2;
constexpr int kTeamSeats = 8;  // seats (one per wave) reserved per workgroup
constexpr int kMaxTeam = 64;   // workgroups per frame

// ROI pixels per lane of the kernel instance that fits an ROI of nroi pixels.
__host__ __device__ inline int ppl_for(int nroi) {
  int p = (nroi + kWave - 1) / kWave;
  if (p <= 2) return 2;
  if (p <= 4) return 4;
  if (p <= 6) return 6;
  if (p <= 8) return 8;
  if (p <= 12) return 12;
  return 16;
}

// An LM turn that may still be undone journals, for each ROI position it
// loads (the start and every new candidate position), the pixels' fit state
// as it was before the turn touched them: a snapshot of (3 PPL + 1) * 64
// words -- words s * 64 + l, (PPL + s) * 64 + l and (2 PPL + s) * 64 + l the
// foreground, background sum and cached transformed fit of slot s of lane l,
// word 3 PPL * 64 + l the lane's stale bits (low 32 bits) and the ROI origin
// (xi | yi << 16, high 32 bits).  Restoring the exact values (not adding the
// stamps back) matters: a rejected candidate's stamp can be so large that
// adding and removing it wipes out the pixel's other contributions, and an
// undo must not do that a second time.  Covering counts are integers and are
// undone by arithmetic.  A turn that would load more positions than
// kJournalSnaps gives up like one that leaves its box.
constexpr int kJournalSnaps = 4;
__host__ __device__ constexpr long long journal_snap_words(int ppl) { return (long long)(3 * ppl + 1) * kWave; }
__host__ __device__ constexpr long long journal_words(int ppl) { return kJournalSnaps * journal_snap_words(ppl); }

struct FitArgs {
  Model m;
  Pix px;
  Peaks pk;
  const Frame* frames;
  const int* roi_tab;
  int* op_peak;         // per-slot scratch: op list of the frame (recentering runs)
  int* run_list;        // per-slot scratch: RUNNING peaks of the sweep (LM runs)
  int4* box;            // per-slot scratch: op boxes
  int* op_delta;        // per-slot scratch: counter increments of an op
  int* op_done;         // per-slot scratch: op finished in the current phase (LM: 1 + snapshots journaled)
  double* journal;      // per-slot scratch: cache snapshots of an LM turn (journal_words(PPL) each)
  int* frame_counters;  // [F][C_COUNT]
  int* frame_sweeps;    // [F]: sweeps run by this launch
  const int* frame_list;
  int max_more;         // doFit's max_iterations; 0 = exactly one sweep
  int* frame_err;       // [F]: set when one of the frame's waits exceeds its bound
  long long spin_limit; // polls before a scheduler wait gives up
  // frame-per-team fitter only
  TeamCtl* team_ctl;            // [F]
  int* team_cur;                // [seats]: whether the seat's wave is inside a run
  const int* team_block_frame;  // [workgroups]: frame of the workgroup
  const int* team_block_seat;   // [workgroups]: first seat of the workgroup
  const int* team_seat0;        // [F + 1]: frame f's seats are [team_seat0[f], team_seat0[f + 1])
  int team_solo = 0;            // hand an LM run of at most this many ops to one workgroup (0: never)
  int team_resume = 0;          // k_fit_frames: continue the frames a team handed off (kTeamHandoff)
};

// One frame's share of a k_add_peaks launch: n start rows (device memory,
// daoNewPeaks layout) become peaks first .. first + n - 1 of frame `frame`.
struct AddPeaksJob {
  const double* rows;
  int frame;
  int n;
  int first;
  int pad;
};

// The multiplane fitter (k_fit_mp): frame-sets of nch channels, channel c of
// set k in frame set_frame0[k] + c; peak i of a set is slot peak0 + i of each
// channel's frame.  cpl: the coupling table in device memory, in this order:
// xt_0toN[3 nch], yt_0toN[3 nch], xt_Nto0[3 nch], yt_Nto0[3 nch],
// width_min[nch], width_max[nch], w_x[n_weights nch], w_y[n_weights nch].
constexpr int kMpMaxChannels = 8;
struct MpArgs {
  Model m;  // 2D Gaussian, MLE; widths per channel from cpl
  Pix px;
  Peaks pk;
  const Frame* frames;
  const int* roi_tab;
  int nch;
  int nsets;
  const int* set_frame0;
  const double* cpl;
  int n_weights;
  double w_z_offset, w_z_scale;
  double tol;
  int max_more;         // doFit's max_iterations; 0 = exactly one sweep
  int* frame_counters;  // [F][C_COUNT]
  int* frame_sweeps;    // [F]: sweeps run, at each set's channel-0 frame
};

}  // namespace dao
