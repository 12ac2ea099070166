// Host runtime of the MI355X fitter: owns the device state of a batch of
// frames and drives the HIP kernels.  The single-frame C-ABI (dao_fit_abi.cpp)
// and the batch / serving API (dao_batch.cpp) are both built on it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "dao_device.h"

#define DAO_CHECK(expr)                                                                       \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) dao::fatal(__FILE__, __LINE__, #expr, hipGetErrorString(e_));      \
  } while (0)

namespace dao {

// Default of Engine::team_solo (daoBatchSetTeamHandoff).
constexpr int kTeamSolo = 32;

// Error codes of the library's entry points (0: success).  No entry point
// aborts the process: a failure inside the engine throws dao::Error, the
// C-ABI function that was called catches it, records it as the calling
// thread's last error (daoGetLastError) and returns (batch entry points
// return the code).
enum ErrCode {
  kOk = 0,
  kErrArg = -1,       // an argument is out of range
  kErrHip = -2,       // a HIP runtime call failed (device memory, launch, ...)
  kErrCapacity = -3,  // more peaks than a frame's slots
  kErrSched = -4,     // the ordered executor stopped making progress
  kErrFrame = -5,     // some frames of a batch failed (daoBatchFailedFrames)
  kErrInternal = -6   // anything else
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

// Throw dao::Error(kErrHip) for a failed HIP call.
[[noreturn]] void fatal(const char* file, int line, const char* what, const char* msg);
// Throw dao::Error(code, msg).
[[noreturn]] void fail(int code, const char* msg);
// The calling thread's last error.
void set_last_error(int code, const char* msg);
int take_last_error(char* msg, int len);

// Run fn; an exception becomes the calling thread's last error and its code
// is returned (0 when fn returns normally).
template <class Fn>
int guarded(Fn&& fn) {
  try {
    fn();
    return kOk;
  } catch (const Error& e) {
    set_last_error(e.code, e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error(kErrInternal, "host memory exhausted");
    return kErrInternal;
  } catch (const std::exception& e) {
    set_last_error(kErrInternal, e.what());
    return kErrInternal;
  } catch (...) {
    set_last_error(kErrInternal, "unknown failure");
    return kErrInternal;
  }
}

// kernel launchers (dao_kernels.hip)

void launch_op_round(int kind, const OpArgs& a, hipStream_t st, const double* rows, int fmt, int first_slot);
void launch_binning(const Sched& sc, const Frame* frames, const Peaks& pk, int ncells, hipStream_t st);
void launch_box(int kind, const Sched& sc, const Peaks& pk, const Frame* frames, const Model& m,
                const int4* override_box, const double* rows, int fmt, int first_slot, hipStream_t st);
void launch_commit(const Sched& sc, const Peaks& pk, int* frame_counters, hipStream_t st);
void launch_collect_escapes(const Sched& sc, const Peaks& pk, int4* override_box, hipStream_t st);
void launch_new_image(const Pix& px, long long pix0, long long n, const double* img, hipStream_t st);
void launch_new_background(const Pix& px, long long pix0, long long n, const double* bg, hipStream_t st);
void launch_new_frame_pixels(const Pix& px, long long pix0, long long n, const double* img, const double* bg,
                             hipStream_t st);
void launch_set_calibration(const Pix& px, long long pix0, long long n, const double* rqe, const double* scm,
                            hipStream_t st);
void launch_set_calibration_const(const Pix& px, long long pix0, long long n, double rqe, double scm,
                                  hipStream_t st);
void launch_anscombe(const Pix& px, long long pix0, long long n, hipStream_t st);
void launch_fit_image(const Pix& px, long long pix0, long long n, double* out, int residual, hipStream_t st);
void launch_peak_sums(const Model& m, const Pix& px, const Peaks& pk, const Frame* frames, const int* roi_tab,
                      int slot0, int n, int what, double* out, hipStream_t st, const int* slots = nullptr);
void launch_packed_slots(const Frame* frames, int nframes, const int* off, int* slots, hipStream_t st);
void launch_status_flag(const Peaks& pk, const Frame* frames, const int* off, long long nslots, const int* status,
                        int* flag, hipStream_t st);
void launch_apply_status(const Peaks& pk, const Frame* frames, const int* off, long long nslots, const int* status,
                         hipStream_t st);
// order-preserving list of the flagged entries of flag[0, n) into out, count in *count
void launch_compact_list(const int* flag, long long n, int* out, int* count, hipStream_t st);
void launch_compact_frames(const Peaks& pk, const Frame* frames, int nframes, int drop, int* counts, int* counters,
                           hipStream_t st);
void launch_peak_jacobian(const Model& m, const Pix& px, const Peaks& pk, const Frame* frames,
                          const int* roi_tab, int slot0, int n, double* out, hipStream_t st);
void launch_gather_peaks(PeakRec* dst, const PeakRec* src, const int* map, int slot0, int n, hipStream_t st);
void launch_set_status(const Peaks& pk, int slot0, int n, const int* status, hipStream_t st);
void launch_set_owner(const Peaks& pk, long long s0, long long s1, int frame, hipStream_t st);
void launch_select(const Peaks& pk, const Frame* frames, long long nslots, int mode, const int* frame_bad,
                   int* flag, int* out, int* count, hipStream_t st);
void launch_fill_i32(int* p, long long n, int v, hipStream_t st);
void launch_iota(int* p, int n, int base, hipStream_t st);
void launch_fit_frames(const FitArgs& a, int nframes, hipStream_t st);
// k_add_peaks over njobs workgroups of nw waves (0: the kernel's maximum)
void launch_add_peaks(const FitArgs& a, const AddPeaksJob* jobs, int njobs, int fmt, int nw, hipStream_t st);
void launch_fit_team(const FitArgs& a, int nblocks, hipStream_t st);
void launch_fit_mp(const MpArgs& a, hipStream_t st);  // dao_mp.hip
void launch_team_init(TeamCtl* ctl, int nframes, int* cur, int nseats, hipStream_t st);
void launch_restore_bad_frames(const Frame* frames, int nframes, const int* bad, PixRW* rw, const PixRW* rw_copy,
                               PeakRec* rec, const PeakRec* rec_copy, hipStream_t st);
void launch_pack_peaks(const Peaks& pk, const Frame* frames, const int* offset, const Model& m, long long nslots,
                       double* out, int* status, hipStream_t st);

// Reference ROI table (multi_fit.c:1031-1060): pixels whose centre lies within
// r/2 of the ROI centre, row-major.
void roi_table(int roi, std::vector<int>& rx, std::vector<int>& ry);

struct FrameSpec {
  int W, H;
  int cap;  // peak slots reserved for the frame
};

class Engine {
 public:
  Engine(const Model& m, const std::vector<FrameSpec>& frames);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  Model m;
  hipStream_t st = nullptr;
  int F = 0;
  std::vector<Frame> fr;
  Frame* d_fr = nullptr;
  Frame* h_fr = nullptr;  // pinned copy of fr for uploads
  Pix px{};
  long long npix = 0;
  Peaks pk{};
  long long cap = 0;  // total peak slots
  std::vector<int> roi_x, roi_y;
  int* d_roi = nullptr;
  Sched sc{};
  int ncells = 0;
  int4* d_override = nullptr;
  int* d_counters = nullptr;  // [F][C_COUNT]
  int* d_flag = nullptr;
  int* d_count = nullptr;
  int* h_pin = nullptr;  // pinned scratch for small read-backs
  size_t h_pin_bytes = 0;
  double* d_stage = nullptr;  // upload staging
  size_t stage_bytes = 0;
  int* d_istage = nullptr;
  size_t istage_bytes = 0;
  int max_attempts_seen = 0;
  long long escapes_total = 0;
  int num_cus = 0;  // compute units of the device
  long long rounds_total = 0;

  void set_model(const Model& mm);
  void upload_frames();
  // image state
  void new_image(int f, const double* img_host);
  void new_image_dev(int f, const double* img_dev);
  void new_background(int f, const double* bg_host);
  void new_background_dev(int f, const double* bg_dev);
  void anscombe(int f);
  void set_calibration(int f, const double* rqe_host, const double* scmos_host);
  // peaks
  // room for `need` peaks in frame f: a single-frame engine (and one marked
  // growable) grows its peak storage, any other refuses
  void ensure_frame_cap(int f, int need);
  void grow_frames(int newcap);  // every frame to newcap slots
  bool growable = false;
  void new_peaks(int f, const double* rows_host, int fmt, int n);
  void new_peaks_multi(const std::vector<int>& frames, const std::vector<const double*>& rows_dev,
                       const std::vector<int>& counts, int fmt);
  void add_peaks(const std::vector<AddPeaksJob>& jobs, int fmt);
  void compact(int f, const std::vector<int32_t>& keep);
  void set_status(int f, const int32_t* status);
  // every frame at once, on the device (rows in frame order, as pack_all):
  // mFitSetPeakStatus from a device status array; mFitRemoveErrorPeaks
  // (drop ST_ERROR) / mFitRemoveRunningPeaks (drop ST_RUNNING); the peaks'
  // significance into a device array
  void set_status_all(const int32_t* status_dev);
  void remove_all(int drop);
  void significance_all(double* out_dev);
  // frames the next fit may touch (empty: all); set by the caller, cleared
  // by nothing
  std::vector<char> fit_mask;
  // fitting
  // one mFitIterateLM over every frame of the batch; returns the number of
  // RUNNING peaks it started with (0: nothing to do)
  int sweep();
  // MultiFitter.doFit (dao_fit_c.py:332-356) for every frame: one sweep, then
  // more while any peak is RUNNING, at most max_iterations more; returns the
  // number of sweeps
  int fit(int max_iterations);
  int running_count();
  // sweeps of every frame with peaks, one persistent workgroup per frame:
  // max_more = 0 runs exactly one sweep (mFitIterateLM), otherwise doFit.
  // A frame whose scheduler wait gave up is left empty (nfit = 0) and
  // flagged in `failed`; the other frames' results stand.
  int fit_frames(int max_more);
  // Arguments of the frame fitters over this engine's state, without a
  // frame list or team tables (the resident fitter takes frames itself).
  FitArgs fitter_args(int max_more);
  std::vector<int> failed;   // per frame: failed in the last fit
  int failed_count = 0;
  long long spin_limit = 0;  // polls before a scheduler wait gives up (0: kSpinLimit)
  // teams hand a frame's LM runs of at most this many ops to one workgroup
  // (0: never): ~22 us per tail sweep against ~42 us through the team on
  // 512^2 Z frames of 1000 emitters, measured on one MI355X; 32 and 64 gave
  // the lowest single-frame latency there, and no change on a dense 2048^2
  // frame or an uneven batch with teams of 500
  int team_solo = kTeamSolo;
  bool use_rounds = false;  // multi-kernel executor instead of fit_frames
  // timing of the LM kernel (HIP events on the engine's stream)
  bool timing = false;
  double lm_ms = 0.0;
  long long lm_launches = 0;
  long long sweeps_total = 0;
  // read-back
  void read_counters(std::vector<int>& out);  // F * C_COUNT, zeroes the device copy
  void read_counters(std::vector<int>& out, int f);  // frame f's C_COUNT, zeroes them
  void get_records(int f, std::vector<PeakRec>& out);  // frame f's live peak records
  void get_status(int f, std::vector<int32_t>& out);
  void get_iterations(int f, std::vector<int32_t>& out);
  void get_error(int f, std::vector<double>& out);
  void get_param(int f, int q, std::vector<double>& out);
  void get_sums(int f, int what, std::vector<double>& out);
  void get_jacobian(int f, std::vector<double>& out);
  void get_image(int f, double* out_host, int residual);
  // all frames' live peaks as rows of 9 doubles + status (host or device
  // buffers)
  void pack_all(double* out, int32_t* status, bool on_device);
  void get_images_all(double* out, int residual, bool on_device);
  long long total_nfit() const;
  // every frame at once from device arrays concatenated at the frames'
  // current sizes
  void new_images_all_dev(const double* img_dev);
  // new_images_all_dev and new_backgrounds_all_dev in one pass
  void new_frames_all_dev(const double* img_dev, const double* bg_dev);
  void new_backgrounds_all_dev(const double* bg_dev);
  long long frame_pixels_total() const;
  // serving: re-size slot f within its capacity (false if it does not fit),
  // empty it, uniform calibration
  bool set_frame_shape(int f, int W, int H, bool upload = true);
  void clear_frame(int f);
  void set_calibration_const(int f, double rqe, double scmos);
  std::vector<long long> pix_cap;  // pixel capacity of each slot
  std::vector<int> cell_cap;       // scheduling-cell capacity of each slot
  int count_status(int f, int status);
  void sync();

 private:
  void init(const std::vector<FrameSpec>& specs);
  void release();  // frees everything (also after a failed construction)
  void alloc_sched(long long c);
  void free_sched();
  void grow_total_cap(long long c);
  int* pin(size_t bytes);
  double* stage(size_t bytes);
  int* istage(size_t bytes);
  // row offsets of the frames (frame order, F + 1 entries) on the device
  int* upload_offsets(std::vector<int>& off);
  int select(int mode);
  bool packed() const;  // every frame at its creation size
  void run_ops(int kind, const double* rows, int fmt, int first_slot);
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int* d_run = nullptr;  // RUNNING list of the frame fitters (per slot)
  double* d_journal = nullptr;  // LM turn journals of the frame fitters (per slot)
  PixRW* d_sweep_rw = nullptr;    // multi-kernel executor: pixel fit state at the sweep's start
  PeakRec* d_sweep_rec = nullptr;  // and the peak records (per slot)
  int* d_sweeps = nullptr;
  int* d_flist = nullptr;
  int* d_err = nullptr;
  // frame-per-team fitter: control words, seats, workgroup tables
  TeamCtl* d_team_ctl = nullptr;
  int* d_team_cur = nullptr;
  int* d_team_tab = nullptr;  // block frame [blocks], block seat [blocks], seat0 [F + 1]
  int team_seat_cap = 0, team_tab_cap = 0;
  int team_blocks(std::vector<int>& tab, int& nseats) const;
  OpArgs args(int round);
};

}  // namespace dao
