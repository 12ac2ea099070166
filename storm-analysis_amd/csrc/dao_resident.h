// Layout shared by the resident fitter kernel (dao_resident.hip) and its host
// runtime (dao_resident.cpp).
//
// The host publishes requests ("tickets") in pinned, device-mapped host
// memory: it fills the descriptor of a free slot of the batch, writes the
// slot's index into the ticket ring and then raises `tail`.  Workgroups of a
// kernel that stays resident take tickets in order (a device counter `head`),
// fit the slot's frame (mFitNewImage, mFitNewBackground, daoNewPeaks and
// MultiFitter.doFit, dao_fit_c.py:332-356) and write the peaks and the
// completion into the same host memory.  No HIP call is made per request.
#pragma once
#include <hip/hip_runtime.h>

#include "dao_device.h"

namespace dao {

// One slot's request and result.  The host writes the request fields before
// publishing the ticket; the kernel writes the result fields and then `done`
// (= seq) last.
struct ResSlot {
  // request
  long long seq;     // ticket number (> 0)
  const double* img;   // width x height, row-major: device memory or mapped host memory
  const double* bg;    // background estimate, same layout
  const double* rows;  // start peaks, n x 4 ("finder"/"testing") or n x 7 ("text"/"hdf5")
  double* out;         // mapped host memory: n x 9 {x, y, z, height, background, xsigma, ysigma, error, iterations}
  int* status;         // mapped host memory: n
  int W, H, n, fmt;
  int team;            // workgroups fitting the frame (1: one workgroup, the frame-per-workgroup
                       // fitter; more: members 0..team-1 of a team, one ticket each)
  int pad0;
  double rqe, scm;     // uniform calibration
  // result
  int counters[C_COUNT];
  int nfit;    // peaks written to out / status
  int sweeps;  // LM sweeps of doFit
  int err;     // 1: a scheduler wait gave up, 2: the request does not fit the slot, 3: invalid
               // input values (no peaks written)
  int pad;
  long long t_take, t_load, t_done;  // device wall clock (ticks): ticket taken, frame loaded, finished
  long long done;            // seq once the result is complete
};

// Device-side state of a team request in slot f: member 0 loads the frame
// and sets `ready` (= the request's seq) once the others may start; the
// member that leaves last (`left` reaches the team size) writes the result.
struct ResTeamSlot {
  long long ready;
  int left;
  int err;              // the load refused the request (as ResSlot::err)
  int W, H, nfit, pad;  // the frame as loaded
  int cnt[C_COUNT];     // counter increments of the load and of every member
};

// A ticket in the ring: slot | member << 16.
constexpr int kResMemberShift = 16;

struct ResQueue {
  long long tail;  // tickets published
  long long stop;  // non-zero: workgroups leave once no ticket is waiting
};

struct ResArgs {
  ResQueue* q;         // mapped host memory
  const int* ring;     // mapped host memory: ticket t at ring[t % qcap] (slot | member << 16)
  ResSlot* slots;      // mapped host memory, one per slot of the batch
  long long* head;     // device: tickets taken
  long long* last;     // device: wall clock when a ticket was last taken
  long long* busy;     // device: workgroups fitting a frame
  ResTeamSlot* team;   // device: one per slot
  double* rows_dev;    // device: start rows of slot f at rows_dev + 7 * peak0
  const long long* pix_cap;  // device: pixel capacity of each slot
  int nslots;
  int qcap;
  long long idle_ticks;  // the workgroups leave once no frame is in flight and no ticket was taken
                         // for this long
};

// resident fitter launchers (dao_resident_{fp64,mixed}{,_team}.hip):
// `blocks` workgroups that take tickets until the server is idle; every
// request fitted by one workgroup, or (team) by a team of workgroups
void launch_resident_fp64(const FitArgs& a, const ResArgs& r, int blocks, hipStream_t st);
void launch_resident_mixed(const FitArgs& a, const ResArgs& r, int blocks, hipStream_t st);
void launch_resident_team_fp64(const FitArgs& a, const ResArgs& r, int blocks, hipStream_t st);
void launch_resident_team_mixed(const FitArgs& a, const ResArgs& r, int blocks, hipStream_t st);

}  // namespace dao
