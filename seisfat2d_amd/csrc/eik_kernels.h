// Launch interface of the eikonal kernels (host side of eik_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "eik_device.h"

namespace seisfat {

// Compressed stiffness of the padded model (src/modeling/eikonal_ani.cuh:10-17).
struct StiffDev {
  const uint16_t* c[6];  // C11, C13, C15, C33, C35, C55
  float mn[6], mx[6];
};
// Shot-independent factors of the qP slowness per padded node:
// C11, C13, C15, C33, C35, C55 each divided twice by Ro, then Ro = C33 s^2.
struct AniTab {
  float* v[7];
};

void launch_fill(float* p, size_t n, float v, hipStream_t st);
void launch_recip(const float* b, float* r, int n, hipStream_t st);
// bad[d] |= 1 unless fdiv(a, zh[d], zl[d]) == a / b[d] for every dividend of
// magnitude in [2^-100, 2^100)
void launch_fastdiv_check(const float* b, const float* zh, const float* zl, int n, int* bad, hipStream_t st);
void launch_expand(const float* s, float* spad, int nz, int nx, hipStream_t st);
void launch_check_positive(const float* s, size_t n, float hmin, float hmax, int* bad, hipStream_t st);
void launch_init(const ShotDev* shots, int nshots, int nzz, float dx, float dz, hipStream_t st);
// Returns the number of kernel launches issued.
int launch_sweeps_plain(const ShotDev& sd, int nzz, const EikConsts& k, bool eqh, hipStream_t st);
// skew: four skewed slowness views (one per sweep) or nullptr for slowness
// staged in LDS from ShotDev's natural view; first_fused: sweep 0 generates
// the initial traveltimes (no fill/init launches needed before it).
int launch_sweeps_tiled(const ShotDev* shots, const int4* tiles, const std::vector<int>& diag_off, int nzz,
                        const EikConsts& k, bool eqh, const SkewView* skew, bool first_fused, hipStream_t st);
// Device state of the persistent schedule: ticket counters (two solves x
// four sweeps, then two solves x two fused sweep pairs x 8 XCD queues), the
// run epoch and the spin-limit error flag.
constexpr int PAIR_COUNTERS = 8;   // first fused-pair counter
constexpr int NUM_COUNTERS = 40;   // counters zeroed at the start of a run
// control block layout (ints): epoch, error flag, tiles skipped by the
// last run (all inputs TINF)
constexpr int CTL_EPOCH = 40, CTL_ERR = 41, CTL_SKIP = 42, CTL_WORDS = 48;
struct PersistRun {
  int* counters;
  unsigned* epoch;
  int* err;
  int wpcu = 0;  // persistent workgroups per CU (0 = as many as fit)
};
// Both sweeps of each pair (0,1) and (2,3) in one launch each, shots
// partitioned over the 8 XCDs: ptiles holds each XCD's tiles of both sweeps
// (w bit 2 = second sweep, bits 0-1 = priority) at [qoff[x], qoff[x+1]);
// shots_b differs from shots_a only in its exchange buffers.  skew: the
// skewed slowness copies, or null for slowness staged from the shots' own
// grids (the qP solve); solve (0 or 1) selects sweep ids and queue counters.
int launch_sweeps_pairs(const ShotDev* shots_a, const ShotDev* shots_b, const int4* ptiles, const int* qoff,
                        int nstrips, int nzz, const EikConsts& k, bool eqh, const SkewView* skew,
                        const PersistRun& pr, int solve, hipStream_t st);
// out[i] = XCC id of workgroup i (i < n)
void launch_xcc_ids(int* out, int n, hipStream_t st);
// changed / skip (converged mode, skewed slowness only): every tile that
// lowers a traveltime sets *changed; the launches exit at once while *skip
// is set.
int launch_sweeps_persist(const ShotDev* shots, const int4* tiles, int ntiles, int nstrips, int nzz,
                          const EikConsts& k, bool eqh, const SkewView* skew, bool first_fused,
                          const PersistRun& pr, int solve, hipStream_t st, int* changed = nullptr,
                          const int* skip = nullptr);
// Converged mode's control words (see k_conv_init / k_conv_gate).
constexpr int CONV_WORDS = 8;
void launch_conv_init(int* cw, int max_passes, hipStream_t st);
void launch_conv_gate(int* cw, hipStream_t st);
void launch_begin_run(const PersistRun& pr, hipStream_t st);
void launch_skew_s(const float* spad, float* sk, int nzz, int nxx, int KP, int NS, int sweep, hipStream_t st);
// rec_shot null: all receivers belong to shot 0; err (nullable) gets bit 1
// for a receiver outside its shot's grid (picked as NaN)
void launch_picks(const ShotDev* shots, const int* rec_shot, const float* rx, const float* rz, float* out,
                  int nrec, int nzz, float dx, float dz, hipStream_t st, int* err = nullptr);
void launch_ani_table(const ShotDev& model, int nzz, int nxx, const StiffDev& stf, const AniTab& tb,
                      hipStream_t st);
void launch_quasi_slowness(const ShotDev* iso, ShotDev* qp, int nshots, int nzz, int nxx, float dx, float dz,
                           const AniTab& tb, hipStream_t st);

}  // namespace seisfat
