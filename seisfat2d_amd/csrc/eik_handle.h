// Handle and batch state of the C ABI (private to seisfat2d_amd/csrc).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/seisfat2d_eik.h"
#include "eik_adjoint.h"
#include "eik_kernels.h"

namespace seisfat {
// Records msg as the calling thread's last error (seisfat_last_error) and
// returns code.
int set_error(int code, const char* msg);
}  // namespace seisfat

using seisfat::AdjShot;
using seisfat::EikConsts;
using seisfat::PersistRun;
using seisfat::ShotDev;
using seisfat::SkewView;

struct seisfat_eik_s {
  int nz = 0, nx = 0, nzz = 0, nxx = 0;
  float dz = 0.f, dx = 0.f;
  EikConsts kc{};
  bool eqh = false;
  int device = 0;
  hipStream_t stream = nullptr;
  float* d_spad = nullptr;  // padded slowness, nzz x nxx (the reference's Modeling::d_S)
  bool has_s = false;
  // the padding of d_spad replicates its interior (every upload through
  // seisfat_eik_set_slowness, and padded uploads whose border is the
  // replication); the adjoint's fold onto the unpadded model assumes it
  bool pad_replicated = true;
  unsigned version = 0;  // incremented by every accepted slowness upload
  uint16_t* d_c[6] = {};
  float cmin[6] = {}, cmax[6] = {};
  float* d_anitab = nullptr;  // 7 planes of nzz*nxx: the qP slowness's shot-independent factors
  bool has_c = false;
  int* d_flag = nullptr;
  // skewed slowness copies, one per sweep orientation (tiled schedule)
  float* d_sk[4] = {};
  int KP = 0, NS = 0;
  SkewView skew[4] = {};
  // 1: workgroups go round robin over 8 XCCs (fused sweep pairs usable), 0:
  // not, -1: not checked yet
  int xcc_ok = -1;
  seisfat_batch_s* single = nullptr;  // one-shot batch of seisfat_eik_solve, reused while its shape holds
};

// Adjoint state of a batch, allocated on its first recorded run.
struct AdjState {
  bool ready = false;
  bool recorded = false;  // a recorded run exists for model `version`
  unsigned version = 0;
  bool pad_replicated = true;  // the recorded model's border replicated its edges
  std::vector<AdjShot> tab;
  AdjShot* d_tab = nullptr;
  int nstr = 0, max_nxx = 0;
  size_t ksize = 0;
  double* d_T64 = nullptr;  // fp64 trajectory: T0 (initial field) .. T4, tsize each
  float* d_lamA = nullptr;
  float* d_lamB = nullptr;
  float4* d_coef = nullptr;
  uint8_t* d_code = nullptr;
  float* d_gc = nullptr;
  float* d_gshot = nullptr;
  size_t* d_seed_node = nullptr;
  int* d_seed_ptr = nullptr;
  int* d_seed_rec = nullptr;
  float* d_seed_w = nullptr;
  int nseed = 0;
  double* d_picks64 = nullptr;  // picks of the fp64 trajectory
  // strip-parallel fp64 walk and reverse sweeps: ticket counter + wait-limit
  // flag, each strip's published last row (tagged words, walk_ldx per strip),
  // launch tag
  int* d_walk_ctl = nullptr;
  unsigned long long* d_walk_x = nullptr;
  int walk_ldx = 0;
  unsigned walk_tag = 0;
  float* d_w = nullptr;      // cotangents of the host-side misfit call
  double* d_grad = nullptr;  // gradient of the host-side misfit call
};

struct seisfat_batch_s {
  seisfat_eik_s* h = nullptr;
  int type = SEISFAT_ISO, schedule = SEISFAT_SCHED_TILED;
  seisfat_sched_opts opts{1, 0, 1, 1};
  EikConsts kc{};  // the geometry's constants with the skip option applied
  int nshots = 0, nrec = 0;
  std::vector<seisfat_shot> shots;
  std::vector<ShotDev> iso, qp;  // host copies of the shot tables
  std::vector<size_t> toff;      // offset of each shot's T grid (floats)
  size_t tsize = 0;
  ShotDev* d_iso = nullptr;
  ShotDev* d_qp = nullptr;
  float* d_T = nullptr;
  float* d_swork = nullptr;
  float* d_rx = nullptr;
  float* d_rz = nullptr;
  int* d_recshot = nullptr;
  float* d_picks = nullptr;
  int4* d_tiles = nullptr;
  int ntiles = 0, nstrips = 0;
  std::vector<int> diag_off;
  // receivers relative to their shot's grid origin, and their shots
  std::vector<float> rxr, rzr;
  std::vector<int> recshot;
  AdjState adj;
  // persistent schedule state
  unsigned long long* d_xch = nullptr;  // exchange words of every shot
  size_t xch_words = 0;
  int* d_ctl = nullptr;  // [0..NUM_COUNTERS) ticket counters, [CTL_EPOCH] epoch, [CTL_ERR] error flag
  // fused sweep pairs (isotropic solve of the tiled schedule): per-XCD queues
  bool pairs = false;
  std::vector<ShotDev> iso_b, qp_b;  // iso / qp with the pairs' second-sweep exchange buffers
  ShotDev* d_iso_b = nullptr;
  ShotDev* d_qp_b = nullptr;
  int4* d_ptiles = nullptr;
  int* d_qoff = nullptr;
  std::vector<int4> pair_tiles;  // host copies of the queues' contents
  std::vector<int> pair_qoff;
  // the last run was a plain pass (run / recorded run), whose fused-pair
  // queues batch_status checks were drained
  bool last_run_plain = false;
  PersistRun pr{};
  int launches = 0, sweep_launches = 0;
  hipEvent_t ev_sweep[2] = {nullptr, nullptr};
  // recorded after the batch's last enqueued work (run, recorded run or
  // vector-Jacobian product), on whatever stream that work went to
  hipEvent_t ev_done = nullptr;
  // The tiled solve is recorded as three graphs: initialisation, sweeps (for
  // anisotropic batches: both solves and the qP slowness between them), and
  // picks; the timing events are recorded on the stream between them.
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  // converged mode: control words, and graphs of the first pass, of `conv_chunk`
  // gated passes and of the picks, captured for pass limit `conv_max`
  int* d_conv = nullptr;
  int conv_chunk = 0, conv_max = 0;
  hipGraph_t conv_graph[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t conv_exec[3] = {nullptr, nullptr, nullptr};
};
