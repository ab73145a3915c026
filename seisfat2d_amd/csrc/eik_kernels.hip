// HIP kernels of the eikonal path (gfx950).
//
//   k_fill            time_set             (src/modeling/modeling.cu:286-291)
//   k_init            time_init            (src/modeling/modeling.cu:293-306)
//   k_sweep_plain     inner_sweep, one launch per anti-diagonal — the
//                     reference schedule (src/modeling/modeling.cu:123-159)
//   k_sweep_tiles     the same sweep as 64x64 skewed tiles staged in LDS,
//                     one launch per tile diagonal over every shot of a batch
//   k_picks           compute_seismogram's Catmull-Rom picks
//                     (src/modeling/modeling.cu:161-194, src/admin/admin.cpp:121-134)
//   k_quasi_slowness  get_quasi_slowness   (src/modeling/eikonal_ani.cu:47-111)
//   k_expand          expand_boundary      (src/modeling/modeling.cu:204-232)
#include "eik_kernels.h"

#include <type_traits>

namespace seisfat {

// ---------------------------------------------------------------------------
__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// expand_boundary of an unpadded nz x nx model into the padded nzz x nxx grid:
// every padded node copies the nearest model node (edge rows and columns
// replicated, src/modeling/modeling.cu:204-232).
__global__ void k_expand(const float* __restrict__ s, float* __restrict__ spad, int nz, int nx) {
  const int nzz = nz + 2 * NB;
  const size_t n = (size_t)nzz * (nx + 2 * NB);
  size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; idx < n; idx += stride) {
    const int i = (int)(idx % nzz);
    const int j = (int)(idx / nzz);
    const int im = min(max(i - NB, 0), nz - 1);
    const int jm = min(max(j - NB, 0), nx - 1);
    spad[idx] = s[im + (size_t)jm * nz];
  }
}

// Flags slowness outside the solver's operating range: non-finite or
// non-positive, or slowness*spacing / slowness/spacing outside [1e-12, 1e12].
__global__ void k_check_positive(const float* __restrict__ s, size_t n, float hmin, float hmax,
                                 int* __restrict__ bad) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  int flag = 0;
  for (; i < n; i += stride) {
    const float v = s[i];
    if (!(v > 0.0f) || !(v < __builtin_inff())) {
      flag = 1;
    } else {
      const double d = v;
      if (d * hmin < 1e-12 || d * hmax > 1e12 || d / hmax < 1e-12 || d / hmin > 1e12) flag = 1;
    }
  }
  if (flag) atomicOr(bad, 1);
}

// ---------------------------------------------------------------------------
// time_init: one thread per (shot, stencil node).  Uses the node's own
// slowness and the exact source position; powf(d,2) is the rounded square.
__global__ void k_init(const ShotDev* __restrict__ shots, int nshots, int nzz, float dx, float dz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nshots * 9) return;
  const ShotDev sd = shots[t / 9];
  const int ii = (t % 9) % 3, jj = (t % 9) / 3;
  const int sIdx = (int)((sd.sx + 0.5f * dx) / dx) + NB;
  const int sIdz = (int)((sd.sz + 0.5f * dz) / dz) + NB;
  const int zi = sIdz + (ii - 1);
  const int xi = sIdx + (jj - 1);
  const float ddx = (float)(xi - NB) * dx - sd.sx;
  const float ddz = (float)(zi - NB) * dz - sd.sz;
  gptr(sd.T)[zi + (size_t)xi * nzz] = s_at(sd, nzz, zi, xi) * sqrtf(ddx * ddx + ddz * ddz);
}

// ---------------------------------------------------------------------------
// The reference schedule: one anti-diagonal `level` of sweep SW per launch,
// element e -> node (z_offset + zd*e, x_offset + xd*e), bounds-checked.
template <int SW, bool EQH>
__global__ __launch_bounds__(256) void k_sweep_plain(ShotDev sd, int nzz, int z_offset, int x_offset,
                                                     EikConsts k) {
  constexpr int zd = (SW == 2 || SW == 3) ? -1 : 1;
  constexpr int xd = (SW == 0 || SW == 2) ? -1 : 1;
  constexpr int a = (SW == 0 || SW == 2) ? 1 : -1;  // sgnt z
  constexpr int b = (SW == 0 || SW == 1) ? 1 : -1;  // sgnt x
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int i = z_offset + zd * e;
  const int j = x_offset + xd * e;
  const int nxx = sd.nxx;
  if ((i > 0) && (i < nzz - 1) && (j > 0) && (j < nxx - 1)) {
    const int i1 = i - (a + 1) / 2;
    const int j1 = j - (b + 1) / 2;
    gfloat* __restrict__ T = gptr(sd.T);
    const float tv = T[(i - a) + (size_t)j * nzz];
    const float te = T[i + (size_t)(j - b) * nzz];
    const float tev = T[(i - a) + (size_t)(j - b) * nzz];
    const float sref = s_at(sd, nzz, i1, j1);
    const float szo = s_at(sd, nzz, i1, j1 + b);
    const float sxo = s_at(sd, nzz, i1 + a, j1);
    const size_t idx = i + (size_t)j * nzz;
    T[idx] = vmin(T[idx], eik_local<EQH>(tv, te, tev, sref, szo, sxo, k));
  }
}

// ---------------------------------------------------------------------------
// Skewed-tile sweep.
//
// In sweep-local coordinates (local row kk = 0..NR-1 ordered along the
// sweep's z direction, local column cc = 0..NC-1 along its x direction;
// NR = nzz-2, NC = nxx-2 interior nodes) every node depends only on
// (kk-1,cc), (kk,cc-1), (kk-1,cc-1).  A tile is strip r (local rows 64r ..
// 64r+63, one per lane) by block bb: lane k visits cc = 64*bb + t - k at step
// t = 0..63, so all 64 lanes stay busy and lane k's upwind-z neighbour is the
// value lane k-1 produced one step earlier (a wave_shr:1 DPP move).  Tile
// (r, bb) needs (r, bb-1) and (r-1, bb+1) and runs in launch d = 2r + bb.
//
// LDS: the tile's old traveltimes and its slowness cells live in 64-slot
// rings indexed by column mod 64 (row stride 66 words, conflict free for the
// diagonal reads); results overwrite the old values in place and are written
// back column-pair by column-pair, so every global access touches at most two
// contiguous 256-byte column segments.
template <int SA, int SB>
__device__ __forceinline__ int node_off(int kk, int cc, int nzz, int nxx) {
  const int i = (SA > 0) ? 1 + kk : nzz - 2 - kk;
  const int j = (SB > 0) ? 1 + cc : nxx - 2 - cc;
  return i + j * nzz;
}

template <int SA, int SB>
__device__ __forceinline__ float cell_at(const ShotDev& sd, int nzz, int kap, int lam) {
  const int i = (SA > 0) ? kap : nzz - 2 - kap;
  const int j = (SB > 0) ? lam : sd.nxx - 2 - lam;
  return s_at(sd, nzz, i, j);
}

constexpr int RING_LD = 64;  // LDS row stride (words) of a tile's traveltimes
constexpr int SS_STEPS = 16; // steps per slowness staging phase of the non-skewed path
constexpr int PF = 16;       // slowness prefetch distance (steps) on the skewed path

// A tile's traveltimes in LDS are instruction-major: row v holds, for lane k,
// the node of step (v + k) & 63 (the order in which 64 coalesced column loads
// deliver them).  Lane k's node of step t is therefore at row (t - k) & 63,
// word k: a step's 64 accesses fall in distinct banks within each half-wave.
__device__ __forceinline__ int trow(int t, int lane) { return ((t - lane) & 63) * RING_LD + lane; }
// Word index of the same lane's node one step later (rows wrap mod 64).
__device__ __forceinline__ int trow_next(int ta) { return (ta + RING_LD) & (64 * RING_LD - 1); }
// The step loops walk the ring by byte offset (no index scaling per step).
__device__ __forceinline__ int trow_next_b(int tab) { return (tab + 4 * RING_LD) & (4 * 64 * RING_LD - 1); }
__device__ __forceinline__ float& ring_at(float* sT, int tab) {
  return *reinterpret_cast<float*>(reinterpret_cast<char*>(sT) + tab);
}
// Per-step broadcast values of a tile in LDS: word 2t = the upwind-row value
// lane 0 needs at step t, word 2t+1 = lane 63's x-other cell at step t (times
// dx when dx == dz); one 8-byte read per step serves all lanes.  Entry 64 is
// padding for the look-ahead read after the last step.
constexpr int HX_WORDS = 2 * 65;

// wave_shl:1 (DPP control 0x130): lane k receives lane k+1's `src`; lane 63
// keeps `old`.  Shifts the upwind row one step along: lane 0 holds step t's.
__device__ __forceinline__ float dpp_shl1(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), 0x130,
                                                    0xf, 0xf, false));
}

typedef __attribute__((address_space(3))) float lfloat;

// A tile is one wave, whose LDS operations execute in order: passing data
// between its lanes through LDS needs no barrier, only that the compiler keep
// the accesses in program order (and the LDS counter drained).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt untouched
  asm volatile("" ::: "memory");
}

// Slowness staging of the non-skewed path, SS_STEPS steps per phase: phase p's
// instruction u fetches, for lane k, its z-other cell of step
// t = p*SS_STEPS + ((u + k) mod SS_STEPS), cell (kap, c0 + 1 - k + t) (at most
// five contiguous column runs per instruction); stored step-major,
// sS[(t mod SS_STEPS)*64 + k], one bank per lane on both sides.
template <int SA, int SB>
__device__ __forceinline__ void ss_load(float (&v)[SS_STEPS], const ShotDev& sd, int nzz, int kap, int c0, int lane,
                                        int NC, int p) {
#pragma unroll
  for (int u = 0; u < SS_STEPS; ++u) {
    const int t = p * SS_STEPS + ((u + lane) & (SS_STEPS - 1));
    v[u] = cell_at<SA, SB>(sd, nzz, kap, min(max(c0 + 1 - lane + t, 0), NC));
  }
}
__device__ __forceinline__ void ss_store(const float (&v)[SS_STEPS], float* __restrict__ sS, int lane) {
#pragma unroll
  for (int u = 0; u < SS_STEPS; ++u) sS[((u + lane) & (SS_STEPS - 1)) * 64 + lane] = v[u];
}

template <int SA>
__device__ __forceinline__ int prow(int kk, int nzz) {
  return (SA > 0) ? 1 + kk : nzz - 2 - kk;
}
template <int SB>
__device__ __forceinline__ int pcol(int cc, int nxx) {
  return (SB > 0) ? 1 + cc : nxx - 2 - cc;
}

// Address (in floats, relative to SkewView::base) of sweep-local cell
// (kap, lam) of a shot in the skewed slowness copy; lam is clamped into the
// shot's cell range and mapped through the window's replicated border.
template <int SB>
__device__ __forceinline__ int skew_addr(const ShotDev& sd, int nxx, int NC, int KP, int kap, int lam) {
  lam = min(max(lam, 0), NC);
  const int j = (SB > 0) ? lam : nxx - 2 - lam;
  const int m = sd.x0 + sd.sofs + min(max(j - sd.sofs, 0), sd.smax);  // padded model column
  const int sig = (SB > 0) ? m + (kap & 63) : m - (kap & 63) + 63;
  return (sig + SK_OFF) * KP + kap;
}

// One step of a tile: lane k updates sweep-local node (kg, c0 + t - k).  The
// step's old traveltime, upwind-row value and slowness cells are loaded by the
// caller ahead of time (software pipelining across the solver's branches).
// ALLCOLS: every column the tile visits is an interior node column, so only
// the row test remains.
template <bool EQH, bool FD, bool ALLCOLS, bool XDPP = false, bool CONV = false>
__device__ __forceinline__ void tile_step(float* __restrict__ sT, int tab, int t, int lane, bool row_ok, int c0,
                                          int NC,
                                          float told, float h, float szo, float sxo, float& te, float& tev,
                                          CellQ& sref, const EikConsts& k, bool& chg) {
  const CellQ zo = cell_q<EQH, FD>(szo, k);  // this step's z-other cell, next step's reference cell
  const float tv = dpp_shr1(h, te);
  const int c = c0 + t - lane;
  const bool act = ALLCOLS ? row_ok : (row_ok && (unsigned)c < (unsigned)NC);
  // XDPP: lane k's x-other cell is lane k+1's z-other cell of this step; sxo
  // is then only lane 63's (times dx when EQH, where zo.dzs = dz*s = dx*s)
  const float dxs_xo = !XDPP ? k.dx * sxo : (EQH ? dpp_shl1(sxo, zo.dzs) : k.dx * dpp_shl1(sxo, szo));
  const float m = eik_local_sel<EQH, FD>(tv, te, tev, sref, zo.dzs, dxs_xo, k, told);  // min(told, candidates)
  const float res = act ? m : told;
  if constexpr (CONV) chg = chg | (res < told);  // converged mode: did this lane lower a traveltime
  ring_at(sT, tab) = res;
  tev = tv;
  te = res;
  sref = zo;
}

// The 64 steps of a tile with slowness from the skewed global copy: each lane
// loads its z-other cell PF steps ahead (none past step 63: every outstanding
// load holds one of the wave's 63 memory-counter slots that the tile's
// write-back and the next tile's staging need); its x-other cell is lane
// k+1's z-other cell of the same step (one DPP move), except lane 63's, which
// lies in the next strip: those 64 cells are gathered once (lane j holds step
// 63-j's) and rotate one lane per step.  FAST: the tile lies inside its
// window's unreplicated columns (so also inside the interior node columns),
// so at step t every lane's cell shares one skewed column and the addresses
// advance by a uniform stride; otherwise each lane computes its clamped
// address.
// Per-step slowness addressing of a tile in the skewed copy.  FAST: the
// z-other cell of step t sits at column sig0 + SB*t, a uniform base plus the
// lane's row (raw buffer loads: per-lane byte offset + per-step uniform
// offset, from a buffer resource based at the tile's lowest column, so the
// 32-bit offsets stay within 64 columns whatever the size of the copy);
// otherwise each lane computes its clamped address.
template <int SB, bool FAST>
struct SkewStream {
  const ShotDev& sd;
  const SkewView& sv;
  int kap, lane, c0, nxx, NC, colbase0;
  __amdgpu_buffer_rsrc_t rs;
  __device__ __forceinline__ SkewStream(const ShotDev& sd_, const SkewView& sv_, int kap_, int lane_, int c0_,
                                        int nxx_, int NC_)
      : sd(sd_), sv(sv_), kap(kap_), lane(lane_), c0(c0_), nxx(nxx_), NC(NC_) {
    const int sig0 = (SB > 0) ? sd.x0 + c0 + 1 : sd.x0 + sd.w + 66 - c0;
    colbase0 = (sig0 + SK_OFF) * sv.KP;
    if (FAST) {
      const size_t clo = (size_t)max(colbase0 - (SB > 0 ? 0 : 63 * sv.KP), 0);  // (edge tiles never load)
      rs = __builtin_amdgcn_make_buffer_rsrc((void*)(sv.base + clo), 0, 0x7fffffff, 0x00020000);
    }
  }
  __device__ __forceinline__ float z(int t) const {
    // column of step t relative to the resource's base column
    if (FAST)
      return __uint_as_float(
          __builtin_amdgcn_raw_buffer_load_b32(rs, kap * 4, ((SB > 0) ? t : 63 - t) * sv.KP * 4, 0));
    return gptr(sv.base)[skew_addr<SB>(sd, nxx, NC, sv.KP, kap, c0 + t - lane + 1)];
  }
  // lane 63's x-other cells (row kap63+1 = next strip, one skew period away),
  // lane j holding step 63-j's
  __device__ __forceinline__ float x63() const {
    const int kx = kap - lane + 64;
    if (FAST) return gptr(sv.base)[colbase0 + kx - SB * sv.KP * (lane + 1)];
    return gptr(sv.base)[skew_addr<SB>(sd, nxx, NC, sv.KP, kx, c0 - lane)];
  }
};

// A tile's first slowness loads, issued before its dependency wait: the
// z-other cells of steps 0..PF-1, lane 63's x-other cells and the reference
// cell of step 0.
struct SkewPre {
  float zq[PF];
  float xr, s0;
};
template <int SB, bool FAST>
__device__ __forceinline__ void skew_prefetch(SkewPre& p, const SkewStream<SB, FAST>& st) {
#pragma unroll
  for (int u = 0; u < PF; ++u) p.zq[u] = st.z(u);
  p.xr = st.x63();
  p.s0 = gptr(st.sv.base)[skew_addr<SB>(st.sd, st.nxx, st.NC, st.sv.KP, st.kap, st.c0 - st.lane)];
}

template <int SB, bool EQH, bool FD, bool FAST, bool CONV = false>
__device__ __forceinline__ void tile_steps_skew(const SkewStream<SB, FAST>& st, SkewPre& pre, float* __restrict__ sT,
                                                const float* __restrict__ sHX, int lane, bool row_ok, int c0,
                                                int NC, float& te, float& tev, CellQ& sref, const EikConsts& k,
                                                bool& chg) {
  float* zq = pre.zq;
  auto load_z = [&](int t) -> float { return st.z(t); };
  int tab = 4 * trow(0, lane);
  float told_n = ring_at(sT, tab);
  float2 hx_n = reinterpret_cast<const float2*>(sHX)[0];
  // PF steps from t0; the last block is a separate copy without loads, so
  // nothing is fetched for steps past 63
  auto block = [&](int t0, auto fetch) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int t = t0 + u;
      const float told = told_n;
      const float2 hx = hx_n;
      const int tn = trow_next_b(tab);
      told_n = ring_at(sT, tn);
      hx_n = reinterpret_cast<const float2*>(sHX)[t + 1];
      const float szo = zq[u];
      if constexpr (decltype(fetch)::value) zq[u] = load_z(t + PF);
      tile_step<EQH, FD, FAST, true, CONV>(sT, tab, t, lane, row_ok, c0, NC, told, hx.x, szo, hx.y, te, tev, sref,
                                           k, chg);
      tab = tn;
    }
  };
  for (int t0 = 0; t0 < 64 - PF; t0 += PF) block(t0, std::true_type{});
  block(64 - PF, std::false_type{});
}

// Tag words of the persistent schedule's boundary exchange: value bits in the
// low half, (epoch << 11 | sweep << 8 | block & 255) in the high half, written
// and read as single 64-bit agent-scope atomics so a consumer that sees the
// expected tag also sees the value (no fences needed).
__device__ __forceinline__ unsigned long long xpack(float v, unsigned tag) {
  return ((unsigned long long)tag << 32) | (unsigned long long)__float_as_uint(v);
}
// global address space: generic (flat) accesses may alias LDS, and the
// compiler would order each one after the tile's LDS-destination loads
typedef __attribute__((address_space(1))) unsigned long long gword;
__device__ __forceinline__ void xstore(unsigned long long* p, float v, unsigned tag) {
  __hip_atomic_store((gword*)p, xpack(v, tag), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
constexpr int SPIN_LIMIT = 1 << 16;  // ~0.1 s of polling (1000x a tile's run time); then flag and go on

// Waits until every lane sees, for each of a tile's four dependency words
// with `need`, the word's `tag`; returns the values (1e6 where not needed).
// Every poll issues all outstanding loads before testing any tag, so a tile
// whose predecessors are done pays one memory round trip, not four.  The wait
// is bounded: past SPIN_LIMIT polls it sets *err and returns whatever was
// read, so a broken dependency cannot hang the device.
struct XDep {
  const unsigned long long* p;
  bool need;
  unsigned tag;
};
__device__ __forceinline__ void xwait4(const XDep (&d)[4], float (&out)[4], int* err) {
  unsigned long long w[4];
  bool ok[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    w[q] = xpack(TINF, d[q].tag);
    ok[q] = !d[q].need;
  }
  for (int it = 0;; ++it) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (!ok[q]) w[q] = __hip_atomic_load((const gword*)d[q].p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int q = 0; q < 4; ++q) ok[q] = ok[q] || (unsigned)(w[q] >> 32) == d[q].tag;
    if (__all(ok[0] && ok[1] && ok[2] && ok[3])) break;
    if (it > SPIN_LIMIT) {
      if (threadIdx.x == 0) atomicOr(err, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = d[q].need ? __uint_as_float((unsigned)w[q]) : TINF;
}

// Control block of a persistent sweep launch.
struct PersistCtl {
  int* counter;            // tile ticket counter of this launch (zeroed by k_begin_run)
  const unsigned* epoch;   // run counter (advanced by k_begin_run)
  int* err;                // spin-limit flag
  int ntiles;
  int nstrips;             // strips per shot (NRS)
  unsigned sweep_id;       // 0..7 within a run
  unsigned xtag;           // fused pairs: tag of the partner sweep's done words
  // converged mode: set to 1 by every tile that lowers a traveltime, and the
  // word that makes the launch exit at once (a pass after convergence)
  int* changed;
  const int* skip;
  // tiles not run because all their inputs are TINF (counted when set)
  int* nskip;
  // host side: persistent workgroups per CU (0 = as many as fit)
  int wpcu;
};

// One 64x64 skewed tile of a sweep.
//
// In sweep-local coordinates (local row kk = 0..NR-1 along the sweep's z
// direction, local column cc = 0..NC-1 along its x direction; NR = nzz-2,
// NC = nxx-2 interior nodes) each node depends only on (kk-1,cc), (kk,cc-1)
// and (kk-1,cc-1), and uses slowness cells (kk,cc), (kk,cc+1), (kk+1,cc).  A
// tile is strip r (local rows 64r .. 64r+63, one per lane) by block bb: lane k
// visits cc = 64*bb + t - k at step t = 0..63, so all 64 lanes stay busy and
// lane k's z-upwind neighbour is the value lane k-1 produced one step earlier
// (a wave_shr:1 DPP move).  Tile (r, bb) needs (r, bb-1) and (r-1, bb+1).
//
// LDS holds the tile's old traveltimes instruction-major (see trow), loaded by
// 64 column loads straight into LDS; results overwrite them in place and are
// written back the same way, so every global access touches at most two
// contiguous 256-byte segments.  SKEW: slowness comes from the skewed global
// copy (one coalesced load per cell per step, prefetched PF steps ahead);
// otherwise it is staged step-major in LDS from the natural layout,
// SS_STEPS steps at a time (sS[t mod SS_STEPS][kap] = cell
// (64r+kap, c0+t-kap+1)).  FIRST: the launch is the first sweep of a
// solve, whose old traveltimes are time_set + time_init (1e6 except the 3x3
// source stencil) and are generated here instead of read.  PERSIST: the
// upwind row and the carry-in from block bb-1 arrive through the tagged
// exchange buffers of the shot (and the tile publishes its own), instead of
// being read from T after a launch boundary.
//
// XMODE (fused sweep pairs, whose two sweeps differ only in the row
// direction): 1 = the pair's first sweep, which marks each tile done once its
// write-back has reached L2; 2 = the second sweep, which before staging waits
// for the done words of the first-sweep tiles holding its nodes and then
// stages with agent-scope loads (cache policy sc1, as the compiler issues
// agent-scope atomic loads), which do not hit L1 lines this CU may hold from
// earlier tiles.  Both sweeps of a shot run on one XCD, whose L2 is shared,
// so completion at L2 is all such a reader needs, and no L1 invalidation
// (which would also drop the other waves' lines) is.
template <int SA, int SB, bool EQH, bool FD, bool SKEW, bool FIRST, bool PERSIST, int XMODE, bool CONV,
          class AfterStage, class BeforeWriteBack>
__device__ __forceinline__ void tile_process(const int4 tl, const ShotDev* __restrict__ shots, int nzz,
                                             const EikConsts& k, const SkewView& sv, const PersistCtl& ctl,
                                             unsigned tagb, float* __restrict__ sT, float* __restrict__ sS,
                                             float* __restrict__ sHX, const AfterStage& after_stage,
                                             const BeforeWriteBack& before_wb) {
  const ShotDev sd = shots[tl.x];
  const int r = tl.y;
  const int bb = tl.z;
  const int c0 = 64 * bb;
  const int lane = threadIdx.x;
  const int nxx = sd.nxx;
  const int NR = nzz - 2, NC = nxx - 2;
  const int kg = 64 * r + lane;
  const bool row_ok = kg < NR;
  gfloat* __restrict__ T = gptr(sd.T);
  const int kgc = min(kg, NR);
  // interior tile: every column it stages or writes lies in [0, NC)
  const bool interior = c0 >= 64 && c0 + 64 <= NC;
  const int jb = (SB > 0) ? 1 + c0 : nxx - 2 - c0;  // grid column of local column c0
  constexpr int STAGE_CPOL = XMODE == 2 ? 16 : 0;   // cache policy of the staging loads: sc1 (agent scope)

  // old traveltimes known to be all TINF: a first sweep's tile without the
  // source stencil, or (pairs) a second-sweep tile whose first-sweep tiles
  // beneath all skipped (done word value 1)
  bool told_tinf = false;
  if (XMODE == 2) {
    // The partner sweep runs the rows the other way, columns the same: this
    // tile's local rows kk1 are its rows NR-1-kk1 (strips r0, at most two),
    // and its node (kk0, cc) lies in block floor((cc + kk0 mod 64) / 64).
    // Over this tile, cc + kk0 - 64 r0 = c0 + t - 2k + C (t, k in 0..63), so
    // each strip needs at most four consecutive blocks: lane j < 8 waits for
    // strip r0lo + j/4, block blo + j%4.
    const int kk1_hi = min(64 * r + 63, NR - 1);
    const int kk0_lo = NR - 1 - kk1_hi, kk0_hi = NR - 1 - 64 * r;
    const int r0 = (kk0_lo >> 6) + (lane >> 2);
    bool need = false;
    const unsigned long long* fp = sd.xd;
    if (lane < 8 && r0 <= (kk0_hi >> 6)) {
      const int klo = NR - 1 - 64 * r - min(kk0_hi, 64 * r0 + 63);
      const int khi = NR - 1 - 64 * r - max(kk0_lo, 64 * r0);
      const int C = NR - 1 - 64 * (r + r0);
      const int blo = max((c0 - 2 * khi + C) >> 6, 0);
      const int bhi = min((c0 + 63 - 2 * klo + C) >> 6, sd.nb - 1);
      const int b0 = blo + (lane & 3);
      need = b0 <= bhi;
      fp = sd.xd + (size_t)r0 * sd.nb + min(b0, sd.nb - 1);
    }
    const XDep deps[4] = {{fp, need, ctl.xtag}, {fp, false, 0u}, {fp, false, 0u}, {fp, false, 0u}};
    float dv[4];
    xwait4(deps, dv, ctl.err);
    told_tinf = __all(!need || dv[0] == 1.0f);
  }

  // ---- old traveltimes, step-major --------------------------------------
  if (FIRST) {
    const int sIdx = (int)((sd.sx + 0.5f * k.dx) / k.dx) + NB;
    const int sIdz = (int)((sd.sz + 0.5f * k.dz) / k.dz) + NB;
    // uniform test: does any of the 9 stencil nodes fall in this tile?
    bool has_src = false;
    for (int q = 0; q < 9; ++q) {
      const int i = sIdz + q % 3 - 1, j = sIdx + q / 3 - 1;
      const int kk = (SA > 0) ? i - 1 : nzz - 2 - i;
      const int cc = (SB > 0) ? j - 1 : nxx - 2 - j;
      const int kl = kk - 64 * r, tt = cc - c0 + kl;
      if (kl >= 0 && kl < 64 && tt >= 0 && tt < 64) has_src = true;
    }
    const int i = prow<SA>(kg, nzz);
    for (int t = 0; t < 64; ++t) {
      float val = TINF;
      if (has_src) {
        const int c = c0 + t - lane;
        const int j = pcol<SB>(c, nxx);
        if (row_ok && c >= 0 && c < NC && abs(i - sIdz) <= 1 && abs(j - sIdx) <= 1) {
          const float ddx = (float)(j - NB) * k.dx - sd.sx;
          const float ddz = (float)(i - NB) * k.dz - sd.sz;
          val = s_at(sd, nzz, i, j) * sqrtf(ddx * ddx + ddz * ddz);
        }
      }
      sT[trow(t, lane)] = val;
    }
    told_tinf = !has_src;
  } else {
    // Column loads straight into LDS: instruction v fetches, for lane k, the
    // node of step (v+k)&63, column c0-k+((v+k)&63) = c0+v or c0+v-64 (two
    // contiguous column segments).  Edge tiles clamp positions outside the
    // interior onto the boundary ring (always 1e6).
    if (interior) {
      const int irow = (SA > 0) ? 1 + kgc : nzz - 2 - kgc;
      const unsigned o0 = (unsigned)(irow + (SB > 0 ? 64 * nzz : 0));
      const unsigned o1 = (unsigned)(irow + (SB > 0 ? 0 : 64 * nzz));
      const gfloat* Tb = T + (jb - (SB > 0 ? 64 : 0)) * nzz;
#pragma unroll 2
      for (int v = 0; v < 64; ++v) {
        __builtin_amdgcn_global_load_lds(Tb + (lane >= 64 - v ? o1 : o0), (lfloat*)(sT + v * RING_LD), 4, 0,
                                         STAGE_CPOL);
        Tb += SB * nzz;
      }
    } else {
#pragma unroll 2
      for (int v = 0; v < 64; ++v) {
        const int c = min(max(c0 - lane + ((v + lane) & 63), -1), NC);
        __builtin_amdgcn_global_load_lds(T + (unsigned)node_off<SA, SB>(kgc, c, nzz, nxx),
                                         (lfloat*)(sT + v * RING_LD), 4, 0, STAGE_CPOL);
      }
    }
    // (waited for below, after the halo and carry loads: their latencies overlap)
  }
  after_stage();

  // skewed slowness: the tile's first loads go out before the dependency wait
  const int kap = 64 * r + lane;
  // unreplicated cell columns: lam in [3, w+2] (SB > 0) or [2, w+1] (SB < 0)
  const int lam_lo = max(0, c0 - 63), lam_hi = min(NC, c0 + 64);
  const bool fast = SKEW && ((SB > 0) ? (lam_lo >= 3 && lam_hi <= sd.w + 2) : (lam_lo >= 2 && lam_hi <= sd.w + 1));
  const SkewStream<SB, true> st_fast(sd, sv, kap, lane, c0, nxx, NC);
  const SkewStream<SB, false> st_edge(sd, sv, kap, lane, c0, nxx, NC);
  SkewPre pre;
  // natural-layout slowness: phase 0's cells and lane 63's x-other cells
  // (lane j: step j's, cell (64r+64, c0-63+j))
  float sv_[SS_STEPS];
  float sx63 = 0.0f;
  if (SKEW) {
    if (fast)
      skew_prefetch(pre, st_fast);
    else
      skew_prefetch(pre, st_edge);
  } else {
    ss_load<SA, SB>(sv_, sd, nzz, kgc, c0, lane, NC, 0);
    sx63 = cell_at<SA, SB>(sd, nzz, min(64 * r + 64, NR), min(max(c0 - 63 + lane, 0), NC));
  }

  // ---- upwind row above the strip (local row 64r-1; lane t holds column
  // c0+t, the value lane 0 needs at step t), and the carry-in from block
  // bb-1: te = T(kg, c0-k-1), tev = T(kg-1, c0-k-1).
  float te, tev, hrow;
  if (PERSIST) {
    // halo: strip r-1's bottom row; lane t fetches column c0+t, lane 0 also c0-1
    const bool hneed = r > 0 && c0 + lane < NC;
    const unsigned long long* xh = sd.xh + (size_t)max(r - 1, 0) * sd.ldh;
    const bool h1need = r > 0 && bb > 0 && lane == 0 && c0 - 1 < NC;
    // carries of block bb-1: word 2k = lane k's step-63 result, 2k+1 = step 62
    const bool cneed = bb > 0;
    const unsigned long long* xc = sd.xc + (size_t)(2 * r + ((bb - 1) & 1)) * 128;
    const unsigned ctag = tagb | ((unsigned)(bb - 1) & 255u);
    const XDep deps[4] = {{xh + min(c0 + lane, NC - 1), hneed, tagb},
                          {xh + min(max(c0 - 1, 0), NC - 1), h1need, tagb},
                          {xc + 2 * lane, cneed, ctag},
                          {xc + 2 * max(lane - 1, 0) + 1, cneed && lane > 0, ctag}};
    float dv[4];
    xwait4(deps, dv, ctl.err);
    hrow = dv[0];
    te = dv[2];
    tev = (lane == 0) ? dv[1] : dv[3];
  } else {
    hrow = T[node_off<SA, SB>(64 * r - 1, min(c0 + lane, NC), nzz, nxx)];
    const int cin = min(max(c0 - lane - 1, -1), NC);
    te = T[node_off<SA, SB>(kgc, cin, nzz, nxx)];
    tev = T[node_off<SA, SB>(min(kg - 1, NR), cin, nzz, nxx)];
  }

  if (!FIRST) __builtin_amdgcn_s_waitcnt(0);  // vmcnt(0), lgkmcnt(0): the column loads have landed in LDS

  // A tile whose inputs are all TINF (old traveltimes, upwind row, carry-in)
  // leaves every node at TINF where the geometry keeps such a neighbourhood
  // at TINF (k.tinf_fixed): its steps are not run, as in a first sweep the
  // region the front has not reached.  A skipped second-sweep tile also
  // leaves T as it is; a first sweep writes its TINF values back.
  const bool skip = k.tinf_fixed && (FIRST || XMODE == 2) && told_tinf &&
                    __all((hrow == TINF) & (te == TINF) & (tev == TINF));
  if (skip && ctl.nskip && lane == 0) atomicAdd(ctl.nskip, 1);

  bool chg = false;  // converged mode: this lane lowered a traveltime
  if (skip) {
  } else if (SKEW) {
    // per-step broadcasts: lane t's upwind-row value is step t's; lane j's
    // x-other cell of lane 63 is step 63-j's
    sHX[2 * lane] = hrow;
    CellQ sref = cell_q<EQH, FD>(pre.s0, k);
    sHX[2 * (63 - lane) + 1] = EQH ? k.dx * pre.xr : pre.xr;
    wave_lds_sync();
    if (fast)
      tile_steps_skew<SB, EQH, FD, true, CONV>(st_fast, pre, sT, sHX, lane, row_ok, c0, NC, te, tev, sref, k, chg);
    else
      tile_steps_skew<SB, EQH, FD, false, CONV>(st_edge, pre, sT, sHX, lane, row_ok, c0, NC, te, tev, sref, k, chg);
  } else {
    // Slowness in SS_STEPS-step phases (the next phase's loads in flight
    // while this one runs); lane k's x-other cell is lane k+1's z-other cell
    // of the same step (one DPP move), lane 63's comes from the broadcasts.
    constexpr int NPH = 64 / SS_STEPS;
    sHX[2 * lane] = hrow;
    CellQ sref = cell_q<EQH, FD>(cell_at<SA, SB>(sd, nzz, kgc, min(max(c0 - lane, 0), NC)), k);
    ss_store(sv_, sS, lane);
    sHX[2 * lane + 1] = EQH ? k.dx * sx63 : sx63;
    int tab = 4 * trow(0, lane);
    float told_n = ring_at(sT, tab);
    float2 hx_n = make_float2(0.0f, 0.0f);
    for (int ph = 0; ph < NPH; ++ph) {
      if (ph + 1 < NPH) ss_load<SA, SB>(sv_, sd, nzz, kgc, c0, lane, NC, ph + 1);
      wave_lds_sync();
      float z_n = sS[lane];
      if (ph == 0) hx_n = reinterpret_cast<const float2*>(sHX)[0];
      for (int t0 = ph * SS_STEPS; t0 < (ph + 1) * SS_STEPS; t0 += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int t = t0 + u;
          const float told = told_n, szo = z_n;
          const float2 hx = hx_n;
          const int tn = trow_next_b(tab);
          told_n = ring_at(sT, tn);
          hx_n = reinterpret_cast<const float2*>(sHX)[t + 1];
          z_n = sS[((t + 1) & (SS_STEPS - 1)) * 64 + lane];
          tile_step<EQH, FD, false, true, CONV>(sT, tab, t, lane, row_ok, c0, NC, told, hx.x, szo, hx.y, te, tev,
                                                sref, k, chg);
          tab = tn;
        }
      }
      if (ph + 1 < NPH) ss_store(sv_, sS, lane);
    }
  }
  wave_lds_sync();
  // converged mode: one wave vote per tile, one global word per pass
  if constexpr (CONV) {
    if (__any(chg) && lane == 0) atomicOr(ctl.changed, 1);
  }

  if (PERSIST) {
    // publish: lane 63's results are strip r+1's upwind row (columns c0+t-63),
    // each lane's last two results are block bb+1's carry-in.
    if (r + 1 < ctl.nstrips) {
      const int c = c0 + lane - 63;
      if (c >= 0 && c < NC) xstore(sd.xh + (size_t)r * sd.ldh + c, sT[trow(lane, 63)], tagb);
    }
    unsigned long long* xc = sd.xc + (size_t)(2 * r + (bb & 1)) * 128;
    const unsigned ctag = tagb | ((unsigned)bb & 255u);
    xstore(xc + 2 * lane, sT[trow(63, lane)], ctag);
    xstore(xc + 2 * lane + 1, sT[trow(62, lane)], ctag);
  }

  before_wb();
  // write-back, row v of LDS = the nodes of column loads v
  if (skip && XMODE == 2) {
    // T already holds these TINF values
  } else if (interior) {
    const int irow = (SA > 0) ? 1 + kgc : nzz - 2 - kgc;
    const unsigned o0 = (unsigned)(irow + (SB > 0 ? 64 * nzz : 0));
    const unsigned o1 = (unsigned)(irow + (SB > 0 ? 0 : 64 * nzz));
    gfloat* Tb = T + (jb - (SB > 0 ? 64 : 0)) * nzz;
#pragma unroll 4
    for (int v = 0; v < 64; ++v) {
      const float val = sT[v * RING_LD + lane];
      if (row_ok) Tb[lane >= 64 - v ? o1 : o0] = val;
      Tb += SB * nzz;
    }
  } else {
#pragma unroll 4
    for (int v = 0; v < 64; ++v) {
      const int c = c0 - lane + ((v + lane) & 63);
      const float val = sT[v * RING_LD + lane];
      if (row_ok && c >= 0 && c < NC) T[(unsigned)node_off<SA, SB>(kg, c, nzz, nxx)] = val;
    }
  }
  if (XMODE == 1) {
    __builtin_amdgcn_s_waitcnt(0);  // the write-back stores have reached L2
    // value 1: every node of the tile is TINF (it was skipped)
    if (lane == 0) xstore(sd.xd + (size_t)r * sd.nb + bb, skip ? 1.0f : 0.0f, tagb);
  }
}

// One tile per workgroup; launch d runs the tiles of tile diagonal d = 2r+bb.
template <int SA, int SB, bool EQH, bool SKEW, bool FIRST>
__global__ __launch_bounds__(64) void k_sweep_tiles(const ShotDev* __restrict__ shots,
                                                    const int4* __restrict__ tiles, int nzz, EikConsts k,
                                                    SkewView sv) {
  __shared__ float sT[64 * RING_LD];
  __shared__ float sS[SKEW ? 1 : SS_STEPS * 64];
  __shared__ __attribute__((aligned(8))) float sHX[HX_WORDS];
  const PersistCtl ctl{};
  tile_process<SA, SB, EQH, false, SKEW, FIRST, false, 0, false>(tiles[blockIdx.x], shots, nzz, k, sv, ctl, 0u, sT,
                                                                 sS, sHX,
                                                       [] {}, [] {});
}

// Persistent sweep: every workgroup claims tiles from a ticket counter in
// tile-diagonal order until the sweep's tiles run out.  A tile only waits on
// tiles with smaller tickets, which running workgroups hold, so the sweep
// always drains.
template <int SA, int SB, bool EQH, bool FD, bool SKEW, bool FIRST, bool CONV = false>
__global__ __launch_bounds__(64) void k_sweep_persist(const ShotDev* __restrict__ shots,
                                                      const int4* __restrict__ tiles, int nzz, EikConsts k,
                                                      SkewView sv, PersistCtl ctl) {
  __shared__ float sT[64 * RING_LD];
  __shared__ float sS[SKEW ? 1 : SS_STEPS * 64];
  __shared__ __attribute__((aligned(8))) float sHX[HX_WORDS];
  // converged mode: a pass after convergence does nothing
  if constexpr (CONV) {
    if (__hip_atomic_load(ctl.skip, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  }
  const unsigned tagb = (*ctl.epoch << 11) | (ctl.sweep_id << 8);
  auto claim_issue = [&]() -> int {
    int t = 0;
    if (threadIdx.x == 0) t = atomicAdd(ctl.counter, 1);
    return t;
  };
  int tk = __builtin_amdgcn_readfirstlane(claim_issue());
  while (tk < ctl.ntiles) {
    const int4 tl = tiles[tk];
    // wave priority from the tile's scheduling slack (0 = most slack)
    switch (tl.w) {
      case 3: __builtin_amdgcn_s_setprio(3); break;
      case 2: __builtin_amdgcn_s_setprio(2); break;
      case 1: __builtin_amdgcn_s_setprio(1); break;
      default: __builtin_amdgcn_s_setprio(0); break;
    }
    // the next ticket is claimed as soon as this tile's staging is issued
    // (its round trip overlaps the dependency wait) and read before the
    // write-back.  Holding one ticket ahead keeps the sweep draining: each
    // workgroup's tickets increase, so the smallest unfinished ticket is
    // always one a workgroup is running, and its predecessors are done.
    int raw = 0, next = 0;
    tile_process<SA, SB, EQH, FD, SKEW, FIRST, true, 0, CONV>(
        tl, shots, nzz, k, sv, ctl, tagb, sT, sS, sHX, [&] { raw = claim_issue(); },
        [&] { next = __builtin_amdgcn_readfirstlane(raw); });
    tk = next;
  }
}

// Fused sweep pair: the two sweeps (1, SB) and (-1, SB) of a solve in one
// persistent launch.  Each shot belongs to one XCD (the workgroup's XCC id,
// read from the hardware register, picks its queue): its tiles of both sweeps
// are claimed there in one ticket order in which every tile follows the tiles
// it waits for, so a second-sweep tile starts as soon as the first-sweep
// tiles under it are written back, not when the whole first sweep is done.
// Per-XCD queues drain for the same reason a single queue does.
template <int SB, bool EQH, bool FD, bool FIRST_A, bool SKEW>
__global__ __launch_bounds__(64) void k_sweep_pair(const ShotDev* __restrict__ shots_a,
                                                   const ShotDev* __restrict__ shots_b,
                                                   const int4* __restrict__ tiles, const int* __restrict__ qoff,
                                                   int nzz, EikConsts k, SkewView sv_a, SkewView sv_b,
                                                   PersistCtl ctl_a, PersistCtl ctl_b) {
  __shared__ float sT[64 * RING_LD];
  __shared__ float sS[SKEW ? 1 : SS_STEPS * 64];
  __shared__ __attribute__((aligned(8))) float sHX[HX_WORDS];
  const int xcd = __builtin_amdgcn_s_getreg((15 << 11) | 20) & 7;  // hwreg(HW_REG_XCC_ID, 0, 16)
  const unsigned tag_a = (*ctl_a.epoch << 11) | (ctl_a.sweep_id << 8);
  const unsigned tag_b = (*ctl_b.epoch << 11) | (ctl_b.sweep_id << 8);
  PersistCtl cb = ctl_b;
  cb.xtag = tag_a;
  int* const counter = ctl_a.counter + xcd;
  const int base = qoff[xcd], n = qoff[xcd + 1] - base;
  auto claim_issue = [&]() -> int {
    int t = 0;
    if (threadIdx.x == 0) t = atomicAdd(counter, 1);
    return t;
  };
  int tk = __builtin_amdgcn_readfirstlane(claim_issue());
  while (tk < n) {
    const int4 tl = tiles[base + tk];
    switch (tl.w & 3) {
      case 3: __builtin_amdgcn_s_setprio(3); break;
      case 2: __builtin_amdgcn_s_setprio(2); break;
      case 1: __builtin_amdgcn_s_setprio(1); break;
      default: __builtin_amdgcn_s_setprio(0); break;
    }
    int raw = 0, next = 0;
    if (tl.w & 4)
      tile_process<-1, SB, EQH, FD, SKEW, false, true, 2, false>(
          tl, shots_b, nzz, k, sv_b, cb, tag_b, sT, sS, sHX, [&] { raw = claim_issue(); },
          [&] { next = __builtin_amdgcn_readfirstlane(raw); });
    else
      tile_process<1, SB, EQH, FD, SKEW, FIRST_A, true, 1, false>(
          tl, shots_a, nzz, k, sv_a, ctl_a, tag_a, sT, sS, sHX, [&] { raw = claim_issue(); },
          [&] { next = __builtin_amdgcn_readfirstlane(raw); });
    tk = next;
  }
}

// Converged mode's control words: [0] skip (set once a pass left every
// traveltime unchanged, or the pass limit is reached), [1] passes run,
// [2], [3] the change words of even and odd passes, [4] index of the next
// pass, [5] the pass limit.  Start of a run: pass 0 is about to run.
__global__ void k_conv_init(int* __restrict__ cw, int max_passes) {
  if (threadIdx.x == 0) {
    cw[0] = 0;
    cw[1] = 1;
    cw[2] = 0;
    cw[3] = 0;
    cw[4] = 1;
    cw[5] = max_passes;
  }
}

// Before pass p = cw[4]: skip it (and every later one) if pass p-1 lowered
// no traveltime or the limit is reached; otherwise count it and clear its
// change word.
__global__ void k_conv_gate(int* __restrict__ cw) {
  if (threadIdx.x == 0) {
    const int p = cw[4];
    if (!cw[0]) {
      if (cw[1] >= cw[5] || cw[2 + ((p - 1) & 1)] == 0) {
        cw[0] = 1;
      } else {
        cw[2 + (p & 1)] = 0;
        cw[1] += 1;
      }
    }
    cw[4] = p + 1;
  }
}

// XCC id of each of n workgroups (the fused pairs rely on workgroups going
// round robin over 8 XCCs; checked once per geometry).
__global__ void k_xcc_ids(int* __restrict__ out) {
  if (threadIdx.x == 0) out[blockIdx.x] = (int)(__builtin_amdgcn_s_getreg((15 << 11) | 20) & 0xffff);
}

// Start of a run: advance the epoch (tags of the previous run become stale)
// and zero the ticket counters.
__global__ void k_begin_run(int* __restrict__ counters, int n, unsigned* __restrict__ epoch) {
  const int i = threadIdx.x;
  if (i < n) counters[i] = 0;
  if (i == 0) {
    *epoch = (*epoch + 1) % 2097151u;
    counters[CTL_SKIP] = 0;
  }
}

// ---------------------------------------------------------------------------
// Catmull-Rom receiver picks: one thread per receiver (positions relative to
// the receiver's shot grid origin).
__device__ __forceinline__ float cubic1d(const float P[4], float d) {
  return P[1] + 0.5f * d *
                    (P[2] - P[0] +
                     d * (2.0f * P[0] - 5.0f * P[1] + 4.0f * P[2] - P[3] +
                          d * (3.0f * (P[1] - P[2]) + P[3] - P[0])));
}

// rec_shot null: every receiver belongs to shot 0.  A receiver whose 4x4
// stencil leaves its shot's grid (possible only for receiver arrays the host
// has not checked) gets NaN and sets *err (when err is not null) to 2.
__global__ void k_picks(const ShotDev* __restrict__ shots, const int* __restrict__ rec_shot,
                        const float* __restrict__ rx, const float* __restrict__ rz,
                        float* __restrict__ out, int nrec, int nzz, float dx, float dz, int* __restrict__ err) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const ShotDev& sd = shots[rec_shot ? rec_shot[r] : 0];
  const gfloat* __restrict__ T = gptr(sd.T);
  const float x = rx[r], z = rz[r];
  const float x0 = floorf((x + 0.5f * dx) / dx) * dx;
  const float z0 = floorf((z + 0.5f * dz) / dz) * dz;
  const float x1 = floorf((x + 0.5f * dx) / dx) * dx + dx;
  const float z1 = floorf((z + 0.5f * dz) / dz) * dz + dz;
  const float xd = (x - x0) / (x1 - x0);
  const float zd = (z - z0) / (z1 - z0);
  // position tests before the int conversions (a NaN or huge float has no int)
  const float fi = (z + 0.5f * dz) / dz, fj = (x + 0.5f * dx) / dx;
  if (!(fi >= 0.0f && fi < (float)(nzz - 2 * NB - 1) + 1.0f && fj >= 0.0f &&
        fj < (float)(sd.nxx - 2 * NB - 1) + 1.0f)) {
    out[r] = __builtin_nanf("");
    if (err) atomicOr(err, 2);
    return;
  }
  const int i = (int)fi + NB;
  const int j = (int)fj + NB;
  float p[4];
  for (int pIdx = 0; pIdx < 4; ++pIdx) {
    float P[4];
    for (int pIdz = 0; pIdz < 4; ++pIdz) P[pIdz] = T[(i + pIdz - 1) + (size_t)(j + pIdx - 1) * nzz];
    p[pIdx] = cubic1d(P, zd);
  }
  out[r] = cubic1d(p, xd);
}

// ---------------------------------------------------------------------------
// qP slowness for every padded node of every shot (blockIdx.y = shot).  Shots
// of an anisotropic batch cover the whole model, so node index = model index.
// Interior nodes except the source get the Christoffel qP slowness; the rest
// keep the isotropic value (src/modeling/eikonal_ani.cu:63-66).
__device__ __forceinline__ float decode_c(uint16_t u, float mn, float mx) {
  return (mn + (static_cast<float>(u) - 1.0f) * (mx - mn) / (65535 - 1));
}

// The shot-independent part of get_quasi_slowness: per padded node, the
// decoded stiffness scaled by Ro = C33 * s^2 (Cij / Ro / Ro) and Ro itself,
// evaluated with the same operations as the per-node expressions, once per
// run instead of once per shot.
__global__ void k_ani_table(ShotDev model, int nzz, int nxx, StiffDev st, AniTab tb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nzz || j >= nxx) return;
  const size_t index = i + (size_t)j * nzz;
  const float s_iso = s_at(model, nzz, i, j);
  const float c11 = decode_c(st.c[0][index], st.mn[0], st.mx[0]);
  const float c13 = decode_c(st.c[1][index], st.mn[1], st.mx[1]);
  const float c15 = decode_c(st.c[2][index], st.mn[2], st.mx[2]);
  const float c33 = decode_c(st.c[3][index], st.mn[3], st.mx[3]);
  const float c35 = decode_c(st.c[4][index], st.mn[4], st.mx[4]);
  const float c55 = decode_c(st.c[5][index], st.mn[5], st.mx[5]);
  const float Ro = c33 * s_iso * s_iso;
  tb.v[0][index] = c11 / Ro / Ro;
  tb.v[1][index] = c13 / Ro / Ro;
  tb.v[2][index] = c15 / Ro / Ro;
  tb.v[3][index] = c33 / Ro / Ro;
  tb.v[4][index] = c35 / Ro / Ro;
  tb.v[5][index] = c55 / Ro / Ro;
  tb.v[6][index] = Ro;
}

// Per shot and padded node (src/modeling/eikonal_ani.cu:47-111): a thread
// takes one node for QP_SHOTS shots, so the shot-independent factors are read
// once for all of them.
constexpr int QP_SHOTS = 4;
__global__ void k_quasi_slowness(const ShotDev* __restrict__ iso, ShotDev* __restrict__ qp, int nshots, int nzz,
                                 float dx, float dz, AniTab tb) {
  const int nxx = iso[0].nxx;  // every anisotropic shot covers the whole model
  const unsigned n = (unsigned)nzz * (unsigned)nxx;  // < 2^31 (checked when the batch is made)
  const int k0 = blockIdx.y * QP_SHOTS, kn = min(nshots - k0, QP_SHOTS);
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned index = blockIdx.x * blockDim.x + threadIdx.x; index < n; index += stride) {
    const int i = (int)(index % (unsigned)nzz);
    const int j = (int)(index / (unsigned)nzz);
    const bool inner = (i >= NB) && (i < nzz - NB) && (j >= NB) && (j < nxx - NB);
    float C00 = 0.0f, C01 = 0.0f, C02 = 0.0f, C11 = 0.0f, C12 = 0.0f, C22 = 0.0f, Ro = 0.0f;
    if (inner) {
      C00 = tb.v[0][index];
      C01 = tb.v[1][index];
      C02 = tb.v[2][index];
      C11 = tb.v[3][index];
      C12 = tb.v[4][index];
      C22 = tb.v[5][index];
      Ro = tb.v[6][index];
    }
    for (int q = 0; q < kn; ++q) {
      const ShotDev& si = iso[k0 + q];
      const ShotDev& so = qp[k0 + q];
      const int sIdx = (int)((si.sx + 0.5f * dx) / dx) + NB;
      const int sIdz = (int)((si.sz + 0.5f * dz) / dz) + NB;
      const gfloat* __restrict__ T = gptr(si.T);
      const float s_iso = s_at(si, nzz, i, j);
      float s_new = s_iso;
      if (inner && !((i == sIdz) && (j == sIdx))) {
        const float dTz = 0.5f * (T[(i + 1) + (size_t)j * nzz] - T[(i - 1) + (size_t)j * nzz]) / dz;
        const float dTx = 0.5f * (T[i + (size_t)(j + 1) * nzz] - T[i + (size_t)(j - 1) * nzz]) / dx;
        const float norm = sqrtf(dTx * dTx + dTz * dTz);
        const float p0 = dTx / norm;
        const float p1 = dTz / norm;
        const float Gxx = C00 * p0 * p0 + C22 * p1 * p1 + 2.0f * C02 * p0 * p1;
        const float Gzz = C22 * p0 * p0 + C11 * p1 * p1 + 2.0f * C12 * p0 * p1;
        const float Gxz = C02 * p0 * p0 + C12 * p1 * p1 + (C01 + C22) * p0 * p1;
        const float coeff1 = Gxx + Gzz;
        const float coeff2 = Gxx - Gzz;
        const float det = sqrtf((coeff2 * coeff2) / 4.0f + Gxz * Gxz);
        float g0 = (float)((double)coeff1 / 2.0 + (double)det);
        const float g1 = (float)((double)coeff1 / 2.0 - (double)det);
        if (g0 < g1) g0 = g1;
        s_new = 1.0f / sqrtf(g0 * Ro);
      }
      gptr(const_cast<float*>(so.sbase))[index] = s_new;
    }
  }
}

// Refined reciprocals of the local solver's divisors (the first three steps of
// the hardware division sequence; see rdiv).
__global__ void k_recip(const float* __restrict__ b, float* __restrict__ r, int n) {
  const int i = threadIdx.x;
  if (i < n) {
    const float r0 = __builtin_amdgcn_rcpf(b[i]);
    const float e = __builtin_fmaf(-b[i], r0, 1.0f);
    r[i] = __builtin_fmaf(e, r0, r0);
  }
}

// Exhaustive check of fdiv for the divisors b[0..n): every positive dividend
// a with 2^-100 <= a < 2^100 (biased exponents 27..226, all 2^23 fractions),
// against IEEE division.  fdiv is odd in a, so negative dividends follow.
// bad[d] is set when divisor d fails for some a.
constexpr unsigned FDIV_LO = 27u << 23, FDIV_HI = 227u << 23;
__global__ void k_fastdiv_check(const float* __restrict__ b, const float* __restrict__ zh,
                                const float* __restrict__ zl, int* __restrict__ bad) {
  const int d = blockIdx.y;  // one divisor per grid row
  const float bd = b[d], h = zh[d], l = zl[d];
  const unsigned stride = gridDim.x * blockDim.x;
  int flag = 0;
  for (unsigned u = FDIV_LO + blockIdx.x * blockDim.x + threadIdx.x; u < FDIV_HI; u += stride) {
    const float a = __uint_as_float(u);
    flag |= __float_as_uint(fdiv(a, h, l)) != __float_as_uint(a / bd);
  }
  if (flag) atomicOr(bad + d, 1);
}

// ---------------------------------------------------------------------------
// Host-side launchers.
void launch_recip(const float* b, float* r, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_recip, dim3(1), dim3(64), 0, st, b, r, n);
}

void launch_fastdiv_check(const float* b, const float* zh, const float* zl, int n, int* bad, hipStream_t st) {
  hipLaunchKernelGGL(k_fastdiv_check, dim3(4096, n), dim3(256), 0, st, b, zh, zl, bad);
}

void launch_fill(float* p, size_t n, float v, hipStream_t st) {
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, p, n, v);
}

void launch_expand(const float* s, float* spad, int nz, int nx, hipStream_t st) {
  const size_t n = (size_t)(nz + 2 * NB) * (nx + 2 * NB);
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_expand, dim3(blocks), dim3(256), 0, st, s, spad, nz, nx);
}

void launch_check_positive(const float* s, size_t n, float hmin, float hmax, int* bad, hipStream_t st) {
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_check_positive, dim3(blocks), dim3(256), 0, st, s, n, hmin, hmax, bad);
}

void launch_init(const ShotDev* shots, int nshots, int nzz, float dx, float dz, hipStream_t st) {
  const int n = nshots * 9;
  hipLaunchKernelGGL(k_init, dim3((n + 255) / 256), dim3(256), 0, st, shots, nshots, nzz, dx, dz);
}

template <int SW, bool EQH>
static void plain_level(const ShotDev& sd, int nzz, int zoff, int xoff, int nblk, const EikConsts& k,
                        hipStream_t st) {
  hipLaunchKernelGGL((k_sweep_plain<SW, EQH>), dim3(nblk), dim3(256), 0, st, sd, nzz, zoff, xoff, k);
}

// Modeling::eikonal_solver's level loop (src/modeling/modeling.cu:123-159).
int launch_sweeps_plain(const ShotDev& sd, int nzz, const EikConsts& k, bool eqh, hipStream_t st) {
  const int nxx = sd.nxx;
  const int total_levels = (nxx - 1) + (nzz - 1);
  const int min_level = std::min(nxx, nzz), max_level = std::max(nxx, nzz);
  int launches = 0;
  for (int sweep = 0; sweep < 4; ++sweep) {
    for (int level = 0; level < total_levels; ++level) {
      const int z_offset = (sweep == 0) ? ((level < nxx) ? 0 : level - nxx + 1)
                           : (sweep == 1) ? ((level < nzz) ? nzz - level - 1 : 0)
                           : (sweep == 2) ? ((level < nzz) ? level : nzz - 1)
                                          : ((level < nxx) ? nzz - 1 : nzz - 1 - (level - nxx + 1));
      const int x_offset = (sweep == 0) ? ((level < nxx) ? level : nxx - 1)
                           : (sweep == 1) ? ((level < nzz) ? 0 : level - nzz + 1)
                           : (sweep == 2) ? ((level < nzz) ? nxx - 1 : nxx - 1 - (level - nzz + 1))
                                          : ((level < nxx) ? nxx - level - 1 : 0);
      const int n_elements = (level < min_level)   ? level + 1
                             : (level >= max_level) ? total_levels - level
                                                    : total_levels - min_level - max_level + level;
      const int nblk = (n_elements + 255) / 256;
      if (nblk <= 0) continue;
      switch (sweep * 2 + (eqh ? 1 : 0)) {
        case 0: plain_level<0, false>(sd, nzz, z_offset, x_offset, nblk, k, st); break;
        case 1: plain_level<0, true>(sd, nzz, z_offset, x_offset, nblk, k, st); break;
        case 2: plain_level<1, false>(sd, nzz, z_offset, x_offset, nblk, k, st); break;
        case 3: plain_level<1, true>(sd, nzz, z_offset, x_offset, nblk, k, st); break;
        case 4: plain_level<2, false>(sd, nzz, z_offset, x_offset, nblk, k, st); break;
        case 5: plain_level<2, true>(sd, nzz, z_offset, x_offset, nblk, k, st); break;
        case 6: plain_level<3, false>(sd, nzz, z_offset, x_offset, nblk, k, st); break;
        default: plain_level<3, true>(sd, nzz, z_offset, x_offset, nblk, k, st); break;
      }
      ++launches;
    }
  }
  return launches;
}

template <int SA, int SB, bool EQH, bool FD, bool SKEW, bool FIRST, bool CONV = false>
static void persist_launch(const ShotDev* shots, const int4* tiles, int nzz, const EikConsts& k,
                           const SkewView& sv, const PersistCtl& ctl, hipStream_t st) {
  static int per_cu = 0, cus = 0;
  if (per_cu == 0) {
    int dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_persist<SA, SB, EQH, FD, SKEW, FIRST, CONV>,
                                                     64, 0) != hipSuccess || per_cu <= 0)
      per_cu = 4;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  const int grid = (ctl.wpcu > 0 ? std::min(ctl.wpcu, per_cu) : per_cu) * cus;
  hipLaunchKernelGGL((k_sweep_persist<SA, SB, EQH, FD, SKEW, FIRST, CONV>), dim3(std::min(grid, ctl.ntiles)),
                     dim3(64), 0, st, shots, tiles, nzz, k, sv, ctl);
}

template <int SA, int SB, bool EQH, bool SKEW, bool FIRST>
static void tiles_diag(const ShotDev* shots, const int4* tiles, int n, int nzz, const EikConsts& k,
                       const SkewView& sv, const PersistCtl* ctl, hipStream_t st) {
  if (ctl) {
    // the fdiv variant exists for the skewed-slowness (isotropic batch) path,
    // and so does converged mode (a change word in the control block)
    if constexpr (SKEW) {
      if (ctl->changed) {
        if (k.fastdiv)
          persist_launch<SA, SB, EQH, true, true, FIRST, true>(shots, tiles, nzz, k, sv, *ctl, st);
        else
          persist_launch<SA, SB, EQH, false, true, FIRST, true>(shots, tiles, nzz, k, sv, *ctl, st);
        return;
      }
    }
    if (SKEW && k.fastdiv)
      persist_launch<SA, SB, EQH, SKEW, SKEW, FIRST>(shots, tiles, nzz, k, sv, *ctl, st);
    else
      persist_launch<SA, SB, EQH, false, SKEW, FIRST>(shots, tiles, nzz, k, sv, *ctl, st);
  } else {
    hipLaunchKernelGGL((k_sweep_tiles<SA, SB, EQH, SKEW, FIRST>), dim3(n), dim3(64), 0, st, shots, tiles, nzz, k,
                       sv);
  }
}

template <bool EQH, bool SKEW>
static void tiles_dispatch(int sweep, bool first, const ShotDev* shots, const int4* tl, int n, int nzz,
                           const EikConsts& k, const SkewView& sv, const PersistCtl* ctl, hipStream_t st) {
  switch (sweep) {
    case 0:
      if (first)
        tiles_diag<1, 1, EQH, SKEW, true>(shots, tl, n, nzz, k, sv, ctl, st);
      else
        tiles_diag<1, 1, EQH, SKEW, false>(shots, tl, n, nzz, k, sv, ctl, st);
      break;
    case 1: tiles_diag<-1, 1, EQH, SKEW, false>(shots, tl, n, nzz, k, sv, ctl, st); break;
    case 2: tiles_diag<1, -1, EQH, SKEW, false>(shots, tl, n, nzz, k, sv, ctl, st); break;
    default: tiles_diag<-1, -1, EQH, SKEW, false>(shots, tl, n, nzz, k, sv, ctl, st); break;
  }
}

static void tiles_launch(int sweep, bool first, bool eqh, const SkewView* skew, const ShotDev* shots,
                         const int4* tl, int n, int nzz, const EikConsts& k, const PersistCtl* ctl,
                         hipStream_t st) {
  const SkewView sv = skew ? skew[sweep] : SkewView{nullptr, 0};
  if (skew) {
    if (eqh) tiles_dispatch<true, true>(sweep, first, shots, tl, n, nzz, k, sv, ctl, st);
    else tiles_dispatch<false, true>(sweep, first, shots, tl, n, nzz, k, sv, ctl, st);
  } else {
    if (eqh) tiles_dispatch<true, false>(sweep, first, shots, tl, n, nzz, k, sv, ctl, st);
    else tiles_dispatch<false, false>(sweep, first, shots, tl, n, nzz, k, sv, ctl, st);
  }
}

int launch_sweeps_tiled(const ShotDev* shots, const int4* tiles, const std::vector<int>& diag_off, int nzz,
                        const EikConsts& k, bool eqh, const SkewView* skew, bool first_fused, hipStream_t st) {
  int launches = 0;
  const int D = (int)diag_off.size() - 1;
  for (int sweep = 0; sweep < 4; ++sweep) {
    const bool first = first_fused && sweep == 0;
    for (int d = 0; d < D; ++d) {
      const int n = diag_off[d + 1] - diag_off[d];
      if (n <= 0) continue;
      tiles_launch(sweep, first, eqh, skew, shots, tiles + diag_off[d], n, nzz, k, nullptr, st);
      ++launches;
    }
  }
  return launches;
}

int launch_sweeps_persist(const ShotDev* shots, const int4* tiles, int ntiles, int nstrips, int nzz,
                          const EikConsts& k, bool eqh, const SkewView* skew, bool first_fused,
                          const PersistRun& pr, int solve, hipStream_t st, int* changed, const int* skip) {
  for (int sweep = 0; sweep < 4; ++sweep) {
    PersistCtl ctl{};
    ctl.changed = changed;
    ctl.skip = skip;
    ctl.wpcu = pr.wpcu;
    ctl.nskip = pr.counters + CTL_SKIP;
    ctl.counter = pr.counters + solve * 4 + sweep;
    ctl.epoch = pr.epoch;
    ctl.err = pr.err;
    ctl.ntiles = ntiles;
    ctl.nstrips = nstrips;
    ctl.sweep_id = (unsigned)(solve * 4 + sweep);
    tiles_launch(sweep, first_fused && sweep == 0, eqh, skew, shots, tiles, ntiles, nzz, k, &ctl, st);
  }
  return 4;
}

template <int SB, bool EQH, bool FD, bool FIRST_A, bool SKEW>
static void pair_launch(const ShotDev* shots_a, const ShotDev* shots_b, const int4* tiles, const int* qoff,
                        int nzz, const EikConsts& k, const SkewView& sv_a, const SkewView& sv_b,
                        const PersistCtl& ca, const PersistCtl& cb, hipStream_t st) {
  static int per_cu = 0, cus = 0;
  if (per_cu == 0) {
    int dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sweep_pair<SB, EQH, FD, FIRST_A, SKEW>, 64, 0) !=
            hipSuccess ||
        per_cu <= 0)
      per_cu = 4;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  const int grid = (ca.wpcu > 0 ? std::min(ca.wpcu, per_cu) : per_cu) * cus;
  hipLaunchKernelGGL((k_sweep_pair<SB, EQH, FD, FIRST_A, SKEW>), dim3(grid), dim3(64), 0, st, shots_a, shots_b, tiles,
                     qoff, nzz, k, sv_a, sv_b, ca, cb);
}

template <int SB, bool FIRST_A>
static void pair_dispatch(bool eqh, bool fd, bool skew, const ShotDev* shots_a, const ShotDev* shots_b,
                          const int4* tiles, const int* qoff, int nzz, const EikConsts& k, const SkewView& sv_a,
                          const SkewView& sv_b, const PersistCtl& ca, const PersistCtl& cb, hipStream_t st) {
#define SEIS_PAIR(E, F, S) pair_launch<SB, E, F, FIRST_A, S>(shots_a, shots_b, tiles, qoff, nzz, k, sv_a, sv_b, ca, cb, st)
  // the known-divisor division where it is exact
  if (skew) {
    if (eqh) {
      if (fd) SEIS_PAIR(true, true, true);
      else SEIS_PAIR(true, false, true);
    } else {
      if (fd) SEIS_PAIR(false, true, true);
      else SEIS_PAIR(false, false, true);
    }
  } else {
    if (eqh) {
      if (fd) SEIS_PAIR(true, true, false);
      else SEIS_PAIR(true, false, false);
    } else {
      if (fd) SEIS_PAIR(false, true, false);
      else SEIS_PAIR(false, false, false);
    }
  }
#undef SEIS_PAIR
}

int launch_sweeps_pairs(const ShotDev* shots_a, const ShotDev* shots_b, const int4* ptiles, const int* qoff,
                        int nstrips, int nzz, const EikConsts& k, bool eqh, const SkewView* skew,
                        const PersistRun& pr, int solve, hipStream_t st) {
  const SkewView none{nullptr, 0};
  for (int p = 0; p < 2; ++p) {
    PersistCtl ca{}, cb{};
    ca.wpcu = cb.wpcu = pr.wpcu;
    ca.nskip = cb.nskip = pr.counters + CTL_SKIP;
    ca.counter = pr.counters + PAIR_COUNTERS + 16 * solve + 8 * p;
    ca.epoch = cb.epoch = pr.epoch;
    ca.err = cb.err = pr.err;
    ca.ntiles = cb.ntiles = 0;
    ca.nstrips = cb.nstrips = nstrips;
    ca.sweep_id = (unsigned)(4 * solve + 2 * p);
    cb.sweep_id = (unsigned)(4 * solve + 2 * p + 1);
    ca.xtag = cb.xtag = 0u;
    cb.counter = ca.counter;
    const SkewView& sa = skew ? skew[2 * p] : none;
    const SkewView& sb = skew ? skew[2 * p + 1] : none;
    if (p == 0)
      pair_dispatch<1, true>(eqh, k.fastdiv != 0, skew != nullptr, shots_a, shots_b, ptiles, qoff, nzz, k, sa, sb,
                             ca, cb, st);
    else
      pair_dispatch<-1, false>(eqh, k.fastdiv != 0, skew != nullptr, shots_a, shots_b, ptiles, qoff, nzz, k, sa, sb,
                               ca, cb, st);
  }
  return 2;
}

void launch_xcc_ids(int* out, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_xcc_ids, dim3(n), dim3(64), 0, st, out);
}

void launch_conv_init(int* cw, int max_passes, hipStream_t st) {
  hipLaunchKernelGGL(k_conv_init, dim3(1), dim3(64), 0, st, cw, max_passes);
}

void launch_conv_gate(int* cw, hipStream_t st) { hipLaunchKernelGGL(k_conv_gate, dim3(1), dim3(64), 0, st, cw); }

void launch_begin_run(const PersistRun& pr, hipStream_t st) {
  hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(64), 0, st, pr.counters, NUM_COUNTERS, pr.epoch);
}

// Skewed slowness copy for sweep orientation (SA, SB) of the padded model
// (nzz x nxx): one thread per entry.
__global__ void k_skew_s(const float* __restrict__ spad, float* __restrict__ sk, int nzz, int nxx, int KP,
                         int NS, int SA, int SB) {
  const size_t n = (size_t)NS * KP;
  size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; idx < n; idx += stride) {
    const int kap = (int)(idx % KP);
    const int sig = (int)(idx / KP) - SK_OFF;
    const int m = (SB > 0) ? sig - (kap & 63) : sig + (kap & 63) - 63;
    const int i = min(max((SA > 0) ? kap : nzz - 2 - kap, 0), nzz - 1);
    sk[idx] = spad[i + (size_t)min(max(m, 0), nxx - 1) * nzz];
  }
}

void launch_skew_s(const float* spad, float* sk, int nzz, int nxx, int KP, int NS, int sweep, hipStream_t st) {
  const int SA = (sweep == 0 || sweep == 2) ? 1 : -1;
  const int SB = (sweep == 0 || sweep == 1) ? 1 : -1;
  const size_t n = (size_t)NS * KP;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_skew_s, dim3(blocks), dim3(256), 0, st, spad, sk, nzz, nxx, KP, NS, SA, SB);
}

void launch_picks(const ShotDev* shots, const int* rec_shot, const float* rx, const float* rz, float* out,
                  int nrec, int nzz, float dx, float dz, hipStream_t st, int* err) {
  if (nrec <= 0) return;
  hipLaunchKernelGGL(k_picks, dim3((nrec + 255) / 256), dim3(256), 0, st, shots, rec_shot, rx, rz, out, nrec,
                     nzz, dx, dz, err);
}


void launch_ani_table(const ShotDev& model, int nzz, int nxx, const StiffDev& stf, const AniTab& tb,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_ani_table, dim3((nzz + 63) / 64, nxx), dim3(64), 0, st, model, nzz, nxx, stf, tb);
}

void launch_quasi_slowness(const ShotDev* iso, ShotDev* qp, int nshots, int nzz, int nxx, float dx, float dz,
                           const AniTab& tb, hipStream_t st) {
  const size_t n = (size_t)nzz * nxx;
  const int bx = (int)std::min<size_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_quasi_slowness, dim3(bx, (nshots + QP_SHOTS - 1) / QP_SHOTS), dim3(256), 0, st, iso, qp,
                     nshots, nzz, dx, dz, tb);
}

}  // namespace seisfat
