// Adjoint-state (reverse-mode) gradient kernels of the single-pass eikonal
// solve.  The forward is SeisFAT2D's time_init + one pass of four sweeps
// (src/modeling/modeling.cu:110-159, local solver :308-357); the gradient is
// the exact derivative of that discrete map, so it is what a finite
// difference of the forward converges to.  The reference's tomography uses a
// ray-traced sensitivity instead (src/inversion/inversion.cpp:95-177), which
// this gradient replaces.
//
// Branch decisions and local derivatives come from an fp64 trajectory of the
// solve recorded on the device; the adjoint recurrence runs in fp32 and the
// cross-shot accumulation in fp64.  Every reduction runs in a fixed order, so
// the gradient is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "eik_adjoint.h"

namespace seisfat {

// ---------------------------------------------------------------------------
__global__ void k_fill_d(double* __restrict__ p, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_adj_seed(const size_t* __restrict__ node, const int* __restrict__ ptr,
                           const int* __restrict__ rec, const float* __restrict__ wt, int nseed,
                           const float* __restrict__ w, float* __restrict__ lam) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nseed) return;
  float acc = 0.0f;
  for (int e = ptr[n]; e < ptr[n + 1]; ++e) acc += w[rec[e]] * wt[e];
  lam[node[n]] = acc;
}

// ---------------------------------------------------------------------------
// The recorded trajectory is an fp64 solve of the same single-pass scheme:
// fp32 rounding decides near-tied branch choices (two sweeps' candidates
// agreeing to a few ulps, common where the wave travels along a grid axis),
// and the derivative of the fp32 map then differs locally from the derivative
// the fp64 map (and finite differences of it) has.  Decisions and partials
// are taken from the fp64 trajectory; the adjoint recurrence runs in fp32.
struct ConstsD {
  double dx, dz, dx2i, dz2i, sum, prod, hyp;
  double r_dx, r_dz, r_sum, r_hyp;  // RN(1/divisor), for cdiv
};

// RN(a / b) for a model-constant divisor b, from y = RN(1/b) (computed on the
// host by IEEE division): q = RN(a y) is within one ulp of a/b, the residual
// a - b q is exact (one fma), and RN(q + (a - b q) y) is the correctly
// rounded quotient (Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al.,
// Handbook of Floating-Point Arithmetic, the FMA-based division theorems),
// for every a whose quotient neither underflows nor overflows.  Three
// instructions instead of the general division's scale / reciprocal /
// iteration / fix-up sequence; the dividends here are traveltime differences
// and slowness products of a validated model, far from both limits.
__device__ __forceinline__ double cdiv(double a, double b, double y) {
  const double q = a * y;
  const double r = __builtin_fma(-b, q, a);
  return __builtin_fma(r, y, q);
}

__device__ __forceinline__ double dpp_shr1_d(double old, double src) {
  const long long o = __double_as_longlong(old), s = __double_as_longlong(src);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)s, 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), 0x138, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// wave_shl:1: lane k receives lane k+1's `src`; lane 63 keeps `old`.
__device__ __forceinline__ float dpp_shl1f(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), 0x130, 0xf, 0xf,
                                                    false));
}

// The local solver (src/modeling/modeling.cu:308-357) in fp64, operation for
// operation as the fp64 oracle evaluates it.
__device__ __forceinline__ double local_d(double tv, double te, double tev, double sref, double szo, double sxo,
                                          const ConstsD& k) {
  const double t1d1 = tv + k.dz * fmin(sref, szo);
  const double t1d2 = te + k.dx * fmin(sref, sxo);
  const double t1D = fmin(t1d1, t1d2);
  double t2D = 1e6;
  if ((tv <= te + k.dx * sref) && (te <= tv + k.dz * sref) && (te - tev >= 0.0) && (tv - tev >= 0.0)) {
    const double ta = tev + te - tv;
    const double tb = tev - te + tv;
    t2D = cdiv((tb * k.dz2i + ta * k.dx2i) + sqrt(4.0 * sref * sref * k.sum - k.prod * (ta - tb) * (ta - tb)), k.sum,
               k.r_sum);
  } else if ((te - tev <= cdiv(sref * k.dz * k.dz, k.hyp, k.r_hyp)) && (te - tev > 0.0)) {
    const double q = cdiv(te - tev, k.dz, k.r_dz);
    t2D = te + k.dx * sqrt(sref * sref - q * q);
  } else if ((tv - tev <= cdiv(sref * k.dx * k.dx, k.hyp, k.r_hyp)) && (tv - tev > 0.0)) {
    const double p = cdiv(tv - tev, k.dx, k.r_dx);
    t2D = tv + k.dz * sqrt(sref * sref - p * p);
  }
  return fmin(t1D, fmin(t2D, 1e6));
}

// time_set + time_init in fp64 (T0 holds 1e6 elsewhere for good).
__global__ void k_fwd64_init(const AdjShot* __restrict__ shots, int nshots, int nzz, double dx, double dz,
                             double* __restrict__ T0) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nshots * 9) return;
  const AdjShot sh = shots[t / 9];
  const int ii = (t % 9) % 3, jj = (t % 9) / 3;
  const double sx = sh.sd.sx, sz = sh.sd.sz;
  const int sIdx = (int)((sx + 0.5 * dx) / dx) + NB;
  const int sIdz = (int)((sz + 0.5 * dz) / dz) + NB;
  const int zi = sIdz + (ii - 1), xi = sIdx + (jj - 1);
  const double ddx = (double)(xi - NB) * dx - sx;
  const double ddz = (double)(zi - NB) * dz - sz;
  T0[sh.toff + zi + (size_t)xi * nzz] = (double)s_at(sh.sd, nzz, zi, xi) * sqrt(ddx * ddx + ddz * ddz);
}

// One fp64 sweep, Tnew = sweep(Told), out of place (ring nodes of Tnew hold
// 1e6 for good).  A wave walks one strip of 64 sweep-local rows of one shot;
// lane k holds row 64r + k and visits column t - k at step t: its z-upwind
// and diagonal neighbours are lane k-1's results one and two steps back (DPP
// shift), its x-upwind neighbour its own previous result, and lane 0 reads
// the previous strip's last row from Tnew.  Inputs are loaded ADJ_PF steps
// ahead.
//
// The strips of a shot run concurrently, each trailing the one above: lane
// 0's upwind inputs are the upper strip's lane-63 results (its last row), which
// that strip publishes as it computes them, column by column, each double as
// two 64-bit words (one 32-bit half + the launch's tag in the high half),
// written and read as single agent-scope atomics: a reader that sees the tag
// in both words has the value, with no fence on either side.  The wave
// fetches the words of a block of ADJ_PF steps in one load (lane u holds step
// u's), one block ahead, takes step s's from its lane, and, while a tag is
// stale, polls that column again (bounded: past the limit it sets the flag and
// stops waiting for the rest of the strip, so a broken dependency cannot stall
// the device).  Lane 0's diagonal input is
// its previous step's upwind value (the ring column 0 holds 1e6).  Strips are
// claimed from a ticket counter strip-major (every shot's strip r before any
// strip r+1), so a strip's upper neighbour always holds a smaller ticket and
// is running: the walk drains for any number of resident waves.
constexpr int WALK_SPIN = 1 << 16;

__global__ void k_walk_begin(int* __restrict__ ctl) {
  if (threadIdx.x == 0) ctl[0] = 0;
}

__device__ __forceinline__ unsigned long long walk_word(unsigned half, unsigned tag) {
  return ((unsigned long long)tag << 32) | half;
}

template <int SA, int SB>
__global__ __launch_bounds__(64) void k_fwd64_sweep(const AdjShot* __restrict__ shots, int nshots, int nzz,
                                                    int nstr, ConstsD K, const double* __restrict__ Told,
                                                    double* __restrict__ Tnew, int* __restrict__ ctl,
                                                    unsigned long long* __restrict__ xw, int ldx, unsigned tag) {
  typedef __attribute__((address_space(1))) unsigned long long gword;
  const int lane = threadIdx.x;
  for (;;) {
    int tk = 0;
    if (lane == 0) tk = atomicAdd(ctl, 1);
    tk = __builtin_amdgcn_readfirstlane(tk);
    if (tk >= nshots * nstr) break;
    const int r = tk / nshots, ks = tk - r * nshots;
    const AdjShot sh = shots[ks];
    const int nxx = sh.sd.nxx;
    const double* __restrict__ To = Told + sh.toff;
    double* __restrict__ Tn = Tnew + sh.toff;
    const int NT = (nxx + 63 + 2 * ADJ_PF - 1) / (2 * ADJ_PF) * (2 * ADJ_PF);  // whole block pairs
    // this strip's published last row, and the upper strip's
    gword* xme = (gword*)(xw + ((size_t)ks * nstr + r) * ldx);
    const gword* xup = (const gword*)(xw + ((size_t)ks * nstr + max(r - 1, 0)) * ldx);
    const bool pub = lane == 63 && r + 1 < nstr;
    // the upper strip's row (lane 0's upwind values): the loads, the tag test
    // and the wait are wave-uniform
    const bool up = r > 0;
    const int fi = 64 * r + lane;
    const bool row_int = fi >= 1 && fi <= nzz - 2;
    const int fic = min(max(fi, 1), nzz - 2);
    const int i = SA > 0 ? fic : nzz - 1 - fic;
    double qto[ADJ_PF];
    float qsz[ADJ_PF];
    // Inputs of step t: the old value at clamped (always valid) coordinates
    // and the z-other cell (i1, j1 + SB), one load per step.  The reference
    // cell (i1, j1) is the previous step's z-other cell and the x-other cell
    // (i1 + SA, j1) is lane k+1's z-other cell of the same step (one DPP
    // move; lane 63's, in the next strip, arrives with the word blocks), so
    // the z-other cell's row and column are clamped to the cell range only,
    // keeping those identities for every interior node.
    const int fis = min(max(fi, 1), nzz - 1);
    const int is1 = (SA > 0 ? fis : nzz - 1 - fis) - (SA + 1) / 2;
    auto load = [&](int t, double& to, float& sz) {
      const int fjc = min(max(t - lane, 1), nxx - 2);
      const int j = SB > 0 ? fjc : nxx - 1 - fjc;
      to = To[i + (size_t)j * nzz];
      const int fjs = min(max(t - lane, 0), nxx - 2);
      sz = s_at(sh.sd, nzz, is1, (SB > 0 ? fjs : nxx - 1 - fjs) - (SB + 1) / 2 + SB);
    };
    // the upper strip's words of steps tb..tb+ADJ_PF-1, one block load: lane
    // u (mod ADJ_PF) holds step tb+u's (column clamp(tb+u, 1, nxx-2)).
    // Unconditional (strip 0 reads its own words and ignores them), so no
    // branch forces a wait for it.
    // With each block, lane 63's x-other cells: the z-other cells of a
    // virtual lane 64 (row 64r+64, column t-64).
    const int f64 = min(64 * r + 64, nzz - 1);
    const int i64 = (SA > 0 ? f64 : nzz - 1 - f64) - (SA + 1) / 2;
    auto xload = [&](int tb, unsigned long long& h, unsigned long long& l, float& x63) {
      const int tt = tb + (lane & (ADJ_PF - 1));
      const int fjx = min(max(tt, 1), nxx - 2);
      h = __hip_atomic_load(xup + 2 * fjx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      l = __hip_atomic_load(xup + 2 * fjx + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int fjs = min(max(tt - 64, 0), nxx - 2);
      x63 = s_at(sh.sd, nzz, i64, (SB > 0 ? fjs : nxx - 1 - fjs) - (SB + 1) / 2 + SB);
    };
#pragma unroll
    for (int u = 0; u < ADJ_PF; ++u) load(u, qto[u], qsz[u]);
    double prev = 1e6, prev2 = 1e6, gv_prev = 1e6;
    float sref = 0.0f;  // the previous step's z-other cell (unused at step 0)
    bool gave_up = false;
    // ADJ_PF steps from t0 with the upper strip's words in (xh, xl).  Each
    // step reloads its queue slot after its last use of the slot (and the
    // caller each word block after its block), so a slot keeps its register
    // across the loop and no copy waits for a load in flight.
    auto block = [&](int t0, unsigned long long xh, unsigned long long xl, float x63) {
#pragma unroll
      for (int u = 0; u < ADJ_PF; ++u) {
        const int t = t0 + u;
        double gv = 0.0, ge = 0.0;
        if (up) {
          // upper strip's value of column clamp(t, 1, nxx-2), published at its step t+63
          const int fjc = min(max(t, 1), nxx - 2);
          unsigned hw = __builtin_amdgcn_readlane((unsigned)xh, u);
          unsigned ht = __builtin_amdgcn_readlane((unsigned)(xh >> 32), u);
          unsigned lw = __builtin_amdgcn_readlane((unsigned)xl, u);
          unsigned lt = __builtin_amdgcn_readlane((unsigned)(xl >> 32), u);
          for (int it = 0; !gave_up && (ht != tag || lt != tag); ++it) {
            if (it > WALK_SPIN) {
              if (lane == 0) atomicOr(ctl + 1, 1);
              gave_up = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            const unsigned long long hi =
                __hip_atomic_load(xup + 2 * fjc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long lo =
                __hip_atomic_load(xup + 2 * fjc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            hw = __builtin_amdgcn_readfirstlane((unsigned)hi);
            ht = __builtin_amdgcn_readfirstlane((unsigned)(hi >> 32));
            lw = __builtin_amdgcn_readfirstlane((unsigned)lo);
            lt = __builtin_amdgcn_readfirstlane((unsigned)(lo >> 32));
          }
          gv = __hiloint2double((int)hw, (int)lw);
          ge = t >= 2 ? gv_prev : 1e6;  // column fjc-1: the previous step's, or the ring
          gv_prev = gv;
        }
        const double tv = dpp_shr1_d(gv, prev);
        const double tev = dpp_shr1_d(ge, prev2);
        const int fj = t - lane;
        double out = 1e6;
        const float szo = qsz[u];
        const float sxo = dpp_shl1f(__uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(x63), u)), szo);
        const float sr = sref;
        sref = szo;
        if (row_int && fj >= 1 && fj <= nxx - 2) {
          out = fmin(qto[u], local_d(tv, prev, tev, sr, szo, sxo, K));
          const int j = SB > 0 ? fj : nxx - 1 - fj;
          Tn[i + (size_t)j * nzz] = out;
        }
        if (pub && fj >= 1 && fj <= nxx - 2) {
          __hip_atomic_store(xme + 2 * fj, walk_word((unsigned)__double2hiint(out), tag), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(xme + 2 * fj + 1, walk_word((unsigned)__double2loint(out), tag), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        prev2 = prev;
        prev = out;
        load(min(t + ADJ_PF, NT - 1), qto[u], qsz[u]);
      }
    };
    // word blocks ping-pong between (ah, al) and (bh, bl), each loaded one
    // block before its use
    unsigned long long ah, al, bh, bl;
    float as, bs;
    xload(0, ah, al, as);
    // Drain once per strip: inside the loop every word block and queue slot
    // was loaded a full block (over 63 memory operations) before its use,
    // which needs no wait; the compiler's count merges the loop-entry path,
    // where the first loads are recent, and would otherwise wait there at
    // every block.
    __builtin_amdgcn_s_waitcnt(0);
    for (int t0 = 0; t0 < NT; t0 += 2 * ADJ_PF) {
      xload(t0 + ADJ_PF, bh, bl, bs);
      block(t0, ah, al, as);
      xload(min(t0 + 2 * ADJ_PF, NT - ADJ_PF), ah, al, as);
      block(t0 + ADJ_PF, bh, bl, bs);
    }
  }
}

// Catmull-Rom picks of the fp64 trajectory's final traveltimes (the residuals
// of the misfit gradient), as the fp64 oracle evaluates them.
__device__ __forceinline__ double cubic1d_d(const double P[4], double d) {
  return P[1] + 0.5 * d * (P[2] - P[0] + d * (2.0 * P[0] - 5.0 * P[1] + 4.0 * P[2] - P[3] +
                                             d * (3.0 * (P[1] - P[2]) + P[3] - P[0])));
}

__global__ void k_picks64(const AdjShot* __restrict__ shots, const int* __restrict__ rec_shot,
                          const float* __restrict__ rx, const float* __restrict__ rz, int nrec, int nzz, double dx,
                          double dz, const double* __restrict__ T4, double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const double* __restrict__ T = T4 + shots[rec_shot[r]].toff;
  const double x = rx[r], z = rz[r];
  const double x0 = floor((x + 0.5 * dx) / dx) * dx;
  const double z0 = floor((z + 0.5 * dz) / dz) * dz;
  const double x1 = floor((x + 0.5 * dx) / dx) * dx + dx;
  const double z1 = floor((z + 0.5 * dz) / dz) * dz + dz;
  const double xd = (x - x0) / (x1 - x0);
  const double zd = (z - z0) / (z1 - z0);
  const int i = (int)((z + 0.5 * dz) / dz) + NB;
  const int j = (int)((x + 0.5 * dx) / dx) + NB;
  double p[4];
  for (int pIdx = 0; pIdx < 4; ++pIdx) {
    double P[4];
    for (int pIdz = 0; pIdz < 4; ++pIdz) P[pIdz] = T[(i + pIdz - 1) + (size_t)(j + pIdx - 1) * nzz];
    p[pIdx] = cubic1d_d(P, zd);
  }
  out[r] = cubic1d_d(p, xd);
}

// ---------------------------------------------------------------------------
// Local partials of one node.  ds < 0 marks a node the sweep left unchanged
// (T_old kept: the adjoint passes through); bit (di | dj << 1) of code marks
// slowness cell (i - di, j - dj) as receiving ds, split evenly when two bits
// are set.
struct AdjCoef {
  float dtv, dte, dtev, ds;
  int code;
};

// The branch the fp64 forward took at a node (same expressions and
// comparisons as local_d) and its partial derivatives.  Ties follow the
// oracle: T_old wins when T_old <= candidate; 1-D before 2-D when equal;
// z-edge before x-edge when equal; min(S, S') of two equal cells splits dS
// half and half (what central differences measure).
template <int SA, int SB>
__device__ __forceinline__ AdjCoef adj_local(double tv, double te, double tev, double sref, double szo,
                                             double sxo, double told, const ConstsD& k) {
  constexpr int di_ref = (SA + 1) / 2, dj_ref = (SB + 1) / 2;
  constexpr int c_ref = di_ref | (dj_ref << 1);
  constexpr int c_zo = di_ref | (((1 - SB) / 2) << 1);
  constexpr int c_xo = ((1 - SA) / 2) | (dj_ref << 1);
  const double t1d1 = tv + k.dz * fmin(sref, szo);
  const double t1d2 = te + k.dx * fmin(sref, sxo);
  const double t1D = fmin(t1d1, t1d2);
  double tbr = 1e6;
  bool taken = false;
  double p_tv = 0.0, p_te = 0.0, p_tev = 0.0, p_s = 0.0;
  if ((tv <= te + k.dx * sref) && (te <= tv + k.dz * sref) && (te - tev >= 0.0) && (tv - tev >= 0.0)) {
    const double ta = tev + te - tv;
    const double tb = tev - te + tv;
    const double sq = sqrt(4.0 * sref * sref * k.sum - k.prod * (ta - tb) * (ta - tb));
    tbr = cdiv((tb * k.dz2i + ta * k.dx2i) + sq, k.sum, k.r_sum);
    taken = true;
    // t = ((tb dz2i + ta dx2i) + sqrt(4 S^2 D - dz2i dx2i (ta - tb)^2)) / D
    const double kk = 2.0 * k.prod * (ta - tb) / sq;
    p_te = cdiv(k.dx2i - k.dz2i - kk, k.sum, k.r_sum);
    p_tv = cdiv(k.dz2i - k.dx2i + kk, k.sum, k.r_sum);
    p_tev = 1.0;
    p_s = 4.0 * sref / sq;
  } else if ((te - tev <= cdiv(sref * k.dz * k.dz, k.hyp, k.r_hyp)) && (te - tev > 0.0)) {
    // t = te + dx sqrt(S^2 - q^2), q = (te - tev) / dz
    const double q = cdiv(te - tev, k.dz, k.r_dz);
    const double Q = sqrt(sref * sref - q * q);
    tbr = te + k.dx * Q;
    taken = true;
    const double r = k.dx * q / (k.dz * Q);
    p_te = 1.0 - r;
    p_tev = r;
    p_s = k.dx * sref / Q;
  } else if ((tv - tev <= cdiv(sref * k.dx * k.dx, k.hyp, k.r_hyp)) && (tv - tev > 0.0)) {
    const double p = cdiv(tv - tev, k.dx, k.r_dx);
    const double P = sqrt(sref * sref - p * p);
    tbr = tv + k.dz * P;
    taken = true;
    const double r = k.dz * p / (k.dx * P);
    p_tv = 1.0 - r;
    p_tev = r;
    p_s = k.dz * sref / P;
  }
  const double t2D = fmin(tbr, 1e6);
  const double cand = fmin(t1D, t2D);
  AdjCoef c{0.0f, 0.0f, 0.0f, -1.0f, 0};
  if (!(cand < told)) return c;
  if (t1D <= t2D) {
    if (t1d1 <= t1d2) {
      c.dtv = 1.0f;
      c.ds = (float)k.dz;
      c.code = sref == szo ? (1 << c_ref) | (1 << c_zo) : sref < szo ? 1 << c_ref : 1 << c_zo;
    } else {
      c.dte = 1.0f;
      c.ds = (float)k.dx;
      c.code = sref == sxo ? (1 << c_ref) | (1 << c_xo) : sref < sxo ? 1 << c_ref : 1 << c_xo;
    }
    return c;
  }
  if (!taken || !(tbr < 1e6)) {  // the 1e6 constant won: no dependence
    c.ds = 0.0f;
    return c;
  }
  c.dtv = (float)p_tv;
  c.dte = (float)p_te;
  c.dtev = (float)p_tev;
  c.ds = (float)p_s;
  c.code = 1 << c_ref;
  return c;
}

// One workgroup per (strip rs, block of ADJ_CB reverse-local columns, shot):
// first each wave takes whole columns of the block in natural order, lane =
// row, so its loads of T and S are contiguous runs (a wave reading one step
// of the strip-skewed layout reads a diagonal instead: 64 cache lines per
// load, which bounded this kernel); the partials go to LDS by column, and
// then the workgroup writes the skewed entries (lane k, step t) whose node
// (64rs + k, t - k) lies in the block, each step a contiguous run of lanes.
// Entries outside the grid are left untouched.
constexpr int ADJ_CB = 32;

template <int SA, int SB>
__global__ __launch_bounds__(256) void k_adj_coef(const AdjShot* __restrict__ shots, int nzz, int nstr, ConstsD k,
                                                  const double* __restrict__ Tnew, const double* __restrict__ Told,
                                                  int pa, int pb, const float* __restrict__ lam_src,
                                                  float* __restrict__ lam_dst, float4* __restrict__ coef,
                                                  uint8_t* __restrict__ code) {
  __shared__ float P[4][ADJ_CB][64];
  __shared__ uint8_t Pc[ADJ_CB][64];
  const AdjShot sh = shots[blockIdx.y];
  const int nxx = sh.sd.nxx, NTP = sh.NTP;
  const int rs = (int)(blockIdx.x % (unsigned)nstr), jp0 = (int)(blockIdx.x / (unsigned)nstr) * ADJ_CB;
  if (jp0 >= nxx) return;  // whole workgroup: blocks past this shot's width
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ip = 64 * rs + lane;
  const double* __restrict__ Tn = Tnew + sh.toff;
  for (int col = w; col < ADJ_CB; col += 4) {
    const int jp = jp0 + col;
    AdjCoef c{0.0f, 0.0f, 0.0f, -1.0f, 0};
    if (ip < nzz && jp < nxx) {
      const int i = adj_ip(ip, SA, nzz), j = adj_jp(jp, SB, nxx);  // (both maps are involutions)
      if (i >= 1 && i <= nzz - 2 && j >= 1 && j <= nxx - 2) {
        const double tv = Tn[(i - SA) + (size_t)j * nzz];
        const double te = Tn[i + (size_t)(j - SB) * nzz];
        const double tev = Tn[(i - SA) + (size_t)(j - SB) * nzz];
        const double told = Told[sh.toff + i + (size_t)j * nzz];
        const int i1 = i - (SA + 1) / 2, j1 = j - (SB + 1) / 2;
        const double sref = s_at(sh.sd, nzz, i1, j1);
        const double szo = s_at(sh.sd, nzz, i1, j1 + SB);
        const double sxo = s_at(sh.sd, nzz, i1 + SA, j1);
        c = adj_local<SA, SB>(tv, te, tev, sref, szo, sxo, told, k);
      }
    }
    P[0][col][lane] = c.dtv;
    P[1][col][lane] = c.dte;
    P[2][col][lane] = c.dtev;
    P[3][col][lane] = c.ds;
    Pc[col][lane] = (uint8_t)c.code;
  }
  __syncthreads();
  // step t = jp0 + e/64 of the skewed strip, lane e%64 (= this thread's lane)
  const size_t sbase = sh.koff + (size_t)rs * NTP * 64;
  for (int e = threadIdx.x; e < (ADJ_CB + 63) * 64; e += 256) {
    const int t = jp0 + (e >> 6), jp = t - lane, col = jp - jp0;
    if (col < 0 || col >= ADJ_CB || jp >= nxx || ip >= nzz) continue;
    const size_t ko = sbase + (size_t)t * 64 + lane;
    if (lam_src) {
      const int i = adj_ip(ip, SA, nzz), j = adj_jp(jp, SB, nxx);
      lam_dst[ko] = lam_src[sh.koff + adj_kidx(adj_ip(i, pa, nzz), adj_jp(j, pb, nxx), NTP)];
    }
    coef[ko] = make_float4(P[0][col][lane], P[1][col][lane], P[2][col][lane], P[3][col][lane]);
    code[ko] = Pc[col][lane];
  }
}

// ---------------------------------------------------------------------------
// Reverse sweep: one wave per strip of 64 reverse-local rows; lane k holds
// row 64r + k and visits column t - k at step t.  The z-upwind and diagonal
// messages come from lane k-1 (DPP shift), the x message from the lane's own
// previous step.  Lane 0's messages are the upper strip's lane-63 ones, which
// that strip publishes column by column as tagged words (word 2jp = the launch's
// tag and mv of column jp, word 2jp+1 the same for mev), written at its step
// jp+63 and read with the forward walker's protocol: one block load per
// ADJ_PF steps (lane u holds step u's words), one block ahead, and a bounded
// poll of the column while a tag is stale.  Strip 0 and the columns no strip
// publishes read zero messages.  Strips are claimed strip-major from the
// ticket counter ctl[0], so every upper neighbour is already running.  Inputs
// of ADJ_PF steps ahead are kept in flight; nodes outside the grid hold zero
// inputs and produce zero outputs.
__global__ __launch_bounds__(64) void k_adj_sweep(const AdjShot* __restrict__ shots, int nshots, int nstr,
                                                  const float* __restrict__ lam_in,
                                                  const float4* __restrict__ coef, float* __restrict__ lam_out,
                                                  float* __restrict__ gc, int* __restrict__ ctl,
                                                  unsigned long long* __restrict__ xw, int ldx, unsigned tag) {
  typedef __attribute__((address_space(1))) unsigned long long gword;
  const int lane = threadIdx.x;
  for (;;) {
    int tk = 0;
    if (lane == 0) tk = atomicAdd(ctl, 1);
    tk = __builtin_amdgcn_readfirstlane(tk);
    if (tk >= nshots * nstr) break;
    const int r = tk / nshots, ks = tk - r * nshots;
    const AdjShot sh = shots[ks];
    const int NTP = sh.NTP, nxx = sh.sd.nxx;
    const size_t base = sh.koff + (size_t)r * NTP * 64 + lane;
    gword* xme = (gword*)(xw + ((size_t)ks * nstr + r) * ldx);
    const gword* xup = (const gword*)(xw + ((size_t)ks * nstr + max(r - 1, 0)) * ldx);
    const bool up = r > 0;
    const bool pub = lane == 63 && r + 1 < nstr;
    float ql[ADJ_PF];
    float4 qc[ADJ_PF];
#pragma unroll
    for (int u = 0; u < ADJ_PF; ++u) {
      ql[u] = lam_in[base + (size_t)u * 64];
      qc[u] = coef[base + (size_t)u * 64];
    }
    // step t's words: mv of column t (word 2t) and mev of column t-1 (word
    // 2t-1), at clamped addresses (the step ignores the columns out of range)
    auto wv = [&](int t) { return xup + 2 * min(t, nxx - 1); };
    auto we = [&](int t) { return xup + max(2 * min(t, nxx) - 1, 0); };
    // unconditional (strip 0 reads its own words and ignores them)
    auto xload = [&](int tb, unsigned long long& v, unsigned long long& e) {
      const int t = tb + (lane & (ADJ_PF - 1));
      v = __hip_atomic_load(wv(t), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      e = __hip_atomic_load(we(t), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    float mv = 0.0f, me = 0.0f, mev = 0.0f, mevp = 0.0f;
    bool gave_up = false;
    auto block = [&](int t0, unsigned long long bv, unsigned long long be) {
#pragma unroll
      for (int u = 0; u < ADJ_PF; ++u) {
        const int t = t0 + u;
        float xv = 0.0f, xe = 0.0f;
        if (up) {
          const bool nv = t < nxx, ne = t >= 1 && t <= nxx;
          unsigned vw = __builtin_amdgcn_readlane((unsigned)bv, u);
          unsigned vt = __builtin_amdgcn_readlane((unsigned)(bv >> 32), u);
          unsigned ew = __builtin_amdgcn_readlane((unsigned)be, u);
          unsigned et = __builtin_amdgcn_readlane((unsigned)(be >> 32), u);
          for (int it = 0; !gave_up && ((nv && vt != tag) || (ne && et != tag)); ++it) {
            if (it > WALK_SPIN) {
              if (lane == 0) atomicOr(ctl + 1, 1);
              gave_up = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            const unsigned long long v = __hip_atomic_load(wv(t), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long e = __hip_atomic_load(we(t), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            vw = __builtin_amdgcn_readfirstlane((unsigned)v);
            vt = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
            ew = __builtin_amdgcn_readfirstlane((unsigned)e);
            et = __builtin_amdgcn_readfirstlane((unsigned)(e >> 32));
          }
          xv = nv ? __uint_as_float(vw) : 0.0f;
          xe = ne ? __uint_as_float(ew) : 0.0f;
        }
        // messages into (ip, jp): from (ip-1, jp) [lane k-1, step t-1],
        // (ip, jp-1) [own lane, step t-1], (ip-1, jp-1) [lane k-1, step t-2]
        const float in_v = dpp_shr1(xv, mv);
        const float in_ev = dpp_shr1(xe, mevp);
        const float mu = ((ql[u] + in_v) + me) + in_ev;
        mevp = mev;
        const float4 cf = qc[u];
        const bool upd = cf.w >= 0.0f;
        mv = upd ? mu * cf.x : 0.0f;
        me = upd ? mu * cf.y : 0.0f;
        mev = upd ? mu * cf.z : 0.0f;
        lam_out[base + (size_t)t * 64] = upd ? 0.0f : mu;
        gc[base + (size_t)t * 64] = upd ? mu * cf.w : 0.0f;
        const int jp = t - 63;
        if (pub && jp >= 0 && jp < nxx) {
          __hip_atomic_store(xme + 2 * jp, walk_word(__float_as_uint(mv), tag), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(xme + 2 * jp + 1, walk_word(__float_as_uint(mev), tag), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        // reload the slot after its last use (it keeps its register)
        const int tn = min(t + ADJ_PF, NTP - 1);
        ql[u] = lam_in[base + (size_t)tn * 64];
        qc[u] = coef[base + (size_t)tn * 64];
      }
    };
    unsigned long long av, ae, bv, be;
    xload(0, av, ae);
    __builtin_amdgcn_s_waitcnt(0);  // once per strip (see k_fwd64_sweep)
    for (int t0 = 0; t0 < NTP; t0 += 2 * ADJ_PF) {
      xload(t0 + ADJ_PF, bv, be);
      block(t0, av, ae);
      if (t0 + ADJ_PF >= NTP) break;  // NTP is a multiple of ADJ_PF only
      xload(t0 + 2 * ADJ_PF, av, ae);
      block(t0 + ADJ_PF, bv, be);
    }
  }
}

// ---------------------------------------------------------------------------
// Slowness contributions of sweep s onto padded cell (i, j): the nodes
// (i + di, j + dj) whose code marks offset (di, dj), in a fixed order.  One
// workgroup per (strip rs, block of ADJ_CB reverse-local columns, shot) owns
// the cells whose nodes lie in the block: it first copies the block's
// entries, with one row and one column of halo on each side, from the
// strip-skewed layout into LDS by column (each step a contiguous run of
// lanes), then sums each cell's contributions from LDS, a wave per column
// with lane = row, so the cell updates are contiguous.
template <int SA, int SB>
__global__ __launch_bounds__(256) void k_adj_fold(const AdjShot* __restrict__ shots, int nzz, int nstr,
                                                  const float* __restrict__ gc, const uint8_t* __restrict__ code,
                                                  float* __restrict__ gshot) {
  constexpr int W = ADJ_CB + 2, H = 66;  // columns jp0-1 .. jp0+ADJ_CB, rows 64rs-1 .. 64rs+64
  __shared__ float G[W][H];
  __shared__ uint8_t C[W][H];
  const AdjShot sh = shots[blockIdx.y];
  const int nxx = sh.sd.nxx, NTP = sh.NTP;
  const int rs = (int)(blockIdx.x % (unsigned)nstr), jp0 = (int)(blockIdx.x / (unsigned)nstr) * ADJ_CB;
  if (jp0 >= nxx) return;  // whole workgroup: blocks past this shot's width
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ip = 64 * rs + lane;
  // the strip's own rows: step t = jp0 - 1 + e/64, lane e%64
  const size_t sbase = sh.koff + (size_t)rs * NTP * 64;
  for (int e = threadIdx.x; e < (W + 63) * 64; e += 256) {
    const int t = jp0 - 1 + (e >> 6), jp = t - lane;
    if (jp < max(jp0 - 1, 0) || jp > jp0 + ADJ_CB || jp >= nxx || ip >= nzz) continue;
    const size_t ko = sbase + (size_t)t * 64 + lane;
    G[jp - jp0 + 1][lane + 1] = gc[ko];
    C[jp - jp0 + 1][lane + 1] = code[ko];
  }
  // halo rows: the last row of strip rs-1 and the first of strip rs+1
  if (threadIdx.x < 2 * W) {
    const int h = threadIdx.x / W, jp = jp0 - 1 + threadIdx.x % W;
    const int iph = h == 0 ? 64 * rs - 1 : 64 * rs + 64;
    if (iph >= 0 && iph < nzz && jp >= 0 && jp < nxx) {
      const size_t ko = sh.koff + adj_kidx(iph, jp, NTP);
      G[jp - jp0 + 1][h == 0 ? 0 : H - 1] = gc[ko];
      C[jp - jp0 + 1][h == 0 ? 0 : H - 1] = code[ko];
    }
  }
  __syncthreads();
  for (int col = w; col < ADJ_CB; col += 4) {
    const int jp = jp0 + col;
    if (ip >= nzz || jp >= nxx) continue;
    const int i = adj_ip(ip, SA, nzz), j = adj_jp(jp, SB, nxx);
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int di = c & 1, dj = c >> 1;
      const int ni = i + di, nj = j + dj;
      if (ni >= 1 && ni <= nzz - 2 && nj >= 1 && nj <= nxx - 2) {
        // node (ni, nj) in reverse-local coordinates, within the LDS window
        const int r = adj_ip(ni, SA, nzz) - 64 * rs + 1, q = adj_jp(nj, SB, nxx) - jp0 + 1;
        const int m = C[q][r];
        if (m & (1 << c)) acc += (m & (m - 1)) ? 0.5f * G[q][r] : G[q][r];
      }
    }
    gshot[sh.toff + i + (size_t)j * nzz] += acc;
  }
}

// T0 = S(node) * distance on the source stencil: dT0/dS(node) = distance.
__global__ void k_adj_source(const AdjShot* __restrict__ shots, int nshots, int nzz, double dx, double dz,
                             const float* __restrict__ lam0, float* __restrict__ gshot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nshots * 9) return;
  const AdjShot sh = shots[t / 9];
  const int ii = (t % 9) % 3, jj = (t % 9) / 3;
  const int sIdx = (int)(((double)sh.sd.sx + 0.5 * dx) / dx) + NB;
  const int sIdz = (int)(((double)sh.sd.sz + 0.5 * dz) / dz) + NB;
  const int zi = sIdz + (ii - 1), xi = sIdx + (jj - 1);
  const double ddx = (double)(xi - NB) * dx - (double)sh.sd.sx;
  const double ddz = (double)(zi - NB) * dz - (double)sh.sd.sz;
  const float lam = lam0[sh.koff + adj_kidx(adj_ip(zi, 1, nzz), adj_jp(xi, 1, sh.sd.nxx), sh.NTP)];
  gshot[sh.toff + zi + (size_t)xi * nzz] += lam * (float)sqrt(ddx * ddx + ddz * ddz);
}

// d/ds of the unpadded model: every padded node of every shot maps to the
// model cell it replicates (rows clamped to the model, columns clamped to the
// shot's window); sums in fp64, shots in order.  fixed_border: the model's
// border was uploaded as given (not the replication of its edges), so its
// rows, and the border columns of shots on the whole model, are constants
// with no derivative; a window's own replicated columns still fold.
__global__ __launch_bounds__(256) void k_adj_reduce(const AdjShot* __restrict__ shots, int nshots, int nz,
                                                    int nx, int nzz, const float* __restrict__ gshot,
                                                    double* __restrict__ grad, int fixed_border) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nz * nx) return;
  const int iz = idx % nz, ix = idx / nz;
  const int plo = (iz == 0 && !fixed_border) ? 0 : iz + NB;
  const int phi = (iz == nz - 1 && !fixed_border) ? nz - 1 + 2 * NB : iz + NB;
  double acc = 0.0;
  for (int s = 0; s < nshots; ++s) {
    const int x0 = shots[s].sd.x0, w = shots[s].sd.w;
    if (ix < x0 || ix >= x0 + w) continue;
    const int lq = ix - x0;
    const bool cfix = fixed_border && shots[s].sd.sofs == 0;  // a whole-model shot reads the border as given
    const int qlo = (lq == 0 && !cfix) ? 0 : lq + NB;
    const int qhi = (lq == w - 1 && !cfix) ? w - 1 + 2 * NB : lq + NB;
    const float* __restrict__ g = gshot + shots[s].toff;
    for (int q = qlo; q <= qhi; ++q)
      for (int p = plo; p <= phi; ++p) acc += (double)g[p + (size_t)q * nzz];
  }
  grad[idx] = acc;
}

// ---------------------------------------------------------------------------
void launch_fill_d(double* p, size_t n, double v, hipStream_t st) {
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_fill_d, dim3(blocks), dim3(256), 0, st, p, n, v);
}

void launch_adj_seed(const size_t* node, const int* ptr, const int* rec, const float* wt, int nseed,
                     const float* w, float* lam, hipStream_t st) {
  if (nseed <= 0) return;
  hipLaunchKernelGGL(k_adj_seed, dim3((nseed + 255) / 256), dim3(256), 0, st, node, ptr, rec, wt, nseed, w, lam);
}

static ConstsD consts_d(double dx, double dz) {
  ConstsD k;
  k.dx = dx;
  k.dz = dz;
  k.dz2i = 1.0 / (dz * dz);
  k.dx2i = 1.0 / (dx * dx);
  k.sum = k.dz2i + k.dx2i;
  k.prod = k.dz2i * k.dx2i;
  k.hyp = std::sqrt(dx * dx + dz * dz);
  k.r_dx = 1.0 / k.dx;
  k.r_dz = 1.0 / k.dz;
  k.r_sum = 1.0 / k.sum;
  k.r_hyp = 1.0 / k.hyp;
  return k;
}

void launch_fwd64_init(const AdjShot* shots, int nshots, int nzz, double dx, double dz, double* T0, hipStream_t st) {
  const int n = nshots * 9;
  hipLaunchKernelGGL(k_fwd64_init, dim3((n + 255) / 256), dim3(256), 0, st, shots, nshots, nzz, dx, dz, T0);
}

void launch_fwd64_sweep(int sweep, const AdjShot* shots, int nshots, int nzz, int nstr, double dx, double dz,
                        const double* Told, double* Tnew, int* walk_ctl, unsigned long long* walk_x, int ldx,
                        unsigned tag, hipStream_t st) {
  const ConstsD k = consts_d(dx, dz);
  hipLaunchKernelGGL(k_walk_begin, dim3(1), dim3(64), 0, st, walk_ctl);
  // one wave per strip, at most what is resident at once (the tickets keep
  // any count correct)
  const int grid = std::min(nshots * nstr, 4096);
  switch (sweep) {
    case 0: hipLaunchKernelGGL((k_fwd64_sweep<1, 1>), dim3(grid), dim3(64), 0, st, shots, nshots, nzz, nstr, k, Told, Tnew, walk_ctl, walk_x, ldx, tag); break;
    case 1: hipLaunchKernelGGL((k_fwd64_sweep<-1, 1>), dim3(grid), dim3(64), 0, st, shots, nshots, nzz, nstr, k, Told, Tnew, walk_ctl, walk_x, ldx, tag); break;
    case 2: hipLaunchKernelGGL((k_fwd64_sweep<1, -1>), dim3(grid), dim3(64), 0, st, shots, nshots, nzz, nstr, k, Told, Tnew, walk_ctl, walk_x, ldx, tag); break;
    default: hipLaunchKernelGGL((k_fwd64_sweep<-1, -1>), dim3(grid), dim3(64), 0, st, shots, nshots, nzz, nstr, k, Told, Tnew, walk_ctl, walk_x, ldx, tag); break;
  }
}

void launch_picks64(const AdjShot* shots, const int* rec_shot, const float* rx, const float* rz, int nrec, int nzz,
                    double dx, double dz, const double* T4, double* out, hipStream_t st) {
  if (nrec <= 0) return;
  hipLaunchKernelGGL(k_picks64, dim3((nrec + 255) / 256), dim3(256), 0, st, shots, rec_shot, rx, rz, nrec, nzz, dx,
                     dz, T4, out);
}

void launch_adj_coef(int sweep, const AdjShot* shots, int nshots, int nstr, int max_nxx, int nzz, double dx,
                     double dz, const double* Tnew, const double* Told, const float* lam_src, float* lam_dst,
                     float4* coef, uint8_t* code, hipStream_t st) {
  const dim3 grid(nstr * ((max_nxx + ADJ_CB - 1) / ADJ_CB), nshots);
  const int pa = adj_a(sweep + 1), pb = adj_b(sweep + 1);
  const ConstsD k = consts_d(dx, dz);
  switch (sweep) {
    case 0: hipLaunchKernelGGL((k_adj_coef<1, 1>), grid, dim3(256), 0, st, shots, nzz, nstr, k, Tnew, Told, pa, pb, lam_src, lam_dst, coef, code); break;
    case 1: hipLaunchKernelGGL((k_adj_coef<-1, 1>), grid, dim3(256), 0, st, shots, nzz, nstr, k, Tnew, Told, pa, pb, lam_src, lam_dst, coef, code); break;
    case 2: hipLaunchKernelGGL((k_adj_coef<1, -1>), grid, dim3(256), 0, st, shots, nzz, nstr, k, Tnew, Told, pa, pb, lam_src, lam_dst, coef, code); break;
    default: hipLaunchKernelGGL((k_adj_coef<-1, -1>), grid, dim3(256), 0, st, shots, nzz, nstr, k, Tnew, Told, pa, pb, lam_src, lam_dst, coef, code); break;
  }
}

void launch_adj_sweep(const AdjShot* shots, int nshots, int nstr, const float* lam_in, const float4* coef,
                      float* lam_out, float* gc, int* walk_ctl, unsigned long long* walk_x, int ldx, unsigned tag,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_walk_begin, dim3(1), dim3(64), 0, st, walk_ctl);
  const int grid = std::min(nshots * nstr, 4096);
  hipLaunchKernelGGL(k_adj_sweep, dim3(grid), dim3(64), 0, st, shots, nshots, nstr, lam_in, coef, lam_out, gc,
                     walk_ctl, walk_x, ldx, tag);
}

void launch_adj_fold(int sweep, const AdjShot* shots, int nshots, int nstr, int max_nxx, int nzz, const float* gc,
                     const uint8_t* code, float* gshot, hipStream_t st) {
  const dim3 grid(nstr * ((max_nxx + ADJ_CB - 1) / ADJ_CB), nshots);
  switch (sweep) {
    case 0: hipLaunchKernelGGL((k_adj_fold<1, 1>), grid, dim3(256), 0, st, shots, nzz, nstr, gc, code, gshot); break;
    case 1: hipLaunchKernelGGL((k_adj_fold<-1, 1>), grid, dim3(256), 0, st, shots, nzz, nstr, gc, code, gshot); break;
    case 2: hipLaunchKernelGGL((k_adj_fold<1, -1>), grid, dim3(256), 0, st, shots, nzz, nstr, gc, code, gshot); break;
    default: hipLaunchKernelGGL((k_adj_fold<-1, -1>), grid, dim3(256), 0, st, shots, nzz, nstr, gc, code, gshot); break;
  }
}

void launch_adj_source(const AdjShot* shots, int nshots, int nzz, double dx, double dz, const float* lam0,
                       float* gshot, hipStream_t st) {
  const int n = nshots * 9;
  hipLaunchKernelGGL(k_adj_source, dim3((n + 255) / 256), dim3(256), 0, st, shots, nshots, nzz, dx, dz, lam0,
                     gshot);
}

void launch_adj_reduce(const AdjShot* shots, int nshots, int nz, int nx, int nzz, const float* gshot,
                       double* grad, bool fixed_border, hipStream_t st) {
  const int n = nz * nx;
  hipLaunchKernelGGL(k_adj_reduce, dim3((n + 255) / 256), dim3(256), 0, st, shots, nshots, nz, nx, nzz, gshot,
                     grad, fixed_border ? 1 : 0);
}

}  // namespace seisfat
