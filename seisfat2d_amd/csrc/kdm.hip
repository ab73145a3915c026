// Kirchhoff depth migration operators on MI355X: the image- and angle-domain
// kernels of src/migration/migration.cu:398-756 and the wavelet
// convolutions of :173-261.
//
// The reference launches one kernel per trace (source, receiver pair), one
// thread per padded node, each evaluating the node's weight from 9-point
// finite differences of both traveltime tables and adding into the image
// (adjoint) or the trace (forward) with atomicAdd.  Here:
//   - the per-table part of the weight is evaluated once per table
//     (k_kdm_prep): the reference's first and second derivatives, Hessian
//     determinant, gradient norm, incidence angle and cosine, folded into
//     one amplitude per node (the Hessian, obliquity and spreading factors of
//     the weight are products of a source and a receiver factor) and the
//     cosine and sine of the half angle (the reflection-angle factor's
//     cosine of the mean angle by the addition formula); a trace's weight is
//     then four products, the angle factor and the aperture taper (7 words
//     per node and table: T, amplitude, half-angle cosine and sine, and the
//     gradient for the angle domain's bin);
//   - the same pass keeps each table's minimum traveltime over every 64
//     consecutive image points; rounded addition and division being
//     monotone, a wave whose chunk minima give a time index >= nt holds no
//     node that reaches the trace and passes over it with two scalar loads
//     (with the reference's recording lengths most node-trace pairs end
//     there); the other quantities are loaded only for nodes that reach it;
//   - adjoint: one thread per image point walks the traces grouped in source
//     blocks (four sources' quantities in registers, each receiver's loaded
//     once for up to four traces), the blocks spread over up to 16 waves per
//     64 points; image domain: each wave's sum stays in a register and the
//     partial images are added in a fixed order, deterministic, no atomics;
//     angle domain: adds into the (depth, angle, CMP) bins with atomics, as
//     the reference (their order varies);
//   - forward: one workgroup (sixteen waves) per trace accumulates the trace
//     in an LDS histogram of its nt samples (LDS atomics) and adds it to the
//     trace once; the tables' minima over 64 chunks let it pass over 4096
//     image points with one test;
//   - convolution: the FFT product of the reference is a linear
//     (cross-)correlation with the wavelet, summed directly in double, the
//     trace and the wavelet staged in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/seisfat2d_kdm.h"
#include "eik_handle.h"

namespace seisfat {

constexpr float KDM_EPS = 1e-6f;  // EPS of src/migration/migration.cuh:7
constexpr int KQ = 7;             // per-table quantities: T, amplitude, cos and sin of half the angle, dT/dx, dT/dz, |grad T| + EPS

struct KdmK {
  int nz, nx, nzz, npts, nt, nang;
  int nchunk;     // 64-node chunks of the image points (one wave each)
  int nsup;       // super-chunks of 64 chunks (4096 image points)
  float thr;      // least x with RN(x / dt) >= nt: a traveltime sum x >= thr has time index >= nt
  float zh_dt, zl_dt;  // two-word reciprocal of dt (fdiv)
  int fastdiv_dt;      // fdiv(x, zh_dt, zl_dt) == RN(x / dt) for every 2^-100 <= x < 2^100 (checked exhaustively)
  int taper_rdiv;      // (x - cmp) / sig_e needs no operand rescaling at any node and trace (kdm_weight)
  int rec_nonneg; // every receiver-table value at an image point is >= 0 (no NaN)
  float dx, dz, dt, da;
  float ntf;     // (float)nt
  float tan_ap;  // tanf(aperture in radians), the reference's sigma / z
  float scale;   // 1 / (2 pi) as the reference rounds it
};

// float -> int rounding down, as the GPU's conversion does it (NaN -> 0,
// saturating): the reference's __float2int_rd.
__device__ __forceinline__ int f2i_rd(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return 2147483647;
  if (v < -2147483648.0f) return (-2147483647 - 1);
  return (int)floorf(v);
}

// The per-table quantities of padded table T at interior node (i, j) (padded
// coordinates).  The reference's per-table expressions (float, its order):
// the derivatives, norm = |grad T| + EPS, det = |det H| + EPS, the angle
// theta = acos(clamp(-uz)) and cos = |-uz|, rs = max(T / S, EPS).  Its
// weight's Hessian, obliquity and spreading factors,
//   sqrt(det_s det_r) S / (norm_s norm_r) * sqrt(max(0, cos_s) max(0, cos_r))
//   / sqrt(rs_s rs_r),
// are S times the product of one amplitude per table,
//   amp = sqrt(det max(0, cos)) / (norm sqrt(rs))   (in double, rounded once),
// and its angle factor's cos((theta_s + theta_r) / 2) is
// ch_s ch_r - sh_s sh_r with ch, sh = cos, sin(theta / 2).  The weight a
// trace forms from them differs from the reference's order of roundings by a
// few ulp.
//
// One wave per chunk of 64 consecutive image points (p = i + j*nz) of a table,
// which also keeps the chunk's minimum traveltime (tmin).  For every node of a
// chunk, RN(RN(Ts + Tr) / dt) >= RN(RN(min_s + min_r) / dt) (rounded addition
// and division are monotone), so a chunk whose minima give a time index >= nt
// holds no node that reaches the trace.  The minimum propagates NaN: a chunk
// holding a NaN traveltime is never passed over (its node's time index would
// be 0).
__device__ __forceinline__ float min_nan(float a, float b) { return (a != a || b != b) ? __builtin_nanf("") : fminf(a, b); }

__global__ __launch_bounds__(256) void k_kdm_prep(const float* __restrict__ tables, int ntab,
                                                  const float* __restrict__ S, KdmK K, float* __restrict__ q,
                                                  float* __restrict__ tmin) {
  const int lane = threadIdx.x & 63;
  const size_t nw = (size_t)ntab * K.nchunk;
  const size_t wstride = (size_t)gridDim.x * (blockDim.x >> 6);
  const size_t tsize = (size_t)K.nzz * (K.nx + 6);
  for (size_t w = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nw; w += wstride) {
    const int tab = (int)(w / K.nchunk);
    const int p = (int)(w - (size_t)tab * K.nchunk) * 64 + lane;
    float tv = __builtin_inff();
    if (p < K.npts) {
      const int i = p % K.nz + 3, j = p / K.nz + 3;
      const float* __restrict__ T = tables + tab * tsize;
      const int nzz = K.nzz;
      const float c = T[i + j * nzz];
      const float up = T[(i - 1) + j * nzz], dn = T[(i + 1) + j * nzz];
      const float lf = T[i + (j - 1) * nzz], rt = T[i + (j + 1) * nzz];
      const float d_z = 0.5f * (dn - up) / K.dz;
      const float d_x = 0.5f * (rt - lf) / K.dx;
      const float d2z = (dn - 2.0f * c + up) / (K.dz * K.dz);
      const float d2x = (rt - 2.0f * c + lf) / (K.dx * K.dx);
      const float dxz = (T[(i + 1) + (j + 1) * nzz] - T[(i - 1) + (j + 1) * nzz] - T[(i + 1) + (j - 1) * nzz] +
                         T[(i - 1) + (j - 1) * nzz]) /
                        (4.0f * K.dx * K.dz);
      const float norm = sqrtf(d_x * d_x + d_z * d_z) + KDM_EPS;
      const float det = fabsf(d2x * d2z - dxz * dxz) + KDM_EPS;
      const float ux = d_x / norm, uz = d_z / norm;
      const float dot = ux * 0.0f + uz * -1.0f;  // unit gradient . (0, -1), the surface normal
      const float cth = fminf(1.0f, fmaxf(-1.0f, dot));  // cos(theta), theta = acos(cth) in the reference
      const float rs = fmaxf(c / S[i + j * nzz], KDM_EPS);
      const double amp = sqrt((double)det * (double)fmaxf(0.0f, fabsf(dot))) / ((double)norm * sqrt((double)rs));
      float* __restrict__ o = q + (size_t)tab * KQ * K.npts + p;
      o[0] = c;
      o[(size_t)1 * K.npts] = (float)amp;
      // cos and sin of theta / 2 by the half-angle identities (1 +- cth exact
      // where it matters, by Sterbenz): within 5e-8 of those of acos(cth), closer
      // than the cosine and sine of the float acosf(cth)
      o[(size_t)2 * K.npts] = sqrtf(0.5f * (1.0f + cth));
      o[(size_t)3 * K.npts] = sqrtf(0.5f * (1.0f - cth));
      o[(size_t)4 * K.npts] = d_x;
      o[(size_t)5 * K.npts] = d_z;
      o[(size_t)6 * K.npts] = norm;
      tv = c;
    }
    for (int off = 32; off >= 1; off >>= 1) tv = min_nan(tv, __shfl_xor(tv, off, 64));
    if (lane == 0) tmin[w] = tv;
  }
}

// Per table, the minimum of the chunk minima over each super-chunk of 64
// chunks (NaN-propagating): one wave per super-chunk.  The super-chunk minima
// bound every chunk's from below, so a super-chunk whose minima give a time
// index >= nt holds no chunk that reaches the trace.
__global__ __launch_bounds__(256) void k_kdm_tmin2(const float* __restrict__ tmin, int ntab, KdmK K,
                                                   float* __restrict__ tmin2) {
  const int lane = threadIdx.x & 63;
  const size_t nw = (size_t)ntab * K.nsup;
  const size_t wstride = (size_t)gridDim.x * (blockDim.x >> 6);
  for (size_t w = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nw; w += wstride) {
    const int tab = (int)(w / K.nsup);
    const int c = (int)(w - (size_t)tab * K.nsup) * 64 + lane;
    float v = c < K.nchunk ? tmin[(size_t)tab * K.nchunk + c] : __builtin_inff();
    for (int off = 32; off >= 1; off >>= 1) v = min_nan(v, __shfl_xor(v, off, 64));
    if (lane == 0) tmin2[w] = v;
  }
}

struct Q7 {
  float t, amp, ch, sh, dx, dz, norm;
};
// The weight's quantities after T (loaded once the node is known to reach the
// trace).
__device__ __forceinline__ void q_amp(const float* __restrict__ q, int tab, int p, int npts, Q7& o) {
  const float* __restrict__ b = q + (size_t)tab * KQ * npts + p;
  o.amp = b[(size_t)npts];
  o.ch = b[(size_t)2 * npts];
  o.sh = b[(size_t)3 * npts];
}
// The gradient, for the angle domain's bin.
__device__ __forceinline__ void q_dir(const float* __restrict__ q, int tab, int p, int npts, Q7& o) {
  const float* __restrict__ b = q + (size_t)tab * KQ * npts + p;
  o.dx = b[(size_t)4 * npts];
  o.dz = b[(size_t)5 * npts];
  o.norm = b[(size_t)6 * npts];
}

// Time index of a node for a trace with traveltime sum xs, as far as the
// kernels use it: the reference's tId = f2i_rd(RN(xs / dt)) where that lies in
// [0, nt) or xs is NaN (0), a value >= nt or negative otherwise.  When every
// lane's sum lies in [2^-100, 2^100) (traveltimes are positive), RN(xs / dt)
// comes from the two-instruction known-divisor division the exhaustive check
// proved exact there and is floored after a clamp to nt: the index is then
// >= nt exactly when xs >= thr.  Otherwise, for the whole wave, IEEE division
// where that division is not proven, and thr, NaN and the sign decided
// explicitly.
__device__ __forceinline__ int t_index(float xs, const KdmK& K) {
  const bool mid = __float_as_uint(xs) - 0x0d800000u < 0x64000000u;  // 2^-100 <= xs < 2^100
  const uint64_t all_mid = __builtin_amdgcn_ballot_w64(mid);
  if (K.fastdiv_dt && all_mid == __builtin_amdgcn_read_exec()) {
    // min(RN(xs / dt), nt) of a positive finite quotient, nt from a scalar
    // register (no NaN: the compiler's canonicalising fminf is not needed)
    float v;
    asm("v_min_f32 %0, %1, %2" : "=v"(v) : "s"(K.ntf), "v"(fdiv(xs, K.zh_dt, K.zl_dt)));
    return (int)floorf(v);
  }
  const float v = (K.fastdiv_dt && mid) ? fdiv(xs, K.zh_dt, K.zl_dt) : xs / K.dt;
  const int it = v != v ? 0 : (int)fminf(fmaxf(floorf(v), -1.0f), (float)K.nt);
  return xs >= K.thr ? K.nt : it;
}

// Refined reciprocal of b (the r of rdiv).
__device__ __forceinline__ float recip_refined(float b) {
  const float r0 = __builtin_amdgcn_rcpf(b);
  return __builtin_fmaf(__builtin_fmaf(-b, r0, 1.0f), r0, r0);
}

// True when every node whose traveltime sum is at least RN(a + b) has a time
// index >= nt: RN(a + b) >= thr.  Exactly f2i_rd(RN(RN(a + b) / dt)) >= nt
// (NaN gives index 0 and fails the compare; +inf passes both).
__device__ __forceinline__ bool beyond(float a, float b, const KdmK& K) { return a + b >= K.thr; }

// Kirchhoff weight of one node for one trace (the product the reference
// kernels form): Hessian amplitude, obliquity and spreading (S times the two
// tables' amplitudes), the reflection-angle factor 1 + 0.2 cos of the mean
// angle (by the half angles' cosines and sines), and the aperture taper about
// the trace's CMP.  sS = S / (2 pi) and sig_e = sigma + EPS (with its refined
// reciprocal r_sig) come in evaluated per node.  The taper is the reference's
// expression and roundings: (x - cmp) / sig_e, by rdiv with the refined
// reciprocal when the host found that no operand needs rescaling (RDIV, the
// operator's taper_rdiv: every divisor in [EPS, 2^40), every dividend 0 or of
// magnitude in [2^-59, 2^41)), then expf(-0.5 a a).
// e^x for a finite x <= 0: 2^(x log2 e), x log2 e split into its rounded
// product hi and the residual fma(x, L, -hi) + x Llo (L + Llo = log2 e), the
// nearest integer n of hi applied by ldexp to the hardware 2^f of the
// remaining f in [-0.5, 0.5].  The same instructions as the library's expf
// without its overflow and non-finite handling: equal results for x >= -103.97;
// below, x is clamped to -104 and the result is 0 or the least subnormals
// (e^x < 2^-149 there).
__device__ __forceinline__ float exp_nonpos(float x) {
#pragma clang fp contract(off)  // hi - n must use the rounded hi: lo holds its residual
  x = fmaxf(x, -104.0f);
  const float hi = x * 0x1.715476p+0f;
  const float lo = __builtin_fmaf(x, 0x1.4ae0bep-26f, __builtin_fmaf(x, 0x1.715476p+0f, -hi));
  const float n = __builtin_rintf(hi);
  return __builtin_ldexpf(__builtin_amdgcn_exp2f((hi - n) + lo), (int)n);
}

template <bool RDIV>
__device__ __forceinline__ float kdm_weight(const Q7& s, const Q7& r, float sS, float x, float sig_e, float r_sig,
                                            float cmp) {
  const float rf = __builtin_fmaf(0.2f, __builtin_fmaf(s.ch, r.ch, -(s.sh * r.sh)), 1.0f);
  const float n = x - cmp;
  // RDIV: |a| < 2^61, so -0.5 a a is finite and <= 0
  const float a = RDIV ? rdiv(n, sig_e, r_sig) : n / sig_e;
  const float taper = RDIV ? exp_nonpos(-0.5f * a * a) : expf(-0.5f * a * a);
  return sS * (s.amp * r.amp) * rf * taper;
}

// Reflection-angle bin of the node for the trace (the reference's aId).
__device__ __forceinline__ int kdm_angle_bin(const Q7& s, const Q7& r, const KdmK& K) {
  const float ra = 0.5f * acosf((s.dx * r.dx + s.dz * r.dz) / (s.norm * r.norm));
  const float ang = (float)((double)(180.0f * ra) / M_PI);
  return f2i_rd(ang / K.da);
}

// Adjoint: thread per image point (a wave holds 64 consecutive points, one
// chunk).  The traces come grouped in source blocks: up to SB runs of equal
// sources (a run: consecutive traces with one source) per block, and per
// block a list of entries sorted by receiver, each one receiver table and,
// per source slot k, the trace pairing it with the block's k-th source (or
// -1): E = {receiver, t_0, ..., t_(SB-1)}.  A wave keeps the block's source
// quantities in registers and loads a receiver's quantities once per entry,
// for up to SB node-trace pairs.  It passes over a pair when the chunk minima
// keep every node beyond the last sample, over a source slot when the source
// minimum alone does (receiver tables non-negative), and loads nothing for
// either.  The blocks are split into gridDim.y consecutive parts, one wave per
// chunk and part, so that a chunk's work spreads over several waves.  Image
// domain: a part's sum stays in a register (no atomics), taking the
// contributions block by block, by receiver, by slot; with one part it is
// added to the model, with several it is written to part[h] and
// k_kdm_sum_parts adds the parts in order.  Angle domain: adds into the
// (depth, angle, CMP) bins with atomics, as the reference.
#ifndef SEISFAT_KDM_SB
#define SEISFAT_KDM_SB 4
#endif
constexpr int KDM_SB = SEISFAT_KDM_SB;
// Parts per chunk, at most (measured on a 1001 x 1001 image with 13 blocks:
// more waves per chunk, up to one block each, keep shortening both the
// 1 s and the 8 s record's adjoint).
#ifndef SEISFAT_KDM_PARTS
#define SEISFAT_KDM_PARTS 16
#endif
constexpr int KDM_PARTS = SEISFAT_KDM_PARTS;

// The angle domain asks the register allocator for 6 waves per SIMD (80
// vector registers instead of 87, no spills): 0.54 against 0.56 ms on the
// 1001 x 1001 survey.  The image domain's scalar registers allow 7; asking for
// 8 spills them into vector lanes (measured slower on long records).
template <bool ANGLE, int SB, bool RDIV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ANGLE ? 6 : 1))) void k_kdm_adjoint(const float* __restrict__ q, const float* __restrict__ tmin,
                                                     const float* __restrict__ S, const int* __restrict__ bsrc,
                                                     const int* __restrict__ bent, const int* __restrict__ ent,
                                                     const float* __restrict__ tcmp, const int* __restrict__ tcid,
                                                     int nb, int nsrc, const float* __restrict__ data,
                                                     float* __restrict__ model, float* __restrict__ part, KdmK K) {
  const int p0 = blockIdx.x * blockDim.x + threadIdx.x;
  const int nparts = gridDim.y, h = blockIdx.y;
  const int b_begin = (int)((long long)nb * h / nparts), b_end = (int)((long long)nb * (h + 1) / nparts);
  const bool live = p0 < K.npts;
  const int p = live ? p0 : K.npts - 1;  // lanes past the end load in bounds and store nothing
  const int c = __builtin_amdgcn_readfirstlane(p0 >> 6);
  if (c >= K.nchunk) return;  // whole wave past the end
  const int lane = threadIdx.x & 63;
  const int i = p % K.nz, j = p / K.nz;
  const float z = (float)i * K.dz, x = (float)j * K.dx;
  const float sS = K.scale * S[(i + 3) + (size_t)(j + 3) * K.nzz];
  const float sig_e = K.tan_ap * z + KDM_EPS;
  const float r_sig = recip_refined(sig_e);
  float acc = (!ANGLE && live && nparts == 1) ? model[p] : 0.0f;
  for (int b = b_begin; b < b_end; ++b) {
    // slot k in use: bit k (one scalar register for the block)
    float ms[SB];
    unsigned use = 0;
#pragma unroll
    for (int k = 0; k < SB; ++k) {
      const int sk = bsrc[b * SB + k];
      ms[k] = sk >= 0 ? tmin[(size_t)sk * K.nchunk + c] : 0.0f;
      const bool u = sk >= 0 && !(K.rec_nonneg && beyond(ms[k], 0.0f, K));
      use |= (unsigned)__builtin_amdgcn_readfirstlane((int)u) << k;
    }
    if (!use) continue;
    // the first group of 64 entries holding a pair that passes the chunk
    // test: the sources' quantities are loaded only when there is one
    const int e1 = bent[b + 1];
    int e_first = bent[b];
    for (; e_first < e1; e_first += 64) {
      bool need = false;
      if (e_first + lane < e1) {
        const int* __restrict__ El = ent + (size_t)(e_first + lane) * (1 + SB);
        const float mr = tmin[(size_t)(nsrc + El[0]) * K.nchunk + c];
#pragma unroll
        for (int k = 0; k < SB; ++k) need |= El[1 + k] >= 0 && (use >> k & 1) && !beyond(ms[k], mr, K);
      }
      if (__ballot(need)) break;
    }
    if (e_first >= e1) continue;
    Q7 qs[SB];
#pragma unroll
    for (int k = 0; k < SB; ++k) {
      if (use >> k & 1) {
        const int sk = bsrc[b * SB + k];
        qs[k].t = q[(size_t)sk * KQ * K.npts + p];
        q_amp(q, sk, p, K.npts, qs[k]);
        if (ANGLE) q_dir(q, sk, p, K.npts, qs[k]);
      }
    }
    for (int e0 = e_first; e0 < e1; e0 += 64) {
      // lane l holds entry e0 + l (receiver, traces, receiver minimum) and
      // tests it: does any of its pairs pass the chunk test?  The entries that
      // pass are then taken one at a time, their words read from that lane.
      int el_pass = 0;  // bit k: slot k holds a trace that passes the chunk test
      int el_r = 0, el_t[SB], el_a[SB];
      float mr_l = 0.0f, el_c[SB];
#pragma unroll
      for (int k = 0; k < SB; ++k) {
        el_t[k] = -1;
        el_a[k] = 0;
        el_c[k] = 0.0f;
      }
      if (e0 + lane < e1) {
        const int* __restrict__ El = ent + (size_t)(e0 + lane) * (1 + SB);
        el_r = nsrc + El[0];
#pragma unroll
        for (int k = 0; k < SB; ++k) el_t[k] = El[1 + k];
        mr_l = tmin[(size_t)el_r * K.nchunk + c];
#pragma unroll
        for (int k = 0; k < SB; ++k) el_pass |= (el_t[k] >= 0 && (use >> k & 1) && !beyond(ms[k], mr_l, K)) << k;
        // the traces' CMP (and bin), read from this lane when the entry is taken
        if (el_pass) {
#pragma unroll
          for (int k = 0; k < SB; ++k) {
            if (el_t[k] >= 0) {
              el_c[k] = tcmp[el_t[k]];
              if (ANGLE) el_a[k] = tcid[el_t[k]];
            }
          }
        }
      }
      uint64_t m = __ballot(el_pass != 0);
      // the next entry's receiver traveltime is loaded while this one computes
      int r_next = 0;
      float rt_next = 0.0f;
      if (m) {
        r_next = __builtin_amdgcn_readlane(el_r, (int)__builtin_ctzll(m));
        rt_next = q[(size_t)r_next * KQ * K.npts + p];
      }
      while (m) {
        const int li = (int)__builtin_ctzll(m);
        m &= m - 1;
        const int r = r_next;
        const float rt = rt_next;
        const int pass = __builtin_amdgcn_readlane(el_pass, li);
        if (m) {
          r_next = __builtin_amdgcn_readlane(el_r, (int)__builtin_ctzll(m));
          rt_next = q[(size_t)r_next * KQ * K.npts + p];
        }
        // first the time indices and the data samples of every pair, then the
        // weights: the gathers of one entry are in flight together
        // (slots whose chunk test passes: bit k, uniform)
        int tk[SB];
        float dv[SB];
        bool ok[SB];
        bool anyok = false;
        unsigned slots = 0;
#pragma unroll
        for (int k = 0; k < SB; ++k) {
          ok[k] = false;
          dv[k] = 0.0f;
          tk[k] = -1;
          if (!(pass >> k & 1)) continue;
          tk[k] = __builtin_amdgcn_readlane(el_t[k], li);
          slots |= 1u << k;
          const float xs = qs[k].t + rt;
          const int it = t_index(xs, K);
          ok[k] = live && (unsigned)it < (unsigned)K.nt;
          // the trace's base (uniform) plus a 32-bit byte offset (it < nt <= 16384)
          const char* __restrict__ dk = (const char*)(data + (size_t)tk[k] * K.nt);
          if (ok[k]) dv[k] = *(const float*)(dk + (unsigned)it * 4u);
          anyok |= ok[k];
        }
        if (!__any(anyok)) continue;  // (the while condition reads the updated m)
        Q7 qr;
        qr.t = rt;
        q_amp(q, r, p, K.npts, qr);
        if (ANGLE) q_dir(q, r, p, K.npts, qr);
        // every lane of a slot computes its weight; lanes that do not reach
        // the trace add nothing (a select, not a product: their weight may be
        // any value)
#pragma unroll
        for (int k = 0; k < SB; ++k) {
          if (!(slots >> k & 1)) continue;
          const float cmp = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(el_c[k]), li));
          const float w = kdm_weight<RDIV>(qs[k], qr, sS, x, sig_e, r_sig, cmp);
          if (ANGLE) {
            const int a = kdm_angle_bin(qs[k], qr, K);
            const int cb = __builtin_amdgcn_readlane(el_a[k], li);
            if (ok[k] && a >= 0 && a < K.nang)
              atomicAdd(model + i + (size_t)a * K.nz + (size_t)cb * K.nang * K.nz, w * dv[k]);
          } else {
            acc += ok[k] ? w * dv[k] : 0.0f;
          }
        }
      }
    }
  }
  if (!ANGLE && live) {
    if (nparts == 1)
      model[p] = acc;
    else
      part[(size_t)h * K.npts + p] = acc;
  }
}

// model[p] += part[0][p] + part[1][p] + ... in that order (the image adjoint's
// partial sums).
__global__ __launch_bounds__(256) void k_kdm_sum_parts(float* __restrict__ model, const float* __restrict__ part,
                                                       int nparts, int npts) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  float v = model[p];
  for (int h = 0; h < nparts; ++h) v += part[(size_t)h * npts + p];
  model[p] = v;
}

// Forward (image or angle domain): one workgroup per trace, the trace's nt
// samples in LDS; a wave passes over 4096 image points whose super-chunk
// minima keep them beyond the last sample with one test, then over 64 when
// the chunk minima do.
// Threads per trace (one workgroup, sixteen waves sharing the trace's LDS
// histogram).  Measured on the 1001 x 1001 survey against 256 and 512: the
// 1 s record's forward 0.224 ms against 0.269 and 0.233 (image), 0.32 against
// 0.44 and 0.37 (angle); the 8 s record's 19.5 ms against 19.9 and 19.7.
#ifndef SEISFAT_KDM_FWD_THREADS
#define SEISFAT_KDM_FWD_THREADS 1024
#endif
constexpr int KDM_FWD_THREADS = SEISFAT_KDM_FWD_THREADS;

template <bool ANGLE, bool RDIV>
__global__ __launch_bounds__(KDM_FWD_THREADS) void k_kdm_forward(const float* __restrict__ q, const float* __restrict__ tmin,
                                                     const float* __restrict__ S, const int* __restrict__ tsrc,
                                                     const int* __restrict__ trec, const float* __restrict__ tcmp,
                                                     const int* __restrict__ tcid, const float* __restrict__ tmin2,
                                                     int nsrc,
                                                     const float* __restrict__ model, float* __restrict__ data,
                                                     KdmK K) {
  extern __shared__ float hist[];
  const int t = blockIdx.x;
  for (int k = threadIdx.x; k < K.nt; k += blockDim.x) hist[k] = 0.0f;
  __syncthreads();
  const int s = tsrc[t], r = nsrc + trec[t];
  const float cmp = tcmp[t];
  const size_t moff = ANGLE ? (size_t)tcid[t] * K.nang * K.nz : 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  // each wave takes super-chunks wave, wave + nwaves, ..., tests them 64 at a
  // time (one per lane), and in each that passes tests its 64 chunks (one per
  // lane) and works through the chunks that pass
  for (int k0 = 0; wave + k0 * nwaves < K.nsup; k0 += 64) {
    const int sl = wave + (k0 + lane) * nwaves;
    const bool need2 = sl < K.nsup && !beyond(tmin2[(size_t)s * K.nsup + sl], tmin2[(size_t)r * K.nsup + sl], K);
    for (uint64_t m2 = __ballot(need2); m2; m2 &= m2 - 1) {
      const int c0 = (wave + (k0 + (int)__builtin_ctzll(m2)) * nwaves) * 64;
      const int cl = c0 + lane;
      const bool need = cl < K.nchunk && !beyond(tmin[(size_t)s * K.nchunk + cl], tmin[(size_t)r * K.nchunk + cl], K);
      for (uint64_t m = __ballot(need); m; m &= m - 1) {
        const int p = (c0 + __builtin_ctzll(m)) * 64 + lane;
        if (p >= K.npts) continue;
        Q7 qs, qr;
        qs.t = q[(size_t)s * KQ * K.npts + p];
        qr.t = q[(size_t)r * KQ * K.npts + p];
        const float xs = qs.t + qr.t;
        const int it = t_index(xs, K);
        if ((unsigned)it >= (unsigned)K.nt) continue;
        q_amp(q, s, p, K.npts, qs);
        q_amp(q, r, p, K.npts, qr);
        if (ANGLE) {
          q_dir(q, s, p, K.npts, qs);
          q_dir(q, r, p, K.npts, qr);
        }
        const int i = p % K.nz, j = p / K.nz;
        const float z = (float)i * K.dz, x = (float)j * K.dx;
        const float sS = K.scale * S[(i + 3) + (size_t)(j + 3) * K.nzz];
        const float sig_e = K.tan_ap * z + KDM_EPS;
        const float w = kdm_weight<RDIV>(qs, qr, sS, x, sig_e, recip_refined(sig_e), cmp);
        if (ANGLE) {
          const int a = kdm_angle_bin(qs, qr, K);
          if (a >= 0 && a < K.nang) atomicAdd(hist + it, w * model[moff + i + (size_t)a * K.nz]);
        } else {
          atomicAdd(hist + it, w * model[p]);
        }
      }
    }
  }
  __syncthreads();
  float* __restrict__ d = data + (size_t)t * K.nt;
  for (int k = threadIdx.x; k < K.nt; k += blockDim.x) d[k] += hist[k];
}

// Trace convolution: one thread per output sample.  Adjoint: out[k] =
// nfft * sum_m w[m] in[k - shift + m] for k >= shift (0 before); forward:
// out[k] = nfft * sum_m in[m] w[k + shift - m].  Sums in double, in the
// order of m.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_trace_conv(const float* __restrict__ in, float* __restrict__ out, int ntr,
                                                    int nt, const float* __restrict__ w, int nw, int shift,
                                                    double nfft) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (g >= (size_t)ntr * nt) return;
  const int t = (int)(g / nt), k = (int)(g - (size_t)t * nt);
  const float* __restrict__ d = in + (size_t)t * nt;
  double acc = 0.0;
  if (ADJ) {
    if (k < shift) {
      out[g] = 0.0f;
      return;
    }
    const int k0 = k - shift;
    const int mmax = min(nw - 1, nt - 1 - k0);
    for (int m = 0; m <= mmax; ++m) acc += (double)w[m] * (double)d[k0 + m];
  } else {
    const int kk = k + shift;
    for (int m = max(0, kk - nw + 1); m <= min(kk, nt - 1); ++m) acc += (double)d[m] * (double)w[kk - m];
  }
  out[g] = (float)(acc * nfft);
}

// The same sums with one workgroup per trace, the trace and the wavelet staged
// in LDS (nt + nw <= 16384 floats): every tap is read from LDS instead of
// global memory.  Same operands, same order: the same bits.
template <bool ADJ>
__global__ __launch_bounds__(256) void k_trace_conv_lds(const float* __restrict__ in, float* __restrict__ out,
                                                        int nt, const float* __restrict__ w, int nw, int shift,
                                                        double nfft) {
  extern __shared__ float sm[];
  float* tr = sm;
  float* wl = sm + nt;
  const size_t t = blockIdx.x;
  const float* __restrict__ d = in + t * nt;
  for (int k = threadIdx.x; k < nt; k += blockDim.x) tr[k] = d[k];
  for (int k = threadIdx.x; k < nw; k += blockDim.x) wl[k] = w[k];
  __syncthreads();
  float* __restrict__ o = out + t * nt;
  for (int k = threadIdx.x; k < nt; k += blockDim.x) {
    double acc = 0.0;
    if (ADJ) {
      if (k < shift) {
        o[k] = 0.0f;
        continue;
      }
      const int k0 = k - shift;
      const int mmax = min(nw - 1, nt - 1 - k0);
      for (int m = 0; m <= mmax; ++m) acc += (double)wl[m] * (double)tr[k0 + m];
    } else {
      const int kk = k + shift;
      for (int m = max(0, kk - nw + 1); m <= min(kk, nt - 1); ++m) acc += (double)tr[m] * (double)wl[kk - m];
    }
    o[k] = (float)(acc * nfft);
  }
}

}  // namespace seisfat

using namespace seisfat;

struct seisfat_kdm_s {
  seisfat_eik_s* h = nullptr;
  seisfat_kdm_params p{};
  int nsrc = 0, nrec = 0, ntr = 0;
  KdmK K{};
  float* d_q = nullptr;    // per-table quantities, (nsrc + nrec) x 7 x npts
  float* d_tmin = nullptr; // per-table chunk minima of T, (nsrc + nrec) x nchunk
  float* d_tmin2 = nullptr; // per-table minima over 64 chunks, (nsrc + nrec) x nsup
  int* d_src = nullptr;    // trace tables
  int* d_rec = nullptr;
  int nb = 0;              // source blocks of the adjoint
  int* d_bsrc = nullptr;   // per block and slot: source table (-1: none)
  int* d_bent = nullptr;   // per block: first entry (nb + 1)
  int* d_ent = nullptr;    // entries {receiver, trace per slot}
  float* d_cmp = nullptr;
  int* d_cid = nullptr;
  int nparts = 1;            // adjoint waves per chunk (block ranges)
  float* d_part = nullptr;   // image adjoint partial sums, nparts x npts (nparts > 1)
  std::vector<char> tab_set;  // per table (sources, then receivers): quantities evaluated
  int tabs_missing = 0;
  unsigned s_version = 0;      // the handle's slowness the set tables were evaluated with
  bool taper_geom = false;     // the divisors and x positions allow rdiv for the taper (taper_geometry_ok)
  bool rec_dirty = true;       // receiver minima changed since rec_nonneg was derived
};

namespace {

int kfail(int code, const std::string& m) { return seisfat::set_error(code, m.c_str()); }

// The least float x with RN(x / dt) >= nt (+inf when none): RN(x / dt) is
// monotone in x, so a bisection over the bit patterns of x >= 0 finds it.
float index_threshold(float dt, int nt) {
  auto ge = [&](float x) {
    volatile float v = x / dt;
    return v >= (float)nt;
  };
  const float big = std::numeric_limits<float>::max();
  if (!ge(big)) return std::numeric_limits<float>::infinity();
  if (ge(0.0f)) return 0.0f;
  uint32_t lo = 0, hi;
  std::memcpy(&hi, &big, 4);
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    float x;
    std::memcpy(&x, &mid, 4);
    (ge(x) ? hi : lo) = mid;
  }
  float x;
  std::memcpy(&x, &hi, 4);
  return x;
}

#define KHIP(expr)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = (expr);                                                                         \
    if (e_ != hipSuccess)                                                                           \
      return kfail(e_ == hipErrorOutOfMemory ? SEISFAT_ENOMEM : SEISFAT_EHIP,                       \
                   std::string(#expr) + ": " + hipGetErrorString(e_));                              \
  } while (0)

template <class T>
void kfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

// The aperture quotient (x - cmp) / sig_e needs no operand rescaling (rdiv is
// then IEEE division) when every divisor sig_e = tan_ap z + EPS lies in
// [EPS, 2^40) (tan_ap >= 0: it grows with z, largest at the last row) and
// every x = j dx and every CMP is 0 or of magnitude in [2^-36, 2^40): the
// difference of two such floats is 0 or of magnitude in [2^-59, 2^41) (both
// are multiples of 2^-59), and the quotient normal.  Here the divisors and the
// x positions; set_traces checks the CMPs.
bool taper_geometry_ok(const KdmK& K) {
  if (!(K.tan_ap >= 0.0f) || !std::isfinite(K.tan_ap)) return false;
  const double zmax = (double)(K.nz - 1) * K.dz, xmax = (double)(K.nx - 1) * K.dx;
  if (!((double)K.tan_ap * zmax + KDM_EPS <= 0x1p39)) return false;
  return (K.nx == 1 || K.dx >= 0x1p-36f) && xmax <= 0x1p39;
}

// The quantities and chunk minima of ntab consecutive tables (padded grids at
// d_tables) into table slots tab0, tab0 + 1, ...: one wave per chunk.
void launch_prep(seisfat_kdm_s* k, const float* d_tables, int tab0, int ntab, hipStream_t st) {
  const size_t nw = (size_t)ntab * k->K.nchunk;
  const int blocks = (int)std::min<size_t>((nw + 3) / 4, 16384);
  hipLaunchKernelGGL(k_kdm_prep, dim3(blocks), dim3(256), 0, st, d_tables, ntab, k->h->d_spad, k->K,
                     k->d_q + (size_t)tab0 * KQ * k->K.npts, k->d_tmin + (size_t)tab0 * k->K.nchunk);
  const size_t ns = (size_t)ntab * k->K.nsup;
  const int blocks2 = (int)std::min<size_t>((ns + 3) / 4, 16384);
  hipLaunchKernelGGL(k_kdm_tmin2, dim3(blocks2), dim3(256), 0, st, k->d_tmin + (size_t)tab0 * k->K.nchunk, ntab,
                     k->K, k->d_tmin2 + (size_t)tab0 * k->K.nsup);
}

// The operators' launches, one instantiation per domain and taper division.
template <bool ANGLE, bool RDIV>
void launch_adjoint(seisfat_kdm_s* k, const float* d_data, float* d_model, hipStream_t st) {
  const dim3 grid((k->K.npts + 255) / 256, k->nparts);
  hipLaunchKernelGGL((k_kdm_adjoint<ANGLE, KDM_SB, RDIV>), grid, dim3(256), 0, st, k->d_q, k->d_tmin, k->h->d_spad,
                     k->d_bsrc, k->d_bent, k->d_ent, k->d_cmp, k->d_cid, k->nb, k->nsrc, d_data, d_model,
                     ANGLE ? nullptr : k->d_part, k->K);
}

template <bool ANGLE, bool RDIV>
void launch_forward(seisfat_kdm_s* k, const float* d_model, float* d_data, hipStream_t st) {
  const size_t lds = (size_t)k->K.nt * sizeof(float);
  hipLaunchKernelGGL((k_kdm_forward<ANGLE, RDIV>), dim3(k->ntr), dim3(KDM_FWD_THREADS), lds, st, k->d_q, k->d_tmin,
                     k->h->d_spad, k->d_src, k->d_rec, k->d_cmp, k->d_cid, k->d_tmin2, k->nsrc, d_model, d_data,
                     k->K);
}

bool cmp_ok(float c) { return c == 0.0f || (std::fabs(c) >= 0x1p-36f && std::fabs(c) < 0x1p40f); }

// The table quantities hold the slowness they were evaluated with (the
// amplitudes): a slowness upload to the handle makes every table set before it
// stale, to be set again.
void track_slowness(seisfat_kdm_s* k) {
  if (k->s_version == k->h->version) return;
  std::fill(k->tab_set.begin(), k->tab_set.end(), 0);
  k->tabs_missing = k->nsrc + k->nrec;
  k->s_version = k->h->version;
}

void free_traces(seisfat_kdm_s* k) {
  kfree(k->d_src);
  kfree(k->d_rec);
  kfree(k->d_cmp);
  kfree(k->d_cid);
  kfree(k->d_bsrc);
  kfree(k->d_bent);
  kfree(k->d_ent);
  kfree(k->d_part);
  k->ntr = k->nb = 0;
  k->nparts = 1;
}

}  // namespace

extern "C" {

int seisfat_kdm_create(seisfat_eik_t h, const seisfat_kdm_params* p, int nsrc, int nrec, seisfat_kdm_t* out) {
  if (!out) return kfail(SEISFAT_EINVAL, "null output pointer");
  *out = nullptr;
  if (!h || !p) return kfail(SEISFAT_EINVAL, "null handle or parameters");
  if (!h->has_s) return kfail(SEISFAT_ESTATE, "slowness not set (seisfat_eik_set_slowness)");
  if (p->domain != SEISFAT_KDM_IMAGE && p->domain != SEISFAT_KDM_ANGLE)
    return kfail(SEISFAT_EINVAL, "domain must be SEISFAT_KDM_IMAGE (0) or SEISFAT_KDM_ANGLE (1)");
  if (p->nt <= 0 || p->nt > 16384) return kfail(SEISFAT_EINVAL, "nt must lie in [1, 16384]");
  if (!(p->dt > 0.0f) || !std::isfinite(p->dt)) return kfail(SEISFAT_EINVAL, "dt must be positive");
  if (!std::isfinite(p->aperture)) return kfail(SEISFAT_EINVAL, "aperture must be finite");
  if (p->domain == SEISFAT_KDM_ANGLE && (p->nang <= 0 || p->ncmp <= 0 || !(p->da > 0.0f)))
    return kfail(SEISFAT_EINVAL, "angle domain: nang and ncmp must be positive, da positive");
  if (nsrc <= 0 || nrec <= 0) return kfail(SEISFAT_EINVAL, "nsrc and nrec must be positive");
  const size_t npts = (size_t)h->nz * h->nx;
  if (npts > (size_t)1 << 30) return kfail(SEISFAT_EINVAL, "model too large");
  auto* k = new (std::nothrow) seisfat_kdm_s();
  if (!k) return kfail(SEISFAT_ENOMEM, "host allocation");
  k->h = h;
  k->p = *p;
  k->nsrc = nsrc;
  k->nrec = nrec;
  KdmK& K = k->K;
  K.nz = h->nz;
  K.nx = h->nx;
  K.nzz = h->nzz;
  K.npts = (int)npts;
  K.nchunk = (int)((npts + 63) / 64);
  K.nsup = (K.nchunk + 63) / 64;
  K.rec_nonneg = 0;
  K.nt = p->nt;
  K.ntf = (float)p->nt;
  K.nang = p->domain == SEISFAT_KDM_ANGLE ? p->nang : 0;
  K.dx = h->dx;
  K.dz = h->dz;
  K.dt = p->dt;
  K.thr = index_threshold(p->dt, p->nt);
  // two-word reciprocal of dt: zh = RN(1/dt); 1 - dt*zh is exact (one fma)
  K.zh_dt = 1.0f / p->dt;
  K.zl_dt = std::fma(-p->dt, K.zh_dt, 1.0f) / p->dt;
  K.fastdiv_dt = 0;
  K.da = p->domain == SEISFAT_KDM_ANGLE ? p->da : 1.0f;
  // sigma = tanf(aperture * pi / 180) * z, the angle rounded to float as the
  // reference's float argument takes it; 1 / (2 pi) rounded from double
  K.tan_ap = std::tan((float)((double)p->aperture * M_PI / 180.0));
  K.scale = (float)(1.0 / (2.0 * M_PI));
  hipError_t e = hipMalloc((void**)&k->d_q, (size_t)(nsrc + nrec) * KQ * npts * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&k->d_tmin, (size_t)(nsrc + nrec) * K.nchunk * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&k->d_tmin2, (size_t)(nsrc + nrec) * K.nsup * sizeof(float));
  if (e == hipSuccess) {
    // exhaustive check of the known-divisor division by dt (every dividend of
    // magnitude in [2^-100, 2^100)) against IEEE division
    float* d = nullptr;
    const float w[3] = {K.dt, K.zh_dt, K.zl_dt};
    int bad = 1;
    e = hipMalloc((void**)&d, 4 * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(d, w, sizeof(w), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d + 3, 0, sizeof(int));
    if (e == hipSuccess) {
      launch_fastdiv_check(d, d + 1, d + 2, 1, (int*)(d + 3), nullptr);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&bad, d + 3, sizeof(int), hipMemcpyDeviceToHost);
    if (d) (void)hipFree(d);
    K.fastdiv_dt = (e == hipSuccess && bad == 0) ? 1 : 0;
  }
  if (e != hipSuccess) {
    kfree(k->d_q);
    kfree(k->d_tmin);
    kfree(k->d_tmin2);
    delete k;
    return kfail(SEISFAT_ENOMEM, std::string("hipMalloc of the table quantities: ") + hipGetErrorString(e));
  }
  k->tab_set.assign(nsrc + nrec, 0);
  k->tabs_missing = nsrc + nrec;
  k->s_version = h->version;
  k->taper_geom = taper_geometry_ok(K);
  K.taper_rdiv = 0;
  *out = k;
  return SEISFAT_OK;
}

int seisfat_kdm_set_tables(seisfat_kdm_t k, const float* d_src, const float* d_rec, void* stream) {
  if (!k || !d_src || !d_rec) return kfail(SEISFAT_EINVAL, "null operator or table pointer");
  hipStream_t st = (hipStream_t)stream;
  track_slowness(k);
  launch_prep(k, d_src, 0, k->nsrc, st);
  launch_prep(k, d_rec, k->nsrc, k->nrec, st);
  KHIP(hipGetLastError());
  std::fill(k->tab_set.begin(), k->tab_set.end(), 1);
  k->tabs_missing = 0;
  k->rec_dirty = true;
  return SEISFAT_OK;
}

int seisfat_kdm_set_table(seisfat_kdm_t k, int which, int index, const float* d_table, void* stream) {
  if (!k || !d_table) return kfail(SEISFAT_EINVAL, "null operator or table pointer");
  if (which != 0 && which != 1) return kfail(SEISFAT_EINVAL, "which must be 0 (source) or 1 (receiver)");
  const int n = which ? k->nrec : k->nsrc;
  if (index < 0 || index >= n) return kfail(SEISFAT_EINVAL, "table index out of range");
  const int tab = (which ? k->nsrc : 0) + index;
  hipStream_t st = (hipStream_t)stream;
  track_slowness(k);
  launch_prep(k, d_table, tab, 1, st);
  KHIP(hipGetLastError());
  if (!k->tab_set[tab]) {
    k->tab_set[tab] = 1;
    --k->tabs_missing;
  }
  if (which) k->rec_dirty = true;
  return SEISFAT_OK;
}

namespace {

// Before a launch: every table set, and rec_nonneg derived from the current
// receiver minima (a source's run of traces may be passed over on the source
// minimum alone when no receiver traveltime is negative or NaN).
int ready(seisfat_kdm_s* k, hipStream_t st) {
  if (k->s_version != k->h->version)
    return kfail(SEISFAT_ESTATE, "the slowness changed since the tables were set (set them again)");
  if (k->tabs_missing) return kfail(SEISFAT_ESTATE, "tables not set (seisfat_kdm_set_tables / _set_table)");
  if (k->rec_dirty) {
    std::vector<float> rmin((size_t)k->nrec * k->K.nchunk);
    KHIP(hipMemcpyAsync(rmin.data(), k->d_tmin + (size_t)k->nsrc * k->K.nchunk, rmin.size() * sizeof(float),
                        hipMemcpyDeviceToHost, st));
    KHIP(hipStreamSynchronize(st));
    k->K.rec_nonneg = std::all_of(rmin.begin(), rmin.end(), [](float v) { return v >= 0.0f; }) ? 1 : 0;
    k->rec_dirty = false;
  }
  return SEISFAT_OK;
}

}  // namespace

int seisfat_kdm_set_traces(seisfat_kdm_t k, int ntr, const int* src, const int* rec, const float* cmp,
                           const int* cmp_id) {
  if (!k) return kfail(SEISFAT_EINVAL, "null operator");
  if (ntr < 0 || (ntr > 0 && (!src || !rec || !cmp))) return kfail(SEISFAT_EINVAL, "bad trace arrays");
  if (k->p.domain == SEISFAT_KDM_ANGLE && ntr > 0 && !cmp_id)
    return kfail(SEISFAT_EINVAL, "the angle domain needs CMP bins");
  for (int t = 0; t < ntr; ++t) {
    if (src[t] < 0 || src[t] >= k->nsrc || rec[t] < 0 || rec[t] >= k->nrec)
      return kfail(SEISFAT_EINVAL, "trace " + std::to_string(t) + ": table index out of range");
    if (!std::isfinite(cmp[t])) return kfail(SEISFAT_EINVAL, "trace " + std::to_string(t) + ": CMP not finite");
    if (k->p.domain == SEISFAT_KDM_ANGLE && (cmp_id[t] < 0 || cmp_id[t] >= k->p.ncmp))
      return kfail(SEISFAT_EINVAL, "trace " + std::to_string(t) + ": CMP bin out of range");
  }
  KHIP(hipDeviceSynchronize());  // runs in flight may read the previous trace tables
  free_traces(k);
  k->K.taper_rdiv = (k->taper_geom && std::all_of(cmp, cmp + ntr, cmp_ok)) ? 1 : 0;
  if (ntr == 0) return SEISFAT_OK;
  KHIP(hipMalloc((void**)&k->d_src, ntr * sizeof(int)));
  KHIP(hipMalloc((void**)&k->d_rec, ntr * sizeof(int)));
  KHIP(hipMalloc((void**)&k->d_cmp, ntr * sizeof(float)));
  KHIP(hipMalloc((void**)&k->d_cid, ntr * sizeof(int)));
  // source blocks for the adjoint: SB runs of equal sources per block, each
  // block's pairs sorted by (receiver, slot), traces in order within equal keys
  std::vector<int> run0;
  for (int t = 0; t < ntr; ++t)
    if (t == 0 || src[t] != src[t - 1]) run0.push_back(t);
  run0.push_back(ntr);
  const int nruns = (int)run0.size() - 1;
  const int nb = (nruns + KDM_SB - 1) / KDM_SB;
  std::vector<int> bsrc((size_t)nb * KDM_SB, -1), bent(nb + 1, 0), ent;
  struct Pair {
    int r, k, t;
  };
  std::vector<Pair> pairs;
  for (int b = 0; b < nb; ++b) {
    pairs.clear();
    for (int k = 0; k < KDM_SB && b * KDM_SB + k < nruns; ++k) {
      const int r0 = run0[b * KDM_SB + k], r1 = run0[b * KDM_SB + k + 1];
      bsrc[(size_t)b * KDM_SB + k] = src[r0];
      for (int t = r0; t < r1; ++t) pairs.push_back({rec[t], k, t});
    }
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const Pair& a, const Pair& c) { return a.r != c.r ? a.r < c.r : a.k < c.k; });
    size_t cur = (size_t)-1;
    for (const Pair& pr : pairs) {
      if (cur == (size_t)-1 || ent[cur] != pr.r || ent[cur + 1 + pr.k] != -1) {
        cur = ent.size();
        ent.push_back(pr.r);
        ent.insert(ent.end(), KDM_SB, -1);
      }
      ent[cur + 1 + pr.k] = pr.t;
    }
    bent[b + 1] = (int)(ent.size() / (1 + KDM_SB));
  }
  KHIP(hipMalloc((void**)&k->d_bsrc, bsrc.size() * sizeof(int)));
  KHIP(hipMalloc((void**)&k->d_bent, bent.size() * sizeof(int)));
  KHIP(hipMalloc((void**)&k->d_ent, ent.size() * sizeof(int)));
  KHIP(hipMemcpy(k->d_bsrc, bsrc.data(), bsrc.size() * sizeof(int), hipMemcpyHostToDevice));
  KHIP(hipMemcpy(k->d_bent, bent.data(), bent.size() * sizeof(int), hipMemcpyHostToDevice));
  KHIP(hipMemcpy(k->d_ent, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice));
  k->nb = nb;
  // adjoint parts: up to KDM_PARTS, no more than the blocks; the image
  // domain's partial sums within 1 GiB
  int np = std::max(1, std::min(KDM_PARTS, nb));
  if (k->p.domain == SEISFAT_KDM_IMAGE) {
    while (np > 1 && (size_t)np * k->K.npts * sizeof(float) > ((size_t)1 << 30)) --np;
    if (np > 1) KHIP(hipMalloc((void**)&k->d_part, (size_t)np * k->K.npts * sizeof(float)));
  }
  k->nparts = np;
  std::vector<int> cid(ntr, 0);
  if (k->p.domain == SEISFAT_KDM_ANGLE) cid.assign(cmp_id, cmp_id + ntr);
  KHIP(hipMemcpy(k->d_src, src, ntr * sizeof(int), hipMemcpyHostToDevice));
  KHIP(hipMemcpy(k->d_rec, rec, ntr * sizeof(int), hipMemcpyHostToDevice));
  KHIP(hipMemcpy(k->d_cmp, cmp, ntr * sizeof(float), hipMemcpyHostToDevice));
  KHIP(hipMemcpy(k->d_cid, cid.data(), ntr * sizeof(int), hipMemcpyHostToDevice));
  k->ntr = ntr;
  return SEISFAT_OK;
}

int seisfat_kdm_adjoint(seisfat_kdm_t k, const float* d_data, float* d_model, void* stream) {
  if (!k) return kfail(SEISFAT_EINVAL, "null operator");
  hipStream_t st = (hipStream_t)stream;
  if (const int rc = ready(k, st)) return rc;
  if (k->ntr == 0) return SEISFAT_OK;  // no traces: the model is unchanged
  if (!d_data || !d_model) return kfail(SEISFAT_EINVAL, "null buffer");
  const bool rd = k->K.taper_rdiv != 0;
  if (k->p.domain == SEISFAT_KDM_IMAGE) {
    if (rd)
      launch_adjoint<false, true>(k, d_data, d_model, st);
    else
      launch_adjoint<false, false>(k, d_data, d_model, st);
    if (k->nparts > 1)
      hipLaunchKernelGGL(k_kdm_sum_parts, dim3((k->K.npts + 255) / 256), dim3(256), 0, st, d_model, k->d_part,
                         k->nparts, k->K.npts);
  } else if (rd) {
    launch_adjoint<true, true>(k, d_data, d_model, st);
  } else {
    launch_adjoint<true, false>(k, d_data, d_model, st);
  }
  KHIP(hipGetLastError());
  return SEISFAT_OK;
}

int seisfat_kdm_forward(seisfat_kdm_t k, const float* d_model, float* d_data, void* stream) {
  if (!k) return kfail(SEISFAT_EINVAL, "null operator");
  hipStream_t st = (hipStream_t)stream;
  if (const int rc = ready(k, st)) return rc;
  if (k->ntr == 0) return SEISFAT_OK;  // no traces to add into
  if (!d_data || !d_model) return kfail(SEISFAT_EINVAL, "null buffer");
  const bool rd = k->K.taper_rdiv != 0;
  if (k->p.domain == SEISFAT_KDM_IMAGE) {
    if (rd)
      launch_forward<false, true>(k, d_model, d_data, st);
    else
      launch_forward<false, false>(k, d_model, d_data, st);
  } else if (rd) {
    launch_forward<true, true>(k, d_model, d_data, st);
  } else {
    launch_forward<true, false>(k, d_model, d_data, st);
  }
  KHIP(hipGetLastError());
  return SEISFAT_OK;
}

int seisfat_kdm_destroy(seisfat_kdm_t k) {
  if (!k) return SEISFAT_OK;
  (void)hipDeviceSynchronize();
  free_traces(k);
  kfree(k->d_q);
  kfree(k->d_tmin);
  kfree(k->d_tmin2);
  delete k;
  return SEISFAT_OK;
}

int seisfat_trace_convolve(int adjoint, int ntr, int nt, const float* wavelet, int nw, const float* d_in,
                           float* d_out, void* stream) {
  if (ntr < 0 || nt <= 0 || nw <= 0 || !wavelet || (ntr > 0 && (!d_in || !d_out || d_in == d_out)))
    return kfail(SEISFAT_EINVAL, "bad convolution arguments (distinct device buffers, positive sizes)");
  if (ntr == 0) return SEISFAT_OK;
  // nfft = the power of two >= nt + nw - 1 (the reference's nextpow2)
  long long nfft = 1;
  while (nfft < (long long)nt + nw - 1) nfft <<= 1;
  const int shift = nw / 2 + nw / 10;
  hipStream_t st = (hipStream_t)stream;
  float* d_w = nullptr;
  KHIP(hipMallocAsync((void**)&d_w, nw * sizeof(float), st));
  KHIP(hipMemcpyAsync(d_w, wavelet, nw * sizeof(float), hipMemcpyHostToDevice, st));
  const size_t n = (size_t)ntr * nt;
  if ((long long)nt + nw <= 16384) {
    const size_t lds = (size_t)(nt + nw) * sizeof(float);
    if (adjoint)
      hipLaunchKernelGGL(k_trace_conv_lds<true>, dim3(ntr), dim3(256), lds, st, d_in, d_out, nt, d_w, nw, shift,
                         (double)nfft);
    else
      hipLaunchKernelGGL(k_trace_conv_lds<false>, dim3(ntr), dim3(256), lds, st, d_in, d_out, nt, d_w, nw, shift,
                         (double)nfft);
  } else {
    const int blocks = (int)((n + 255) / 256);
    if (adjoint)
      hipLaunchKernelGGL(k_trace_conv<true>, dim3(blocks), dim3(256), 0, st, d_in, d_out, ntr, nt, d_w, nw, shift,
                         (double)nfft);
    else
      hipLaunchKernelGGL(k_trace_conv<false>, dim3(blocks), dim3(256), 0, st, d_in, d_out, ntr, nt, d_w, nw, shift,
                         (double)nfft);
  }
  KHIP(hipGetLastError());
  KHIP(hipFreeAsync(d_w, st));
  KHIP(hipStreamSynchronize(st));  // the host wavelet is the caller's
  return SEISFAT_OK;
}

int seisfat_ricker(float dt, float fmax, float* out, int cap) {
  if (!(dt > 0.0f) || !(fmax > 0.0f) || !std::isfinite(dt) || !std::isfinite(fmax) || cap < 0 || (cap > 0 && !out))
    return kfail(SEISFAT_EINVAL, "ricker: dt and fmax must be positive, out non-null when cap > 0");
  const float t0 = 2.0f * std::sqrt((float)M_PI) / fmax;
  const float fc = fmax / (3.0f * std::sqrt((float)M_PI));
  const float len = (t0 - 0.5f * dt) / dt;
  if (!(len < 1e9f)) return kfail(SEISFAT_EINVAL, "ricker: wavelet too long (dt too small for fmax)");
  const int nw = std::max(2 * (int)len, 0);
  for (int k = 0; k < std::min(nw, cap); ++k) {
    const float td = k * dt - t0;
    const float arg = (float)(M_PI * M_PI * M_PI * fc * fc * td * td);  // double product, as the reference's
    out[k] = (1.0f - 2.0f * arg) * std::exp(-arg);
  }
  return nw;
}

}  // extern "C"
