// Device-side types and the eikonal local solver shared by every sweep kernel.
//
// Arithmetic contract: all kernels are compiled with -ffp-contract=off and
// HIP's correctly rounded fp32 division and square root, so each expression
// below rounds exactly as the reference's C++ reads (no FMA contraction).
// The result is bit-identical to the CPU oracle for every node.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace seisfat {

constexpr int NB = 3;  // replicated border width (src/modeling/modeling.cu:113)
constexpr float TINF = 1e6f;  // "not reached" (src/modeling/modeling.cu:286-291)

// Uniform constants of one model geometry.  hyp, sum and prod are the
// sub-expressions the reference evaluates per thread from dx, dz
// (src/modeling/modeling.cu:337-351); precomputing them with the same float
// operations gives the same values.
struct EikConsts {
  float dx, dz;
  float dx2i, dz2i;  // 1/dx^2, 1/dz^2   (src/modeling/modeling.cu:64-65)
  float hyp;         // sqrtf(dx*dx + dz*dz)
  float sum;         // dz2i + dx2i
  float prod;        // dz2i * dx2i
  float sum4;        // 4 * sum (exact)
  // refined reciprocals of the four divisors (see rdiv below), produced on
  // the device by k_recip so they are exactly the values the hardware
  // division sequence would build
  float r_sum, r_hyp, r_dz, r_dx;
  // two-word reciprocals zh = RN(1/b), zl = RN(1/b - zh) of the four
  // divisors (see fdiv below)
  float zh_sum, zl_sum, zh_hyp, zl_hyp, zh_dz, zl_dz, zh_dx, zl_dx;
  // host-side: fdiv reproduces the correctly rounded quotient for all four
  // divisors (checked exhaustively when the geometry is created)
  int fastdiv;
  // a node whose old traveltime and upwind neighbours are all TINF keeps
  // TINF: its 1-D candidates are >= TINF, branches B and C are not taken
  // (te - tev = tv - tev = 0) and branch A's value is at least
  // RN(RN(RN(TINF dz2i) + RN(TINF dx2i)) / sum), which this geometry keeps
  // >= TINF (set when the geometry is made)
  int tinf_fixed;
};

// ---------------------------------------------------------------------------
// Correctly rounded division and square root for normal-range operands.
//
// HIP's correctly rounded fp32 division is, for operands whose exponents need
// no rescaling and whose result is finite and normal (v_div_scale and
// v_div_fixup are then identities):
//   r0 = rcp(b); e = fma(-b, r0, 1); r = fma(e, r0, r0)        (refined 1/b)
//   q0 = a*r; e1 = fma(-b, q0, a); q1 = fma(e1, r, q0); e2 = fma(-b, q1, a);
//   q  = fma(e2, r, q1)
// Every divisor of the local solver is a per-model constant, so r is computed
// once (k_recip) and a division costs five instructions.  The square root
// likewise keeps only the hardware estimate and its two one-ulp corrections
// (no denormal rescaling, no zero/inf class fix-up).  The host rejects models
// whose slowness x spacing leaves the range where these identities hold, and
// the GPU parity tests check every node against the CPU oracle's IEEE
// operations bit for bit.
__device__ __forceinline__ float rdiv(float a, float b, float r) {
  const float q0 = a * r;
  const float e1 = __builtin_fmaf(-b, q0, a);
  const float q1 = __builtin_fmaf(e1, r, q0);
  const float e2 = __builtin_fmaf(-b, q1, a);
  return __builtin_fmaf(e2, r, q1);
}

// Division by a known divisor b in two instructions (Brisebarre, Muller and
// Raina, "Accelerating correctly rounded floating-point division when the
// divisor is known in advance", IEEE Trans. Computers 53(8), 2004):
//   q = RN(a * zh + RN(a * zl)),  zh = RN(1/b),  zl = RN(1/b - zh).
// It returns RN(a/b) for every a of most divisors but not of all, so the
// kernels use it only for a geometry whose four divisors passed an exhaustive
// check against IEEE division over every dividend of magnitude in
// [2^-100, 2^100) (k_fastdiv_check); otherwise they use rdiv.
__device__ __forceinline__ float fdiv(float a, float zh, float zl) {
  return __builtin_fmaf(a, zh, a * zl);
}

__device__ __forceinline__ float rsqrt_rn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sdn = __int_as_float(__float_as_int(s) - 1);
  const float sup = __int_as_float(__float_as_int(s) + 1);
  const float r1 = __builtin_fmaf(-sdn, s, x);
  const float r2 = __builtin_fmaf(-sup, s, x);
  const float t = (r1 <= 0.0f) ? sdn : s;
  return (r2 > 0.0f) ? sup : t;
}

// IEEE minNum of two non-NaN values without the sNaN-quieting canonicalise
// the compiler would add for loaded operands (no operand here is a NaN).
__device__ __forceinline__ float vmin(float a, float b) {
  float r;
  asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Minimum of three non-NaN values (one instruction).  With no NaN operand
// the minimum is exact and independent of grouping, so it equals any nesting
// of vmin over the same three values.
__device__ __forceinline__ float vmin3(float a, float b, float c) {
  float r;
  asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// wave_shr:1 (DPP control 0x138): lane k receives lane k-1's `src`; lane 0
// keeps `old`.
__device__ __forceinline__ float dpp_shr1(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), 0x138,
                                                    0xf, 0xf, false));
}

// A shot as the kernels see it.  T is the shot's padded traveltime grid
// (nzz x nxx, column stride nzz).  Slowness of padded node/cell (i, j) is
//   sbase[i + clamp(j - sofs, 0, smax) * nzz]
// which covers a window of the padded model (sbase = padded model column
// x0 + 3, sofs = 3, smax = width-1: the window's own replicated x border), a
// shot on the whole model (sbase = the padded model, sofs = 0, smax = nxx-1:
// the model's padding as uploaded) and a per-shot padded grid (the same).
struct ShotDev {
  float* T;
  const float* sbase;
  int sofs, smax;
  int nxx;
  float sx, sz;  // source position relative to the shot's grid origin
  int x0, w;     // model columns of the shot's window (skewed slowness addressing)
  // persistent schedule: tagged exchange words of this shot
  unsigned long long* xh;  // strip r's bottom row at [r * ldh + c]
  unsigned long long* xc;  // strip r's carries at [(2r + block parity) * 128 + 2*lane + {0,1}]
  int ldh;
  // fused sweep pairs: done word of the pair's first-sweep tile (r, bb) at
  // [r * nb + bb]; nb = blocks per strip
  unsigned long long* xd;
  int nb;
};

// Skewed copy of the padded model slowness for one sweep orientation
// (SA, SB).  In sweep-local cell coordinates (kap = cell row along the sweep's
// z direction, padded model column m), cell (kap, m) is stored at
//   base[(sig + SK_OFF) * KP + kap],  sig = m + (kap & 63)       for SB > 0
//                                     sig = m - (kap & 63) + 63  for SB < 0
// so the 64 cells a tile's lanes need at one step share one sig and sit in
// consecutive words.
constexpr int SK_OFF = 160;  // column margin: tile reach (66) + prefetch (8) + slack
struct SkewView {
  const float* base;
  int KP;  // rows per skewed column (>= 64 * strips + 64)
};

// Pointers held in ShotDev are read from memory, so the compiler cannot see
// that they address global memory; these casts give global (not flat) loads
// and stores.
typedef __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ gfloat* gptr(float* p) { return (gfloat*)p; }
__device__ __forceinline__ const gfloat* gptr(const float* p) { return (const gfloat*)p; }

__device__ __forceinline__ float s_at(const ShotDev& sd, int nzz, int i, int j) {
  const int jm = min(max(j - sd.sofs, 0), sd.smax);
  return gptr(sd.sbase)[i + (size_t)jm * nzz];
}

// Local solver of the reference's inner_sweep (src/modeling/modeling.cu:
// 308-357) for one node, given its upwind neighbours in the current sweep
// (tv: along z, te: along x, tev: diagonal) and the three slowness cells it
// reads: sref = the upwind cell S[i1][j1], szo = the other cell along the
// z-edge (S[i1][j1+b]), sxo = the other cell along the x-edge (S[i1+a][j1]).
// The reference's max(j-1,1)/max(i-1,1) clamps select cell column/row 1 in
// place of 0; replicated padding makes those cells equal, so they are not
// repeated here.  Returns min(t1D, t2D) — the caller takes min with T_old.
template <bool EQH>
__device__ __forceinline__ float eik_local(float tv, float te, float tev, float sref, float szo,
                                           float sxo, const EikConsts& k) {
  const float t1d1 = tv + k.dz * vmin(sref, szo);
  const float t1d2 = te + k.dx * vmin(sref, sxo);
  const float t1D = vmin(t1d1, t1d2);
  const float de = te - tev;
  const float dv = tv - tev;
  float t2D = TINF;
  if ((tv <= te + k.dx * sref) && (te <= tv + k.dz * sref) && (de >= 0.0f) && (dv >= 0.0f)) {
    // branch A: plane-wave update from both upwind neighbours
    const float ta = tev + te - tv;
    const float tb = tev - te + tv;
    const float dab = ta - tb;
    t2D = rdiv((tb * k.dz2i + ta * k.dx2i) + rsqrt_rn(4.0f * sref * sref * k.sum - k.prod * dab * dab), k.sum,
               k.r_sum);
  } else {
    // branches B (from te) and C (from tv) are exclusive else-ifs in the
    // reference; one evaluation serves whichever a lane takes.
    const float thrB = rdiv(sref * k.dz * k.dz, k.hyp, k.r_hyp);
    const bool cB = (de <= thrB) && (de > 0.0f);
    bool cC = false;
    if (!cB) {
      const float thrC = EQH ? thrB : rdiv(sref * k.dx * k.dx, k.hyp, k.r_hyp);
      cC = (dv <= thrC) && (dv > 0.0f);
    }
    if (cB || cC) {
      const float dd = cB ? de : dv;
      const float h1 = cB ? k.dz : k.dx;
      const float r1 = cB ? k.r_dz : k.r_dx;
      const float h2 = cB ? k.dx : k.dz;
      const float t0 = cB ? te : tv;
      const float q = rdiv(dd, h1, r1);
      t2D = t0 + h2 * rsqrt_rn(sref * sref - q * q);
    }
  }
  return vmin(t1D, vmin(t2D, TINF));
}

// Quantities of one slowness cell that the local solver derives from it,
// computed once when the cell enters a tile's prefetch queue: the cell is the
// z-other cell of one step and the reference cell of the next.  Each is the
// same rounded expression eik_local evaluates.
struct CellQ {
  float dzs;   // dz * s
  float dxs;   // dx * s
  float s4;    // 4 * s * s * (dz2i + dx2i)
  float thrB;  // s * dz * dz / hyp
  float thrC;  // s * dx * dx / hyp
  float s2;    // s * s
};

template <bool EQH, bool FD>
__device__ __forceinline__ CellQ cell_q(float s, const EikConsts& k) {
  CellQ c;
  c.dzs = s * k.dz;
  c.dxs = EQH ? c.dzs : s * k.dx;  // EQH: dx == dz
  // RN(RN(4s * s) * sum) with 4s exact and RN(4 s^2) = 4 RN(s^2) (s = 0 or
  // s^2 normal), so it equals RN(RN(s^2) * 4sum): one product fewer
  c.s2 = s * s;
  c.s4 = c.s2 * k.sum4;
  c.thrB = FD ? fdiv(c.dzs * k.dz, k.zh_hyp, k.zl_hyp) : rdiv(c.dzs * k.dz, k.hyp, k.r_hyp);
  c.thrC = EQH ? c.thrB : (FD ? fdiv(c.dxs * k.dx, k.zh_hyp, k.zl_hyp) : rdiv(c.dxs * k.dx, k.hyp, k.r_hyp));
  return c;
}

// eik_local without branches, for the tile kernels: both candidate branches
// are evaluated on every lane and selected, sharing one square root (branch A
// and branches B/C each take exactly one, of different radicands).  With
// h > 0, RN(h * min(a, b)) = min(RN(h * a), RN(h * b)), so the 1-D candidates
// can use the cells' precomputed products.  Bit-identical to eik_local: the
// selected lanes evaluate the same operations in the same order; the others
// compute values that the selects discard.  FD: divisions by fdiv.  Returns
// min(told, eik_local): the minimum of told, both 1-D candidates, the 2-D
// candidate and TINF (no operand is a NaN, so two three-way minima give the
// same value as the nested pairwise ones).
template <bool EQH, bool FD>
__device__ __forceinline__ float eik_local_sel(float tv, float te, float tev, const CellQ& r, float dzs_zo,
                                               float dxs_xo, const EikConsts& k, float told) {
  const float t1d1 = tv + vmin(r.dzs, dzs_zo);
  const float t1d2 = te + vmin(r.dxs, dxs_xo);
  const float t1D = vmin3(t1d1, t1d2, TINF);
  const float de = te - tev;
  const float dv = tv - tev;
  // non-short-circuit conjunctions: plain mask arithmetic, no control flow
  // Branch A's four conditions as one test: for finite floats a, b,
  // RN(a - b) >= 0 exactly when a >= b (a difference of floats rounds to
  // zero only when it is zero), and a minimum is >= 0 exactly when every
  // operand is.  B and C: with thr > 0 finite, 0 < d <= thr exactly when
  // bits(d) - 1 < bits(thr) as unsigned words (positive floats order as their
  // bit patterns; +0, -0 and negatives wrap to at least 0x7fffffff).
  const bool cA = vmin3(te + r.dxs - tv, tv + r.dzs - te, vmin(de, dv)) >= 0.0f;
  const bool cB = __float_as_uint(de) - 1u < __float_as_uint(r.thrB);
  const bool cC = __float_as_uint(dv) - 1u < __float_as_uint(r.thrC);
  // branch B precedes C (the reference's else-if order)
  const float dd = cB ? de : dv;
  const float h1 = EQH ? k.dz : (cB ? k.dz : k.dx);
  const float r1 = EQH ? k.r_dz : (cB ? k.r_dz : k.r_dx);
  const float zh1 = EQH ? k.zh_dz : (cB ? k.zh_dz : k.zh_dx);
  const float zl1 = EQH ? k.zl_dz : (cB ? k.zl_dz : k.zl_dx);
  const float h2 = EQH ? k.dz : (cB ? k.dx : k.dz);
  const float t0 = cB ? te : tv;
  const float q = FD ? fdiv(dd, zh1, zl1) : rdiv(dd, h1, r1);
  const float ta = tev + te - tv;
  const float tb = tev - te + tv;
  const float dab = ta - tb;
  // radicand a - b * c: A: 4 s^2 (dz2i + dx2i) - (dz2i dx2i dab) dab; B/C: s^2 - q q
  const float ra = cA ? r.s4 : r.s2;
  const float rb = cA ? k.prod * dab : q;
  const float rc = cA ? dab : q;
  const float sq = rsqrt_rn(ra - rb * rc);
  const float num = (tb * k.dz2i + ta * k.dx2i) + sq;
  const float tA = FD ? fdiv(num, k.zh_sum, k.zl_sum) : rdiv(num, k.sum, k.r_sum);
  const float tBC = t0 + h2 * sq;
  const float t2D = cA ? tA : ((cB | cC) ? tBC : TINF);
  return vmin3(t1D, t2D, told);
}

}  // namespace seisfat
