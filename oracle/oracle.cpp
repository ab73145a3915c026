// ORACLE — test infrastructure only.
//
// A CPU restatement of SeisFAT2D's eikonal path, used as the checker for the
// HIP implementation in seisfat2d_amd/.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load this library; the product path never
// calls it.
//
// Every routine is templated on the working real type R (float = the
// reference's FP32 arithmetic, double = the FP64 oracle).  The FP32 build uses
// -ffp-contract=off so every a*b+c is two correctly rounded operations, exactly
// as the C++ expressions in the reference read; fminf semantics (NaN-ignoring)
// stand in for CUDA's device min(float,float).
//
// Parity pins (what checks each function; see tests/test_oracle.py):
//   expand/reduce boundary ...... hand-checked 4x5 fixture, round trip
//   init ......................... closed form S*dist on the 3x3 stencil
//   fsm_pass (raster) ............ literal diagonal schedule (fsm_pass_diag)
//                                  bit-exact; plane-wave exactness (branch A);
//                                  exact one-edge plane waves that only
//                                  branch B / branch C can return, with
//                                  dx != dz; 1-D edge running sums (cell pair
//                                  and clamps, all four sweeps); t=r/v for
//                                  constant velocity, also far-field on grids
//                                  with dz:dx far from 1 (branch A's guard);
//                                  head-wave formula for a two-layer model;
//                                  FP32 vs FP64; monotone under extra passes;
//                                  x-mirror symmetry
//   compress/decode .............. closed-form fixture, exact min/max,
//                                  constant component
//   (tests/test_oracle_mutants.py plants 27 plausible mistakes in copies of
//   this file and checks that each one fails at least one pin)
//   quasi_slowness ............... isotropic stiffness -> S unchanged; VTI
//                                  qP phase velocity (Thomsen exact formula)
//   picks (Catmull-Rom) .......... on-node receivers exact; quadratic fields
//                                  reproduced exactly
//   ray_paths .................... constant-velocity straight rays: row sums
//                                  equal the source-receiver distance
//   adjoint ...................... FP64 central finite differences of the
//                                  FP64 single-pass forward
//   batch_misfit ................. FP64 central finite differences of its own
//                                  J over window shots (pins the window fold)
//   ricker ....................... closed form of the Ricker wavelet at its
//                                  zeros and peak; the reference's sample count
//   trace_convolve ............... numpy FFT product of the reference's
//                                  method (unnormalised, nfft, shift)
//   kdm_trace .................... adjointness <L m, d> = <m, L^T d> in both
//                                  domains; isochron support of a spike
//                                  trace; weights against an FP64 evaluation
//                                  of the published factors; angle bins
//                                  against the angle between the gradients
//
// Reference files restated (paths relative to the SeisFAT2D repository):
//   src/modeling/modeling.cu, src/modeling/eikonal_iso.cu,
//   src/modeling/eikonal_ani.cu, src/admin/admin.cpp,
//   src/inversion/inversion.cpp.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int NB = 3;          // modeling.cu:113
constexpr int COMPRESS = 65535; // modeling.cuh:13

template <class R> inline R rmin(R a, R b) { return std::fmin(a, b); }
template <class R> inline R rsqrt_(R x) { return std::sqrt(x); }

// ---------------------------------------------------------------------------
// A1: Modeling::expand_boundary  (src/modeling/modeling.cu:204-232)
// Interior copy, then edge rows (z) replicated for interior columns, then edge
// columns (x) replicated for every row.  Storage is z-fastest: idx = i + j*n.
template <class R>
void expand_boundary(const R* in, R* out, int nz, int nx) {
  const int nzz = nz + 2 * NB, nxx = nx + 2 * NB;
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < nz; ++i) out[(i + NB) + (j + NB) * nzz] = in[i + j * nz];
  for (int i = 0; i < NB; ++i)
    for (int j = NB; j < nxx - NB; ++j) {
      out[i + j * nzz] = out[NB + j * nzz];
      out[(nzz - i - 1) + j * nzz] = out[(nzz - NB - 1) + j * nzz];
    }
  for (int i = 0; i < nzz; ++i)
    for (int j = 0; j < NB; ++j) {
      out[i + j * nzz] = out[i + NB * nzz];
      out[i + (nxx - j - 1) * nzz] = out[i + (nxx - NB - 1) * nzz];
    }
}

// Modeling::reduce_boundary  (src/modeling/modeling.cu:234-244)
template <class R>
void reduce_boundary(const R* in, R* out, int nz, int nx) {
  const int nzz = nz + 2 * NB;
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < nz; ++i) out[i + j * nz] = in[(i + NB) + (j + NB) * nzz];
}

// ---------------------------------------------------------------------------
// A3: Modeling::initialization + kernels time_set / time_init
// (src/modeling/modeling.cu:110-121, 286-306).  The source snaps to the nearest
// node; the 3x3 nodes around it get S[node] * |node - (sx,sz)| using the node's
// own slowness and the exact source position.  powf(d, 2.0f) is evaluated as
// d*d (the correctly rounded square).
template <class R>
void init_time(R* T, const R* S, int nzz, int nxx, R dx, R dz, R sx, R sz,
               int* out_sIdx, int* out_sIdz) {
  const R half = R(0.5);
  const int sIdx = (int)((sx + half * dx) / dx) + NB;
  const int sIdz = (int)((sz + half * dz) / dz) + NB;
  const size_t matsize = (size_t)nzz * nxx;
  for (size_t k = 0; k < matsize; ++k) T[k] = R(1e6);
  for (int jj = 0; jj < 3; ++jj)
    for (int ii = 0; ii < 3; ++ii) {
      const int zi = sIdz + (ii - 1);
      const int xi = sIdx + (jj - 1);
      const size_t index = (size_t)zi + (size_t)xi * nzz;
      const R ddx = (R)(xi - NB) * dx - sx;
      const R ddz = (R)(zi - NB) * dz - sz;
      T[index] = S[index] * rsqrt_(ddx * ddx + ddz * ddz);
    }
  if (out_sIdx) *out_sIdx = sIdx;
  if (out_sIdz) *out_sIdz = sIdz;
}

// ---------------------------------------------------------------------------
// A5: the local solver of kernel inner_sweep (src/modeling/modeling.cu:308-357)
// for node (i,j) with sweep signs (a,b) = sgnt; sgnv = (sgnt+1)/2
// (src/modeling/modeling.cu:62-102).
template <class R>
struct Consts {
  R dx, dz, dx2i, dz2i;
};

template <class R>
inline R local_candidate(const R* T, const R* S, int i, int j, int nzz, int nxx,
                         int a, int b, const Consts<R>& c) {
  const int i1 = i - (a + 1) / 2;
  const int j1 = j - (b + 1) / 2;
  const R tv = T[(i - a) + (size_t)j * nzz];
  const R te = T[i + (size_t)(j - b) * nzz];
  const R tev = T[(i - a) + (size_t)(j - b) * nzz];
  const R t1d1 = tv + c.dz * rmin(S[i1 + (size_t)std::max(j - 1, 1) * nzz],
                                  S[i1 + (size_t)std::min(j, nxx - 1) * nzz]);
  const R t1d2 = te + c.dx * rmin(S[std::max(i - 1, 1) + (size_t)j1 * nzz],
                                  S[std::min(i, nzz - 1) + (size_t)j1 * nzz]);
  const R t1D = rmin(t1d1, t1d2);
  R t1 = R(1e6), t2 = R(1e6), t3 = R(1e6);
  const R Sref = S[i1 + (size_t)j1 * nzz];
  if ((tv <= te + c.dx * Sref) && (te <= tv + c.dz * Sref) && (te - tev >= R(0)) &&
      (tv - tev >= R(0))) {
    const R ta = tev + te - tv;
    const R tb = tev - te + tv;
    t1 = ((tb * c.dz2i + ta * c.dx2i) +
          rsqrt_(R(4) * Sref * Sref * (c.dz2i + c.dx2i) -
                 c.dz2i * c.dx2i * (ta - tb) * (ta - tb))) /
         (c.dz2i + c.dx2i);
  } else if ((te - tev <= Sref * c.dz * c.dz / rsqrt_(c.dx * c.dx + c.dz * c.dz)) &&
             (te - tev > R(0))) {
    t2 = te + c.dx * rsqrt_(Sref * Sref - ((te - tev) / c.dz) * ((te - tev) / c.dz));
  } else if ((tv - tev <= Sref * c.dx * c.dx / rsqrt_(c.dx * c.dx + c.dz * c.dz)) &&
             (tv - tev > R(0))) {
    t3 = tv + c.dz * rsqrt_(Sref * Sref - ((tv - tev) / c.dx) * ((tv - tev) / c.dx));
  }
  const R t2D = rmin(t1, rmin(t2, t3));
  return rmin(t1D, t2D);
}

inline void sweep_signs(int sweep, int* a, int* b) {
  // sgnt = {{1,1},{-1,1},{1,-1},{-1,-1}}  (src/modeling/modeling.cu:70)
  *a = (sweep == 0 || sweep == 2) ? 1 : -1;
  *b = (sweep == 0 || sweep == 1) ? 1 : -1;
}

template <class R>
Consts<R> make_consts(R dx, R dz) {
  Consts<R> c;
  c.dx = dx;
  c.dz = dz;
  c.dz2i = R(1) / (dz * dz);  // modeling.cu:64-65
  c.dx2i = R(1) / (dx * dx);
  return c;
}

// A4, raster form: one directional sweep visiting nodes in any order that
// respects the sweep's dependencies (each node reads only its three upwind
// neighbours, which lie on the two previous anti-diagonals).
template <class R>
void sweep_raster(R* T, const R* S, int nzz, int nxx, int sweep, const Consts<R>& c) {
  int a, b;
  sweep_signs(sweep, &a, &b);
  for (int jj = 1; jj < nxx - 1; ++jj) {
    const int j = (b > 0) ? jj : nxx - 1 - jj;
    for (int ii = 1; ii < nzz - 1; ++ii) {
      const int i = (a > 0) ? ii : nzz - 1 - ii;
      const size_t k = i + (size_t)j * nzz;
      T[k] = rmin(T[k], local_candidate(T, S, i, j, nzz, nxx, a, b, c));
    }
  }
}

// A4, literal form: the per-level launch schedule of Modeling::eikonal_solver
// (src/modeling/modeling.cu:123-159), element by element.
template <class R>
void sweep_diag(R* T, const R* S, int nzz, int nxx, int sweep, const Consts<R>& c) {
  int a, b;
  sweep_signs(sweep, &a, &b);
  const int total_levels = (nxx - 1) + (nzz - 1);
  const int min_level = std::min(nxx, nzz), max_level = std::max(nxx, nzz);
  const int zd = (sweep == 2 || sweep == 3) ? -1 : 1;
  const int xd = (sweep == 0 || sweep == 2) ? -1 : 1;
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
    const int nthreads = ((n_elements + 255) / 256) * 256;  // whole blocks launched
    // Threads of one level are independent: compute every candidate from the
    // state before the level, then write (what a launch guarantees).
    std::vector<std::pair<size_t, R>> upd;
    for (int e = 0; e < nthreads; ++e) {
      const int i = z_offset + zd * e;
      const int j = x_offset + xd * e;
      if ((i > 0) && (i < nzz - 1) && (j > 0) && (j < nxx - 1)) {
        const size_t k = i + (size_t)j * nzz;
        upd.emplace_back(k, rmin(T[k], local_candidate(T, S, i, j, nzz, nxx, a, b, c)));
      }
    }
    for (auto& u : upd) T[u.first] = u.second;
  }
}

// A4/A6: one pass of the four sweeps (ref mode), or passes repeated until no
// value changes (conv mode).  Returns the number of passes executed.
template <class R>
int fsm(R* T, const R* S, int nzz, int nxx, R dx, R dz, int mode, int max_passes,
        int schedule) {
  const Consts<R> c = make_consts(dx, dz);
  const size_t matsize = (size_t)nzz * nxx;
  std::vector<R> prev;
  int passes = 0;
  for (;;) {
    if (mode == 1) prev.assign(T, T + matsize);
    for (int s = 0; s < 4; ++s) {
      if (schedule == 1)
        sweep_diag(T, S, nzz, nxx, s, c);
      else
        sweep_raster(T, S, nzz, nxx, s, c);
    }
    ++passes;
    if (mode == 0 || passes >= max_passes) break;
    if (std::memcmp(prev.data(), T, matsize * sizeof(R)) == 0) break;
  }
  return passes;
}

// ---------------------------------------------------------------------------
// A7: Modeling::compression (src/modeling/modeling.cu:264-279) with exact
// (serial) min/max; the reference's OpenMP min/max race is not reproduced.
void compress(const float* in, uint16_t* out, size_t n, float* maxv, float* minv) {
  float mx = -1e20f, mn = 1e20f;
  for (size_t k = 0; k < n; ++k) {
    mn = std::min(in[k], mn);
    mx = std::max(in[k], mx);
  }
  // a constant input would make the reference's expression 0/0 (a NaN cast
  // to uint16_t); code 1 decodes to mn exactly
  for (size_t k = 0; k < n; ++k)
    out[k] = (mx == mn) ? uint16_t(1) : static_cast<uint16_t>(1.0f + (COMPRESS - 1) * (in[k] - mn) / (mx - mn));
  *maxv = mx;
  *minv = mn;
}

// Decode of one stiffness element (src/modeling/eikonal_ani.cu:76-83).
inline float decode(uint16_t u, float mn, float mx) {
  return (mn + (static_cast<float>(u) - 1.0f) * (mx - mn) / (COMPRESS - 1));
}

// A8: kernel get_quasi_slowness (src/modeling/eikonal_ani.cu:47-111), FP32 as
// in the reference including the double-precision eigenvalue sum at :103-104.
void quasi_slowness(const float* T, float* S, int nzz, int nxx, float dx, float dz, int sIdx,
                    int sIdz, const uint16_t* const* C, const float* mn, const float* mx) {
  for (int j = NB; j < nxx - NB; ++j)
    for (int i = NB; i < nzz - NB; ++i) {
      if ((i == sIdz) && (j == sIdx)) continue;
      const size_t index = i + (size_t)j * nzz;
      const float dTz = 0.5f * (T[(i + 1) + (size_t)j * nzz] - T[(i - 1) + (size_t)j * nzz]) / dz;
      const float dTx = 0.5f * (T[i + (size_t)(j + 1) * nzz] - T[i + (size_t)(j - 1) * nzz]) / dx;
      const float norm = std::sqrt(dTx * dTx + dTz * dTz);
      const float p0 = dTx / norm, p1 = dTz / norm;
      const float c11 = decode(C[0][index], mn[0], mx[0]);
      const float c13 = decode(C[1][index], mn[1], mx[1]);
      const float c15 = decode(C[2][index], mn[2], mx[2]);
      const float c33 = decode(C[3][index], mn[3], mx[3]);
      const float c35 = decode(C[4][index], mn[4], mx[4]);
      const float c55 = decode(C[5][index], mn[5], mx[5]);
      float Cm[9] = {c11, c13, c15, c13, c33, c35, c15, c35, c55};  // C[r + c*3]
      const float Ro = c33 * S[index] * S[index];
      for (int k = 0; k < 9; ++k) Cm[k] = Cm[k] / Ro / Ro;
      const float Gxx = Cm[0] * p0 * p0 + Cm[8] * p1 * p1 + 2.0f * Cm[6] * p0 * p1;
      const float Gzz = Cm[8] * p0 * p0 + Cm[4] * p1 * p1 + 2.0f * Cm[7] * p0 * p1;
      const float Gxz = Cm[6] * p0 * p0 + Cm[7] * p1 * p1 + (Cm[3] + Cm[8]) * p0 * p1;
      const float coeff1 = Gxx + Gzz;
      const float coeff2 = Gxx - Gzz;
      const float det = std::sqrt((coeff2 * coeff2) / 4.0f + Gxz * Gxz);
      float g0 = (float)((double)coeff1 / 2.0 + (double)det);
      float g1 = (float)((double)coeff1 / 2.0 - (double)det);
      if (g0 < g1) std::swap(g0, g1);
      S[index] = 1.0f / std::sqrt(g0 * Ro);
    }
}

// ---------------------------------------------------------------------------
// A10: cubic1d / cubic2d (src/admin/admin.cpp:121-134) and the receiver stencil
// of Modeling::compute_seismogram (src/modeling/modeling.cu:161-194).
template <class R>
inline R cubic1d(const R P[4], R d) {
  return P[1] + R(0.5) * d *
                    (P[2] - P[0] +
                     d * (R(2) * P[0] - R(5) * P[1] + R(4) * P[2] - P[3] +
                          d * (R(3) * (P[1] - P[2]) + P[3] - P[0])));
}

template <class R>
R pick_one(const R* T, int nzz, R dx, R dz, R x, R z) {
  const R half = R(0.5);
  const R x0 = std::floor((x + half * dx) / dx) * dx;
  const R z0 = std::floor((z + half * dz) / dz) * dz;
  const R x1 = std::floor((x + half * dx) / dx) * dx + dx;
  const R z1 = std::floor((z + half * dz) / dz) * dz + dz;
  const R xd = (x - x0) / (x1 - x0);
  const R zd = (z - z0) / (z1 - z0);
  const int i = (int)((z + half * dz) / dz) + NB;
  const int j = (int)((x + half * dx) / dx) + NB;
  R P[4][4];
  for (int pIdx = 0; pIdx < 4; ++pIdx)
    for (int pIdz = 0; pIdz < 4; ++pIdz)
      P[pIdx][pIdz] = T[(i + pIdz - 1) + (size_t)(j + pIdx - 1) * nzz];
  R p[4];
  for (int k = 0; k < 4; ++k) p[k] = cubic1d(P[k], zd);
  return cubic1d(p, xd);
}

// Catmull-Rom weights of cubic1d (the pick is linear in the 16 stencil values).
inline void cubic_weights(double d, double w[4]) {
  w[0] = 0.5 * d * (-1.0 + d * (2.0 - d));
  w[1] = 1.0 + 0.5 * d * d * (-5.0 + 3.0 * d);
  w[2] = 0.5 * d * (1.0 + d * (4.0 - 3.0 * d));
  w[3] = 0.5 * d * d * (-1.0 + d);
}

// ---------------------------------------------------------------------------
// Discrete adjoint of one single-pass solve (init + 4 sweeps in raster order).
// J = 1/2 sum_r (pick_r - obs_r)^2; returns J and accumulates dJ/dS (padded
// grid, double) into g.  Tie rules: T_old wins when T_old <= candidate; a min
// of two exactly equal slownesses splits dS half and half between the two
// cells (the derivative central differences converge to: such ties are
// systematic wherever the model is constant along an axis).
template <class R>
struct NodeDeriv {
  int kind;  // 0 old, 1 d1z, 2 d1x, 3 branch A, 4 branch B, 5 branch C, 6 constant
  double dtv, dte, dtev;  // d(new)/d(neighbour)
  int scell;              // slowness cell index receiving dS
  int scell2;             // second cell of an exact slowness tie, or -1
  double ds;              // d(new)/dS[scell] (halved between scell and scell2)
};

template <class R>
NodeDeriv<R> node_deriv(const R* T, const R* S, int i, int j, int nzz, int nxx, int a, int b,
                        const Consts<R>& c, R told) {
  NodeDeriv<R> d{};
  d.scell2 = -1;
  const int i1 = i - (a + 1) / 2;
  const int j1 = j - (b + 1) / 2;
  const R tv = T[(i - a) + (size_t)j * nzz];
  const R te = T[i + (size_t)(j - b) * nzz];
  const R tev = T[(i - a) + (size_t)(j - b) * nzz];
  const size_t cz_a = i1 + (size_t)std::max(j - 1, 1) * nzz, cz_b = i1 + (size_t)std::min(j, nxx - 1) * nzz;
  const size_t cx_a = std::max(i - 1, 1) + (size_t)j1 * nzz, cx_b = std::min(i, nzz - 1) + (size_t)j1 * nzz;
  const R t1d1 = tv + c.dz * rmin(S[cz_a], S[cz_b]);
  const R t1d2 = te + c.dx * rmin(S[cx_a], S[cx_b]);
  const R t1D = rmin(t1d1, t1d2);
  R t1 = R(1e6), t2 = R(1e6), t3 = R(1e6);
  int br = 0;
  const size_t cref = i1 + (size_t)j1 * nzz;
  const R Sref = S[cref];
  double dd_tv = 0, dd_te = 0, dd_tev = 0, dd_s = 0;
  if ((tv <= te + c.dx * Sref) && (te <= tv + c.dz * Sref) && (te - tev >= R(0)) &&
      (tv - tev >= R(0))) {
    const R ta = tev + te - tv;
    const R tb = tev - te + tv;
    const R rad = R(4) * Sref * Sref * (c.dz2i + c.dx2i) - c.dz2i * c.dx2i * (ta - tb) * (ta - tb);
    t1 = ((tb * c.dz2i + ta * c.dx2i) + rsqrt_(rad)) / (c.dz2i + c.dx2i);
    br = 3;
    const double D = (double)c.dz2i + (double)c.dx2i;
    const double sq = std::sqrt((double)rad);
    const double k = 2.0 * (double)c.dz2i * (double)c.dx2i * ((double)ta - (double)tb) / sq;
    dd_te = ((double)c.dx2i - (double)c.dz2i - k) / D;
    dd_tv = ((double)c.dz2i - (double)c.dx2i + k) / D;
    dd_tev = 1.0;
    dd_s = 4.0 * (double)Sref / sq;
  } else if ((te - tev <= Sref * c.dz * c.dz / rsqrt_(c.dx * c.dx + c.dz * c.dz)) && (te - tev > R(0))) {
    const R q = (te - tev) / c.dz;
    const R Q = rsqrt_(Sref * Sref - q * q);
    t2 = te + c.dx * Q;
    br = 4;
    const double r = (double)c.dx * (double)q / ((double)c.dz * (double)Q);
    dd_te = 1.0 - r;
    dd_tev = r;
    dd_s = (double)c.dx * (double)Sref / (double)Q;
  } else if ((tv - tev <= Sref * c.dx * c.dx / rsqrt_(c.dx * c.dx + c.dz * c.dz)) && (tv - tev > R(0))) {
    const R p = (tv - tev) / c.dx;
    const R P = rsqrt_(Sref * Sref - p * p);
    t3 = tv + c.dz * P;
    br = 5;
    const double r = (double)c.dz * (double)p / ((double)c.dx * (double)P);
    dd_tv = 1.0 - r;
    dd_tev = r;
    dd_s = (double)c.dz * (double)Sref / (double)P;
  }
  const R t2D = rmin(t1, rmin(t2, t3));
  const R cand = rmin(t1D, t2D);
  if (!(cand < told)) {
    d.kind = 0;
    return d;
  }
  if (t1D <= t2D) {
    if (t1d1 <= t1d2) {
      d.kind = 1;
      d.dtv = 1.0;
      d.scell = (int)((S[cz_a] <= S[cz_b]) ? cz_a : cz_b);
      if (S[cz_a] == S[cz_b]) d.scell2 = (int)cz_b;
      d.ds = (double)c.dz;
    } else {
      d.kind = 2;
      d.dte = 1.0;
      d.scell = (int)((S[cx_a] <= S[cx_b]) ? cx_a : cx_b);
      if (S[cx_a] == S[cx_b]) d.scell2 = (int)cx_b;
      d.ds = (double)c.dx;
    }
    return d;
  }
  // t2D = min(t1, t2, t3) with the untaken ones at 1e6: a taken branch whose
  // value is not below 1e6 loses to the constant.
  const R tbr = (br == 3) ? t1 : (br == 4) ? t2 : (br == 5) ? t3 : R(1e6);
  if (br == 0 || !(tbr < R(1e6))) {
    d.kind = 6;
    return d;
  }
  d.kind = br;
  d.dtv = dd_tv;
  d.dte = dd_te;
  d.dtev = dd_tev;
  d.scell = (int)cref;
  d.ds = dd_s;
  return d;
}

// Which candidate decided each node update of one single-pass solve (init +
// four raster sweeps): counts[kind] over NodeDeriv's kinds (0 old value kept,
// 1 1-D along z, 2 1-D along x, 3 branch A, 4 branch B, 5 branch C, 6 the
// 1e6 constant).  Lets tests show that an input exercises branches B and C.
template <class R>
void sweep_counted(R* T, const R* S, int nzz, int nxx, int sweep, const Consts<R>& c, long long counts[7]) {
  int a, b;
  sweep_signs(sweep, &a, &b);
  for (int jj = 1; jj < nxx - 1; ++jj) {
    const int j = (b > 0) ? jj : nxx - 1 - jj;
    for (int ii = 1; ii < nzz - 1; ++ii) {
      const int i = (a > 0) ? ii : nzz - 1 - ii;
      const size_t k = i + (size_t)j * nzz;
      const NodeDeriv<R> d = node_deriv(T, S, i, j, nzz, nxx, a, b, c, T[k]);
      ++counts[d.kind];
      T[k] = rmin(T[k], local_candidate(T, S, i, j, nzz, nxx, a, b, c));
    }
  }
}

template <class R>
void branch_counts(const R* S, int nzz, int nxx, R dx, R dz, R sx, R sz, long long counts[7]) {
  const Consts<R> c = make_consts(dx, dz);
  std::vector<R> T((size_t)nzz * nxx);
  init_time(T.data(), S, nzz, nxx, dx, dz, sx, sz, nullptr, nullptr);
  for (int k = 0; k < 7; ++k) counts[k] = 0;
  for (int s = 0; s < 4; ++s) sweep_counted(T.data(), S, nzz, nxx, s, c, counts);
}

template <class R>
double adjoint(const R* S, int nzz, int nxx, R dx, R dz, R sx, R sz, const R* rx, const R* rz,
               int nrec, const R* obs, double* g, R* picks_out) {
  const Consts<R> c = make_consts(dx, dz);
  const size_t matsize = (size_t)nzz * nxx;
  std::vector<R> snap(5 * matsize);  // T after init, after sweep 0..3
  int sIdx, sIdz;
  init_time(snap.data(), S, nzz, nxx, dx, dz, sx, sz, &sIdx, &sIdz);
  for (int s = 0; s < 4; ++s) {
    std::copy(snap.begin() + s * matsize, snap.begin() + (s + 1) * matsize,
              snap.begin() + (s + 1) * matsize);
    sweep_raster(snap.data() + (s + 1) * matsize, S, nzz, nxx, s, c);
  }
  const R* Tf = snap.data() + 4 * matsize;
  std::vector<double> L(matsize, 0.0);
  double J = 0.0;
  for (int r = 0; r < nrec; ++r) {
    const R pk = pick_one(Tf, nzz, dx, dz, rx[r], rz[r]);
    if (picks_out) picks_out[r] = pk;
    const double res = (double)pk - (double)obs[r];
    J += 0.5 * res * res;
    // Same stencil position and fractional offsets as pick_one.
    const R half = R(0.5);
    const R x0 = std::floor((rx[r] + half * dx) / dx) * dx;
    const R z0 = std::floor((rz[r] + half * dz) / dz) * dz;
    const R x1 = std::floor((rx[r] + half * dx) / dx) * dx + dx;
    const R z1 = std::floor((rz[r] + half * dz) / dz) * dz + dz;
    const double xd = (double)((rx[r] - x0) / (x1 - x0));
    const double zd = (double)((rz[r] - z0) / (z1 - z0));
    const int i = (int)((rz[r] + half * dz) / dz) + NB;
    const int j = (int)((rx[r] + half * dx) / dx) + NB;
    double wx[4], wz[4];
    cubic_weights(xd, wx);
    cubic_weights(zd, wz);
    for (int pIdx = 0; pIdx < 4; ++pIdx)
      for (int pIdz = 0; pIdz < 4; ++pIdz)
        L[(i + pIdz - 1) + (size_t)(j + pIdx - 1) * nzz] += res * wx[pIdx] * wz[pIdz];
  }
  for (int s = 3; s >= 0; --s) {
    int a, b;
    sweep_signs(s, &a, &b);
    const R* Tnew = snap.data() + (s + 1) * matsize;  // state after sweep s
    const R* Told = snap.data() + s * matsize;        // state before sweep s
    for (int jj = nxx - 2; jj >= 1; --jj) {
      const int j = (b > 0) ? jj : nxx - 1 - jj;
      for (int ii = nzz - 2; ii >= 1; --ii) {
        const int i = (a > 0) ? ii : nzz - 1 - ii;
        const size_t k = i + (size_t)j * nzz;
        const double lam = L[k];
        if (lam == 0.0) continue;
        // Neighbour values seen by the forward update are the sweep-s values.
        const NodeDeriv<R> d = node_deriv(Tnew, S, i, j, nzz, nxx, a, b, c, Told[k]);
        if (d.kind == 0) continue;  // adjoint passes to T_old unchanged
        L[k] = 0.0;
        if (d.kind == 6) continue;
        if (d.dtv != 0.0) L[(i - a) + (size_t)j * nzz] += lam * d.dtv;
        if (d.dte != 0.0) L[i + (size_t)(j - b) * nzz] += lam * d.dte;
        if (d.dtev != 0.0) L[(i - a) + (size_t)(j - b) * nzz] += lam * d.dtev;
        if (d.scell2 >= 0) {
          g[d.scell] += lam * d.ds * 0.5;
          g[d.scell2] += lam * d.ds * 0.5;
        } else {
          g[d.scell] += lam * d.ds;
        }
      }
    }
  }
  // Source initialisation: T = S[node] * dist on the 3x3 stencil.
  for (int jj = 0; jj < 3; ++jj)
    for (int ii = 0; ii < 3; ++ii) {
      const int zi = sIdz + (ii - 1), xi = sIdx + (jj - 1);
      const size_t index = (size_t)zi + (size_t)xi * nzz;
      const R ddx = (R)(xi - NB) * dx - sx;
      const R ddz = (R)(zi - NB) * dz - sz;
      g[index] += L[index] * (double)rsqrt_(ddx * ddx + ddz * ddz);
    }
  return J;
}

// Batch of window shots with the batched GPU path's semantics: shot k solves
// columns [x0[k], x0[k]+w[k]) of the unpadded model s (nz x nx), padded as a
// standalone model, and picks receivers rec_off[k] .. rec_off[k]+nrec[k]-1
// (absolute positions).  Returns J = 1/2 sum (pick - obs)^2 over all
// receivers; when grad is non-null it also writes dJ/ds (nz*nx, unpadded):
// each shot's padded gradient folded onto the model cells its padding
// replicates.  Shots run in parallel; sums are taken in shot order.
template <class R>
double batch_misfit(const R* s, int nz, int nx, R dx, R dz, int nshots, const R* sx, const R* sz, const int* x0,
                    const int* w, const int* rec_off, const int* nrec, const R* rx, const R* rz, const R* obs,
                    double* grad, R* picks, int nthreads) {
  std::vector<double> Jk(nshots, 0.0);
  std::vector<std::vector<double>> gk(grad ? nshots : 0);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads != 1)
  for (int k = 0; k < nshots; ++k) {
    const int nzz = nz + 2 * NB, nxx = w[k] + 2 * NB;
    std::vector<R> crop((size_t)nz * w[k]), S((size_t)nzz * nxx);
    std::memcpy(crop.data(), s + (size_t)x0[k] * nz, crop.size() * sizeof(R));
    expand_boundary(crop.data(), S.data(), nz, w[k]);
    const R shift = (R)x0[k] * dx;
    const int n = nrec[k], ro = rec_off[k];
    std::vector<R> rxk(n), rzk(n), obk(n), pk(n);
    for (int r = 0; r < n; ++r) {
      rxk[r] = rx[ro + r] - shift;
      rzk[r] = rz[ro + r];
      obk[r] = obs[ro + r];
    }
    const R sxk = sx[k] - shift;
    if (grad) {
      std::vector<double> g((size_t)nzz * nxx, 0.0);
      Jk[k] = adjoint(S.data(), nzz, nxx, dx, dz, sxk, sz[k], rxk.data(), rzk.data(), n, obk.data(), g.data(),
                      pk.data());
      std::vector<double>& out = gk[k];
      out.assign((size_t)nz * w[k], 0.0);
      for (int j = 0; j < nxx; ++j) {
        const int jm = std::min(std::max(j - NB, 0), w[k] - 1);
        for (int i = 0; i < nzz; ++i) {
          const int im = std::min(std::max(i - NB, 0), nz - 1);
          out[im + (size_t)jm * nz] += g[i + (size_t)j * nzz];
        }
      }
    } else {
      std::vector<R> T((size_t)nzz * nxx);
      init_time(T.data(), S.data(), nzz, nxx, dx, dz, sxk, sz[k], nullptr, nullptr);
      fsm(T.data(), S.data(), nzz, nxx, dx, dz, 0, 1, 0);
      double J = 0.0;
      for (int r = 0; r < n; ++r) {
        pk[r] = pick_one(T.data(), nzz, dx, dz, rxk[r], rzk[r]);
        const double res = (double)pk[r] - (double)obk[r];
        J += 0.5 * res * res;
      }
      Jk[k] = J;
    }
    if (picks) std::copy(pk.begin(), pk.end(), picks + ro);
  }
  double J = 0.0;
  for (int k = 0; k < nshots; ++k) J += Jk[k];
  if (grad) {
    std::fill(grad, grad + (size_t)nz * nx, 0.0);
    for (int k = 0; k < nshots; ++k)
      for (size_t c = 0; c < gk[k].size(); ++c) grad[(size_t)x0[k] * nz + c] += gk[k][c];
  }
  return J;
}

// ---------------------------------------------------------------------------
// A11: Inversion::gradient_ray_tracing (src/inversion/inversion.cpp:95-177) for
// one shot: steepest descent on T from each receiver, visited cells sorted and
// run-length counted; the first run carries one extra step and the source cell
// gets the final residual distance.  Emits COO triplets (receiver, cell, len).
// max_steps guards against a ray that never reaches the source cell.
int ray_paths(const float* T, int nz, int nx, float dx, float dz, float sx, float sz,
              const float* rx, const float* rz, int nrec, int max_steps, int* out_rec,
              int* out_cell, float* out_len, int cap) {
  const int nzz = nz + 2 * NB;
  const int sIdx = (int)((sx + 0.5f * dx) / dx);
  const int sIdz = (int)((sz + 0.5f * dz) / dz);
  const int sId = sIdz + sIdx * nz;
  const float rayStep = 0.2f * dz;
  int n = 0;
  std::vector<int> ray_index;
  for (int r = 0; r < nrec; ++r) {
    float xi = rx[r], zi = rz[r];
    if ((sz == zi) && (sx == xi)) continue;
    ray_index.clear();
    int steps = 0;
    for (;;) {
      const int j = (int)((xi + 0.5f * dx) / dx) + NB;
      const int i = (int)((zi + 0.5f * dz) / dz) + NB;
      const float dTx = (T[i + (size_t)(j + 1) * nzz] - T[i + (size_t)(j - 1) * nzz]) / (2.0f * dx);
      const float dTz = (T[(i + 1) + (size_t)j * nzz] - T[(i - 1) + (size_t)j * nzz]) / (2.0f * dz);
      const float norm = std::sqrt(dTx * dTx + dTz * dTz);
      xi -= rayStep * dTx / norm;
      zi -= rayStep * dTz / norm;
      const int jm = (int)((xi + 0.5f * dx) / dx);
      const int im = (int)((zi + 0.5f * dz) / dz);
      ray_index.push_back(im + jm * nz);
      if (ray_index.back() == sId) break;
      if (++steps >= max_steps || im < 0 || im >= nz || jm < 0 || jm >= nx) return -1 - r;
    }
    const float dzs = zi - sz, dxs = xi - sx;
    const float final_distance = std::sqrt(dzs * dzs + dxs * dxs);
    std::sort(ray_index.begin(), ray_index.end());
    int current = ray_index[0];
    float dist = rayStep;
    auto emit = [&](int cell, float len) {
      if (n < cap) {
        out_rec[n] = r;
        out_cell[n] = cell;
        out_len[n] = len;
      }
      ++n;
    };
    for (size_t k = 0; k < ray_index.size(); ++k) {
      if (ray_index[k] == current) {
        dist += rayStep;
      } else {
        emit(current, (current == sId) ? final_distance : dist);
        dist = rayStep;
        current = ray_index[k];
      }
    }
    emit(current, (current == sId) ? final_distance : dist);
  }
  return n;
}

// ---------------------------------------------------------------------------
// Tomography operators.
//
// Inversion::set_regularization_matrix (src/inversion/inversion.cpp:209-262):
// order 0 is the identity; order k >= 1 is the k-th forward difference over
// the flattened (z-fastest) model vector, n - k rows, row r holding df[q] at
// column r + q, where df = [-1, 1] convolved with [-1, 1] k - 1 times (built
// by the reference's df1/df2 recurrence).  y = D x.
void tikhonov_coefficients(int order, std::vector<float>& df) {
  df.assign(order + 1, 0.0f);
  if (order <= 0) {
    df[0] = 1.0f;
    return;
  }
  std::vector<float> d1(order + 2, 0.0f), d2(order + 2, 0.0f);
  df[0] = -1.0f;
  df[1] = 1.0f;
  for (int it = 1; it < order; ++it)
    for (int q = 0; q < order + 1; ++q) {
      d2[q] = df[q];
      d1[q + 1] = df[q];
      df[q] = d1[q] - d2[q];
    }
}

void tikhonov_apply(int order, const double* x, long long n, double* y) {
  std::vector<float> df;
  tikhonov_coefficients(order, df);
  const long long rows = n - std::max(order, 0);
  for (long long r = 0; r < rows; ++r) {
    double acc = 0.0;
    for (int q = 0; q < (int)df.size(); ++q) acc += (double)df[q] * x[r + q];
    y[r] = acc;
  }
}

// D^T y (n outputs from n - order inputs).
void tikhonov_adjoint(int order, const double* y, long long n, double* x) {
  std::vector<float> df;
  tikhonov_coefficients(order, df);
  const long long rows = n - std::max(order, 0);
  for (long long c = 0; c < n; ++c) x[c] = 0.0;
  for (long long r = 0; r < rows; ++r)
    for (int q = 0; q < (int)df.size(); ++q) x[r + q] += (double)df[q] * y[r];
}

// Inversion::model_smoothing + smooth_matrix (src/inversion/inversion.cpp:
// 347-447): the model is embedded in a zero border of `samples` cells, then
// convolved with a samples x samples Gaussian (weights exp(-r^2/(2 stdv^2))
// / (pi stdv), normalised to sum 1; a unit delta when stdv = 0) at every node
// at least samples/2 from the embedding's edge (the others keep their
// value), and the model region is read back.  FP32 as the reference.
void smooth_model(const float* in, float* out, int nz, int nx, float stdv, int samples) {
  const int anx = nx + 2 * samples, anz = nz + 2 * samples;
  std::vector<float> aux((size_t)anx * anz, 0.0f), sm((size_t)anx * anz);
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < nz; ++i) aux[(i + samples) + (size_t)(j + samples) * anz] = in[i + (size_t)j * nz];
  const int nk = samples * samples, mid = samples / 2;
  std::vector<float> kernel(nk, 0.0f);
  sm = aux;
  kernel[mid + mid * samples] = 1.0f;
  if (stdv != 0.0f) {
    float sum = 0.0f;
    for (int x = -mid; x <= mid; ++x)
      for (int z = -mid; z <= mid; ++z) {
        const int index = (z + mid) + (x + mid) * samples;
        const float r = std::sqrt((float)(x * x + z * z));
        // the reference's 1.0f / (M_PI * stdv) is a double expression
        kernel[index] = (float)(1.0 / (M_PI * (double)stdv) * (double)std::exp(-((r * r) / (2.0f * stdv * stdv))));
        sum += kernel[index];
      }
    for (int q = 0; q < nk; ++q) kernel[q] /= sum;
  }
  for (int j = mid; j < anx - mid; ++j)
    for (int i = mid; i < anz - mid; ++i) {
      float acc = 0.0f;
      for (int xk = 0; xk < samples; ++xk)
        for (int zk = 0; zk < samples; ++zk)
          acc += aux[(i - mid + zk) + (size_t)(j - mid + xk) * anz] * kernel[zk + xk * samples];
      sm[i + (size_t)j * anz] = acc;
    }
  for (int j = 0; j < nx; ++j)
    for (int i = 0; i < nz; ++i) out[i + (size_t)j * nz] = sm[(i + samples) + (size_t)(j + samples) * anz];
}

// ---------------------------------------------------------------------------
// Kirchhoff depth migration (src/migration/migration.cu): the Ricker wavelet of
// Migration::set_wavelet (:54-73); the wavelet convolutions of
// forward_convolution / adjoint_convolution (:197-261), whose FFT product
// (unnormalised, size nfft) is the linear (cross-)correlation summed here
// directly, times nfft; and one trace of the four Kirchhoff kernels
// (:398-756), every node's weight from 9-point differences of both tables as
// the kernels evaluate it, FP32, nodes in index order.
constexpr float KDM_EPS = 1e-6f;  // migration.cuh:7

void ricker(float dt, float fmax, std::vector<float>& w) {
  const float t0 = 2.0f * std::sqrt((float)M_PI) / fmax;
  const float fc = fmax / (3.0f * std::sqrt((float)M_PI));
  const int nw = 2 * ((int)((t0 - 0.5f * dt) / dt));
  w.assign(std::max(nw, 0), 0.0f);
  for (int k = 0; k < nw; ++k) {
    const float td = k * dt - t0;
    const float arg = (float)(M_PI * M_PI * M_PI * fc * fc * td * td);  // a double expression in the reference
    w[k] = (1.0f - 2.0f * arg) * std::exp(-arg);
  }
}

long long nextpow2_ll(long long n) {
  long long p = 1;
  while (p < n) p <<= 1;
  return p;
}

void trace_convolve(int adjoint, int ntr, int nt, const float* w, int nw, const float* in, float* out) {
  const double nfft = (double)nextpow2_ll((long long)nt + nw - 1);
  const int shift = nw / 2 + nw / 10;
  for (int t = 0; t < ntr; ++t) {
    const float* d = in + (size_t)t * nt;
    float* o = out + (size_t)t * nt;
    for (int k = 0; k < nt; ++k) {
      double acc = 0.0;
      if (adjoint) {
        if (k < shift) {
          o[k] = 0.0f;
          continue;
        }
        const int k0 = k - shift;
        for (int m = 0; m <= std::min(nw - 1, nt - 1 - k0); ++m) acc += (double)w[m] * (double)d[k0 + m];
      } else {
        const int kk = k + shift;
        for (int m = std::max(0, kk - nw + 1); m <= std::min(kk, nt - 1); ++m) acc += (double)d[m] * (double)w[kk - m];
      }
      o[k] = (float)(acc * nfft);
    }
  }
}

// float -> int rounding down as the GPU converts (NaN -> 0, saturating).
int f2i_rd(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return 2147483647;
  if (v < -2147483648.0f) return (-2147483647 - 1);
  return (int)std::floor(v);
}

// Image-domain adjoint columns may run on several threads (nthreads > 1):
// each image point receives its own sum, so the result is the same; the angle
// domain and the forward add into shared bins and stay serial.
void kdm_trace(int domain, int adjoint, const float* S, const float* Ts, const float* Tr, float* data, float* model,
               float dt, int nt, float dx, float dz, int nxx, int nzz, float aperture, float cmp, int cmpId, float da,
               int na, int nthreads = 1) {
  const int nz = nzz - 2 * NB;
  const float tan_ap = std::tan((float)((double)aperture * M_PI / 180.0));
  const float scale = (float)(1.0 / (2.0 * M_PI));
  const int nth = (domain == 0 && adjoint) ? std::max(nthreads, 1) : 1;
#pragma omp parallel for schedule(static) num_threads(nth)
  for (int j = NB; j < nxx - NB; ++j)
    for (int i = NB; i < nzz - NB; ++i) {
      const size_t c = i + (size_t)j * nzz;
      const float z = (float)(i - NB) * dz, x = (float)(j - NB) * dx;
      const float T = Ts[c] + Tr[c];
      const int tId = f2i_rd(T / dt);
      if (tId < 0 || tId >= nt) continue;
      auto at = [&](const float* A, int ii, int jj) { return A[ii + (size_t)jj * nzz]; };
      const float dTs_dz = 0.5f * (at(Ts, i + 1, j) - at(Ts, i - 1, j)) / dz;
      const float dTs_dx = 0.5f * (at(Ts, i, j + 1) - at(Ts, i, j - 1)) / dx;
      const float d2Ts_dz2 = (at(Ts, i + 1, j) - 2.0f * at(Ts, i, j) + at(Ts, i - 1, j)) / (dz * dz);
      const float d2Ts_dx2 = (at(Ts, i, j + 1) - 2.0f * at(Ts, i, j) + at(Ts, i, j - 1)) / (dx * dx);
      const float d2Ts_dxdz = (at(Ts, i + 1, j + 1) - at(Ts, i - 1, j + 1) - at(Ts, i + 1, j - 1) +
                               at(Ts, i - 1, j - 1)) / (4.0f * dx * dz);
      const float dTr_dz = 0.5f * (at(Tr, i + 1, j) - at(Tr, i - 1, j)) / dz;
      const float dTr_dx = 0.5f * (at(Tr, i, j + 1) - at(Tr, i, j - 1)) / dx;
      const float d2Tr_dz2 = (at(Tr, i + 1, j) - 2.0f * at(Tr, i, j) + at(Tr, i - 1, j)) / (dz * dz);
      const float d2Tr_dx2 = (at(Tr, i, j + 1) - 2.0f * at(Tr, i, j) + at(Tr, i, j - 1)) / (dx * dx);
      const float d2Tr_dxdz = (at(Tr, i + 1, j + 1) - at(Tr, i - 1, j + 1) - at(Tr, i + 1, j - 1) +
                               at(Tr, i - 1, j - 1)) / (4.0f * dx * dz);
      const float norm_Ts = std::sqrt(dTs_dx * dTs_dx + dTs_dz * dTs_dz) + KDM_EPS;
      const float norm_Tr = std::sqrt(dTr_dx * dTr_dx + dTr_dz * dTr_dz) + KDM_EPS;
      const float detHs = std::fabs(d2Ts_dx2 * d2Ts_dz2 - d2Ts_dxdz * d2Ts_dxdz) + KDM_EPS;
      const float detHr = std::fabs(d2Tr_dx2 * d2Tr_dz2 - d2Tr_dxdz * d2Tr_dxdz) + KDM_EPS;
      const float Jterm = std::sqrt(detHs * detHr) * S[c] / (norm_Ts * norm_Tr);
      const float ux_s = dTs_dx / norm_Ts, uz_s = dTs_dz / norm_Ts;
      const float ux_r = dTr_dx / norm_Tr, uz_r = dTr_dz / norm_Tr;
      const float nx_n = 0.0f, nz_n = -1.0f;  // surface normal
      const float cos_s = std::fabs(ux_s * nx_n + uz_s * nz_n);
      const float cos_r = std::fabs(ux_r * nx_n + uz_r * nz_n);
      const float obliquity = std::sqrt(std::fmax(0.0f, cos_s) * std::fmax(0.0f, cos_r));
      const float R_s = std::fmax(Ts[c] / S[c], KDM_EPS);
      const float R_r = std::fmax(Tr[c] / S[c], KDM_EPS);
      const float G = 1.0f / std::sqrt(R_s * R_r);
      const float theta_s = std::acos(std::fmin(1.0f, std::fmax(-1.0f, ux_s * nx_n + uz_s * nz_n)));
      const float theta_r = std::acos(std::fmin(1.0f, std::fmax(-1.0f, ux_r * nx_n + uz_r * nz_n)));
      const float theta = 0.5f * (theta_s + theta_r);
      const float Rf = 1.0f + 0.2f * std::cos(theta);
      const float sigma = tan_ap * z;
      const float taper = std::exp(-0.5f * ((x - cmp) / (sigma + KDM_EPS)) * ((x - cmp) / (sigma + KDM_EPS)));
      const float weights = scale * Jterm * obliquity * G * Rf * taper;
      size_t mId;
      if (domain == 0) {
        mId = (size_t)(i - NB) + (size_t)(j - NB) * nz;
      } else {
        const float ra = 0.5f * std::acos((dTs_dx * dTr_dx + dTs_dz * dTr_dz) / (norm_Ts * norm_Tr));
        const float ang = (float)((double)(180.0f * ra) / M_PI);
        const int aId = f2i_rd(ang / da);
        if (aId < 0 || aId >= na) continue;
        mId = (size_t)(i - NB) + (size_t)aId * nz + (size_t)cmpId * na * nz;
      }
      if (adjoint)
        model[mId] += weights * data[tId];
      else
        data[tId] += weights * model[mId];
    }
}

}  // namespace

// ===========================================================================
// C entry points (ctypes).  All grids are z-fastest; "padded" grids are
// (nz+6) x (nx+6).
extern "C" {

void orc_expand_boundary_f32(const float* in, float* out, int nz, int nx) { expand_boundary(in, out, nz, nx); }
void orc_expand_boundary_f64(const double* in, double* out, int nz, int nx) { expand_boundary(in, out, nz, nx); }
void orc_reduce_boundary_f32(const float* in, float* out, int nz, int nx) { reduce_boundary(in, out, nz, nx); }

void orc_init_f32(float* T, const float* S, int nzz, int nxx, float dx, float dz, float sx, float sz, int* sIdx, int* sIdz) {
  init_time(T, S, nzz, nxx, dx, dz, sx, sz, sIdx, sIdz);
}
void orc_init_f64(double* T, const double* S, int nzz, int nxx, double dx, double dz, double sx, double sz, int* sIdx, int* sIdz) {
  init_time(T, S, nzz, nxx, dx, dz, sx, sz, sIdx, sIdz);
}

// mode 0 = ref (one pass), 1 = conv; schedule 0 = raster, 1 = literal diagonal
int orc_fsm_f32(float* T, const float* S, int nzz, int nxx, float dx, float dz, int mode, int max_passes, int schedule) {
  return fsm(T, S, nzz, nxx, dx, dz, mode, max_passes, schedule);
}
int orc_fsm_f64(double* T, const double* S, int nzz, int nxx, double dx, double dz, int mode, int max_passes, int schedule) {
  return fsm(T, S, nzz, nxx, dx, dz, mode, max_passes, schedule);
}
void orc_sweep_f32(float* T, const float* S, int nzz, int nxx, float dx, float dz, int sweep) {
  sweep_raster(T, S, nzz, nxx, sweep, make_consts(dx, dz));
}

// Full iso solve on a padded slowness grid: init + fsm (Eikonal_ISO::time_propagation,
// src/modeling/eikonal_iso.cu:9-13).
int orc_solve_iso_f32(float* T, const float* S, int nzz, int nxx, float dx, float dz, float sx, float sz, int mode, int max_passes) {
  init_time(T, S, nzz, nxx, dx, dz, sx, sz, nullptr, nullptr);
  return fsm(T, S, nzz, nxx, dx, dz, mode, max_passes, 0);
}
int orc_solve_iso_f64(double* T, const double* S, int nzz, int nxx, double dx, double dz, double sx, double sz, int mode, int max_passes) {
  init_time(T, S, nzz, nxx, dx, dz, sx, sz, nullptr, nullptr);
  return fsm(T, S, nzz, nxx, dx, dz, mode, max_passes, 0);
}

// Shot-parallel batch of iso solves on windows of one model (the CPU timing
// path).  Shot k uses columns [x0[k], x0[k]+w[k]) of the unpadded model s
// (nz x nx), padded by replication like a standalone model; T for shot k is
// written at T + toff[k] with shape (nz+6) x (w[k]+6).  nthreads<=0 uses the
// OpenMP default.
void orc_solve_windows_f32(const float* s, int nz, int nx, float dx, float dz, int nshots,
                           const float* sx, const float* sz, const int* x0, const int* w,
                           const long long* toff, float* T, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads != 1)
  for (int k = 0; k < nshots; ++k) {
    const int nzz = nz + 2 * NB, nxx = w[k] + 2 * NB;
    std::vector<float> crop((size_t)nz * w[k]), S((size_t)nzz * nxx);
    std::memcpy(crop.data(), s + (size_t)x0[k] * nz, crop.size() * sizeof(float));
    expand_boundary(crop.data(), S.data(), nz, w[k]);
    float* Tk = T + toff[k];
    init_time(Tk, S.data(), nzz, nxx, dx, dz, sx[k] - x0[k] * dx, sz[k], nullptr, nullptr);
    fsm(Tk, S.data(), nzz, nxx, dx, dz, 0, 1, 0);
  }
}

void orc_compress(const float* in, uint16_t* out, long long n, float* maxv, float* minv) {
  compress(in, out, (size_t)n, maxv, minv);
}

void orc_quasi_slowness(const float* T, float* S, int nzz, int nxx, float dx, float dz, int sIdx, int sIdz,
                        const uint16_t* c11, const uint16_t* c13, const uint16_t* c15, const uint16_t* c33,
                        const uint16_t* c35, const uint16_t* c55, const float* mn, const float* mx) {
  const uint16_t* C[6] = {c11, c13, c15, c33, c35, c55};
  quasi_slowness(T, S, nzz, nxx, dx, dz, sIdx, sIdz, C, mn, mx);
}

// Eikonal_ANI::time_propagation (src/modeling/eikonal_ani.cu:33-45) on padded
// grids; S_work receives the qP slowness used by the second solve.
void orc_solve_ani_f32(float* T, const float* S_iso, float* S_work, int nzz, int nxx, float dx, float dz,
                       float sx, float sz, const uint16_t* c11, const uint16_t* c13, const uint16_t* c15,
                       const uint16_t* c33, const uint16_t* c35, const uint16_t* c55, const float* mn,
                       const float* mx) {
  const size_t matsize = (size_t)nzz * nxx;
  std::memcpy(S_work, S_iso, matsize * sizeof(float));
  int sIdx, sIdz;
  init_time(T, S_work, nzz, nxx, dx, dz, sx, sz, &sIdx, &sIdz);
  fsm(T, S_work, nzz, nxx, dx, dz, 0, 1, 0);
  const uint16_t* C[6] = {c11, c13, c15, c33, c35, c55};
  quasi_slowness(T, S_work, nzz, nxx, dx, dz, sIdx, sIdz, C, mn, mx);
  init_time(T, S_work, nzz, nxx, dx, dz, sx, sz, nullptr, nullptr);
  fsm(T, S_work, nzz, nxx, dx, dz, 0, 1, 0);
}

void orc_picks_f32(const float* T, int nzz, float dx, float dz, const float* rx, const float* rz, int nrec, float* out) {
  for (int r = 0; r < nrec; ++r) out[r] = pick_one(T, nzz, dx, dz, rx[r], rz[r]);
}
void orc_picks_f64(const double* T, int nzz, double dx, double dz, const double* rx, const double* rz, int nrec, double* out) {
  for (int r = 0; r < nrec; ++r) out[r] = pick_one(T, nzz, dx, dz, rx[r], rz[r]);
}

double orc_adjoint_f32(const float* S, int nzz, int nxx, float dx, float dz, float sx, float sz, const float* rx,
                       const float* rz, int nrec, const float* obs, double* g, float* picks) {
  return adjoint(S, nzz, nxx, dx, dz, sx, sz, rx, rz, nrec, obs, g, picks);
}
double orc_adjoint_f64(const double* S, int nzz, int nxx, double dx, double dz, double sx, double sz, const double* rx,
                       const double* rz, int nrec, const double* obs, double* g, double* picks) {
  return adjoint(S, nzz, nxx, dx, dz, sx, sz, rx, rz, nrec, obs, g, picks);
}

double orc_batch_misfit_f32(const float* s, int nz, int nx, float dx, float dz, int nshots, const float* sx,
                            const float* sz, const int* x0, const int* w, const int* rec_off, const int* nrec,
                            const float* rx, const float* rz, const float* obs, double* grad, float* picks,
                            int nthreads) {
  return batch_misfit(s, nz, nx, dx, dz, nshots, sx, sz, x0, w, rec_off, nrec, rx, rz, obs, grad, picks, nthreads);
}
double orc_batch_misfit_f64(const double* s, int nz, int nx, double dx, double dz, int nshots, const double* sx,
                            const double* sz, const int* x0, const int* w, const int* rec_off, const int* nrec,
                            const double* rx, const double* rz, const double* obs, double* grad, double* picks,
                            int nthreads) {
  return batch_misfit(s, nz, nx, dx, dz, nshots, sx, sz, x0, w, rec_off, nrec, rx, rz, obs, grad, picks, nthreads);
}

void orc_branch_counts_f32(const float* S, int nzz, int nxx, float dx, float dz, float sx, float sz,
                           long long* counts) {
  branch_counts(S, nzz, nxx, dx, dz, sx, sz, counts);
}
// One raster sweep on T in place, adding each node's deciding candidate to
// counts[7] (kinds as in branch_counts).
void orc_sweep_counted_f32(float* T, const float* S, int nzz, int nxx, float dx, float dz, int sweep,
                           long long* counts) {
  sweep_counted(T, S, nzz, nxx, sweep, make_consts(dx, dz), counts);
}

void orc_tikhonov_apply(int order, const double* x, long long n, double* y) { tikhonov_apply(order, x, n, y); }
void orc_tikhonov_adjoint(int order, const double* y, long long n, double* x) { tikhonov_adjoint(order, y, n, x); }
void orc_smooth_model(const float* in, float* out, int nz, int nx, float stdv, int samples) {
  smooth_model(in, out, nz, nx, stdv, samples);
}

int orc_ray_paths(const float* T, int nz, int nx, float dx, float dz, float sx, float sz, const float* rx,
                  const float* rz, int nrec, int max_steps, int* out_rec, int* out_cell, float* out_len, int cap) {
  return ray_paths(T, nz, nx, dx, dz, sx, sz, rx, rz, nrec, max_steps, out_rec, out_cell, out_len, cap);
}

int orc_ricker(float dt, float fmax, float* out, int cap) {
  std::vector<float> w;
  ricker(dt, fmax, w);
  if ((int)w.size() <= cap) std::copy(w.begin(), w.end(), out);
  return (int)w.size();
}
void orc_trace_convolve(int adjoint, int ntr, int nt, const float* w, int nw, const float* in, float* out) {
  trace_convolve(adjoint, ntr, nt, w, nw, in, out);
}
// Traces in order, each the reference's per-trace kernel launch: data is
// ntr x nt, tables are consecutive padded grids.
void orc_kdm(int domain, int adjoint, const float* S, const float* src_tables, const float* rec_tables, int ntr,
             const int* isrc, const int* irec, const float* cmp, const int* cmp_id, float* data, float* model,
             int nt, float dt, float dx, float dz, int nzz, int nxx, float aperture, float da, int na, int nthreads) {
  const size_t m = (size_t)nzz * nxx;
  for (int t = 0; t < ntr; ++t)
    kdm_trace(domain, adjoint, S, src_tables + isrc[t] * m, rec_tables + irec[t] * m, data + (size_t)t * nt, model,
              dt, nt, dx, dz, nxx, nzz, aperture, cmp[t], cmp_id ? cmp_id[t] : 0, da, na, nthreads);
}
}  // extern "C"
