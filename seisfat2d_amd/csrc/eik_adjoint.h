// Adjoint-state gradient of the single-pass eikonal solve: device layout and
// launchers (host side of eik_adjoint.hip).
//
// The forward solve is init + four directional sweeps; node (i, j) of sweep s
// becomes min(T_old, cand(tv, te, tev, S)) with tv, te, tev its upwind
// neighbours after sweep s.  The adjoint walks the sweeps backwards; for each
// sweep it
//   1. re-derives, for every node in parallel, the branch the forward took
//      and its local partials d(new)/d(tv, te, tev, S) from a recorded fp64
//      trajectory of the solve (k_fwd64_sweep) — k_adj_coef;
//   2. runs the linear recurrence
//        mu(n) = lam_in(n) + sum over the three nodes m that read n of
//                mu(m) * d(new_m)/d(n)
//      in reverse dependency order — k_adj_sweep, one wave per 64-row strip;
//   3. folds mu(n) * d(new_n)/dS into the per-shot slowness gradient — k_adj_fold.
// The source initialisation T0 = S(node) * distance and the FP64 cross-shot
// reduction onto the unpadded model close the chain.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eik_device.h"

namespace seisfat {

constexpr int ADJ_PF = 8;  // prefetch distance and word-block length of the strip walkers (steps)

// A shot as the adjoint kernels see it.
struct AdjShot {
  ShotDev sd;   // final traveltimes (sd.T), slowness view, width, source, window
  size_t toff;  // offset of the shot in the natural-layout arrays (snapshots, gshot)
  size_t koff;  // offset in the strip-skewed arrays
  int NTP;      // steps per strip: nxx + 63 rounded up to ADJ_PF
};

// Forward sweep s runs rows in direction a and columns in direction b
// (s0 (+,+), s1 (-,+), s2 (+,-), s3 (-,-)).  Its adjoint processes
// "reverse-local" coordinates (ip, jp) in increasing order; a node reads the
// messages of (ip-1, jp), (ip, jp-1), (ip-1, jp-1).
__host__ __device__ inline int adj_a(int s) { return (s & 1) ? -1 : 1; }
__host__ __device__ inline int adj_b(int s) { return (s & 2) ? -1 : 1; }
__host__ __device__ inline int adj_ip(int i, int a, int nzz) { return a > 0 ? nzz - 1 - i : i; }
__host__ __device__ inline int adj_jp(int j, int b, int nxx) { return b > 0 ? nxx - 1 - j : j; }

// Strip-skewed layout: reverse-local node (ip, jp) is held by strip ip/64,
// lane ip%64, at step jp + lane, so the 64 nodes one wave visits at one step
// are 64 consecutive words.
__host__ __device__ inline size_t adj_kidx(int ip, int jp, int NTP) {
  const int lane = ip & 63;
  return ((size_t)(ip >> 6) * NTP + (size_t)(jp + lane)) * 64 + lane;
}

// Pick seeds: lam[node[n]] = sum_{e in ptr[n]..ptr[n+1]} w[rec[e]] * wt[e].
void launch_adj_seed(const size_t* node, const int* ptr, const int* rec, const float* wt, int nseed,
                     const float* w, float* lam, hipStream_t st);
// fp64 trajectory: T0 stencil (T0 otherwise 1e6), and sweep s: Tnew = sweep(Told).
void launch_fill_d(double* p, size_t n, double v, hipStream_t st);
void launch_fwd64_init(const AdjShot* shots, int nshots, int nzz, double dx, double dz, double* T0, hipStream_t st);
// walk_ctl: 2 ints (ticket counter, reset by the launch; wait-limit flag);
// walk_x: nshots * nstr * ldx exchange words (ldx >= 2 * nxx); tag: a value
// no earlier launch on the same walk_x used
void launch_fwd64_sweep(int sweep, const AdjShot* shots, int nshots, int nzz, int nstr, double dx, double dz,
                        const double* Told, double* Tnew, int* walk_ctl, unsigned long long* walk_x, int ldx,
                        unsigned tag, hipStream_t st);
// Picks of the fp64 trajectory's final traveltimes (rx, rz relative to each
// receiver's shot grid).
void launch_picks64(const AdjShot* shots, const int* rec_shot, const float* rx, const float* rz, int nrec, int nzz,
                    double dx, double dz, const double* T4, double* out, hipStream_t st);
// Local partials of sweep s from the fp64 trajectory (Tnew = T after sweep s,
// Told = before); also moves lam_src (layout of sweep s+1) into lam_dst
// (layout of sweep s) when lam_src is non-null.  nstr: 64-row strips per
// shot; max_nxx: the widest shot's padded width.
void launch_adj_coef(int sweep, const AdjShot* shots, int nshots, int nstr, int max_nxx, int nzz, double dx,
                     double dz, const double* Tnew, const double* Told, const float* lam_src, float* lam_dst,
                     float4* coef, uint8_t* code, hipStream_t st);
void launch_adj_sweep(const AdjShot* shots, int nshots, int nstr, const float* lam_in, const float4* coef,
                      float* lam_out, float* gc, int* walk_ctl, unsigned long long* walk_x, int ldx, unsigned tag,
                      hipStream_t st);
void launch_adj_fold(int sweep, const AdjShot* shots, int nshots, int nstr, int max_nxx, int nzz, const float* gc,
                     const uint8_t* code, float* gshot, hipStream_t st);
void launch_adj_source(const AdjShot* shots, int nshots, int nzz, double dx, double dz, const float* lam0,
                       float* gshot, hipStream_t st);
// fixed_border: the model's border is not the replication of its edges (a
// padded upload), so only interior nodes (and windows' replicated columns) fold
void launch_adj_reduce(const AdjShot* shots, int nshots, int nz, int nx, int nzz, const float* gshot,
                       double* grad, bool fixed_border, hipStream_t st);

}  // namespace seisfat
