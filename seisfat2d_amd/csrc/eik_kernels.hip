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
//   k_zpad            the row part of expand_boundary (src/modeling/modeling.cu:204-232)
#include "eik_kernels.h"

namespace seisfat {

// ---------------------------------------------------------------------------
__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Row-pad an unpadded nz x nx model into nzz x nx (rows 0..2 copy model row 0,
// the last three copy row nz-1).  Columns are padded virtually by ShotDev.
__global__ void k_zpad(const float* __restrict__ s, float* __restrict__ szp, int nz, int nx) {
  const int nzz = nz + 2 * NB;
  const size_t n = (size_t)nzz * nx;
  size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; idx < n; idx += stride) {
    const int i = (int)(idx % nzz);
    const int j = (int)(idx / nzz);
    const int im = min(max(i - NB, 0), nz - 1);
    szp[idx] = s[im + (size_t)j * nz];
  }
}

// Flags non-finite or non-positive slowness (bit 0 of *bad).
__global__ void k_check_positive(const float* __restrict__ s, size_t n, int* __restrict__ bad) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  int flag = 0;
  for (; i < n; i += stride) {
    const float v = s[i];
    if (!(v > 0.0f) || !(v < __builtin_inff())) flag = 1;
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
    T[idx] = fminf(T[idx], eik_local<EQH>(tv, te, tev, sref, szo, sxo, k));
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

__device__ __forceinline__ float dpp_shr1(float old, float src) {
  // wave_shr:1 (DPP control 0x138): lane k receives lane k-1; lane 0 keeps `old`.
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), 0x138,
                                                    0xf, 0xf, false));
}

constexpr int RING_LD = 66;

template <int SA, int SB, bool EQH>
__global__ __launch_bounds__(64) void k_sweep_tiles(const ShotDev* __restrict__ shots,
                                                    const int4* __restrict__ tiles, int nzz,
                                                    EikConsts k) {
  __shared__ float sT[64 * RING_LD];
  __shared__ float sS[64 * RING_LD];
  __shared__ float sH[64];

  const int4 tl = tiles[blockIdx.x];
  const ShotDev sd = shots[tl.x];
  const int r = tl.y;
  const int c0 = 64 * tl.z;
  const int lane = threadIdx.x;
  const int nxx = sd.nxx;
  const int NR = nzz - 2, NC = nxx - 2;
  const int kg = 64 * r + lane;
  const bool row_ok = kg < NR;
  gfloat* __restrict__ T = gptr(sd.T);

  // Staging.  Every load is unconditional: positions outside the interior are
  // clamped onto the boundary ring (local row -1/NR, column -1/NC), whose
  // traveltime is always 1e6, and slowness cells are clamped into the grid
  // (those values only feed inactive lanes).  All 129 ring loads are issued
  // before the first LDS write so their latencies overlap.
  const int kgc = min(kg, NR);
  float tv_[64], sv_[64];
#pragma unroll
  for (int v = 0; v < 64; ++v) {
    // old traveltimes: ring slot v of lane k holds column c = c0-k+((v+k)&63)
    const int c = min(max(c0 - lane + ((v + lane) & 63), -1), NC);
    tv_[v] = T[node_off<SA, SB>(kgc, c, nzz, nxx)];
  }
  const int kapc = min(64 * r + lane, NR);
#pragma unroll
  for (int v = 0; v < 64; ++v) {
    // slowness cells: ring row kap (0..64) holds cell columns
    // lam in [c0-kap+1, c0-kap+64]; slot = lam & 63
    const int lam = min(max(c0 + 1 - lane + ((v + lane - 1) & 63), 0), NC);
    sv_[v] = cell_at<SA, SB>(sd, nzz, kapc, lam);
  }
  // ring row 64 (lane = slot), the upwind row above the strip for lane 0
  // (local row 64r-1, columns c0..c0+63) and the carry-in from block bb-1:
  // te = T(kg, c0-k-1), tev = T(kg-1, c0-k-1), sref = cell (kg, c0-k).
  const float s64 = cell_at<SA, SB>(sd, nzz, min(64 * r + 64, NR),
                                    min(max(c0 + 1 - 64 + ((lane + 63) & 63), 0), NC));
  const float hval = T[node_off<SA, SB>(64 * r - 1, min(c0 + lane, NC), nzz, nxx)];
  const int cin = min(max(c0 - lane - 1, -1), NC);
  float te = T[node_off<SA, SB>(kgc, cin, nzz, nxx)];
  float tev = T[node_off<SA, SB>(min(kg - 1, NR), cin, nzz, nxx)];
  float sref = cell_at<SA, SB>(sd, nzz, min(kg, NR), min(max(c0 - lane, 0), NC));
#pragma unroll
  for (int v = 0; v < 64; ++v) sT[v * RING_LD + lane] = tv_[v];
#pragma unroll
  for (int v = 0; v < 64; ++v) sS[v * RING_LD + lane] = sv_[v];
  sS[lane * RING_LD + 64] = s64;
  sH[lane] = hval;
  __syncthreads();

#pragma unroll 4
  for (int t = 0; t < 64; ++t) {
    const int slot = (t - lane) & 63;
    const float told = sT[slot * RING_LD + lane];
    const float szo = sS[((slot + 1) & 63) * RING_LD + lane];
    const float sxo = sS[slot * RING_LD + lane + 1];
    const float tv = dpp_shr1(sH[t], te);
    const int c = c0 + t - lane;
    const bool act = row_ok && (unsigned)c < (unsigned)NC;
    const float cand = eik_local<EQH>(tv, te, tev, sref, szo, sxo, k);
    const float res = act ? fminf(told, cand) : told;
    sT[slot * RING_LD + lane] = res;
    tev = tv;
    te = res;
    sref = szo;
  }
  __syncthreads();

#pragma unroll 8
  for (int v = 0; v < 64; ++v) {
    const int c = c0 - lane + ((v + lane) & 63);
    if (row_ok && c >= 0 && c < NC) T[node_off<SA, SB>(kg, c, nzz, nxx)] = sT[v * RING_LD + lane];
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

__global__ void k_picks(const ShotDev* __restrict__ shots, const int* __restrict__ rec_shot,
                        const float* __restrict__ rx, const float* __restrict__ rz,
                        float* __restrict__ out, int nrec, int nzz, float dx, float dz) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const gfloat* __restrict__ T = gptr(shots[rec_shot[r]].T);
  const float x = rx[r], z = rz[r];
  const float x0 = floorf((x + 0.5f * dx) / dx) * dx;
  const float z0 = floorf((z + 0.5f * dz) / dz) * dz;
  const float x1 = floorf((x + 0.5f * dx) / dx) * dx + dx;
  const float z1 = floorf((z + 0.5f * dz) / dz) * dz + dz;
  const float xd = (x - x0) / (x1 - x0);
  const float zd = (z - z0) / (z1 - z0);
  const int i = (int)((z + 0.5f * dz) / dz) + NB;
  const int j = (int)((x + 0.5f * dx) / dx) + NB;
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

__global__ void k_quasi_slowness(const ShotDev* __restrict__ iso, ShotDev* __restrict__ qp, int nzz,
                                 float dx, float dz, StiffDev st) {
  const ShotDev si = iso[blockIdx.y];
  const ShotDev so = qp[blockIdx.y];
  const int nxx = si.nxx;
  const size_t n = (size_t)nzz * nxx;
  const int sIdx = (int)((si.sx + 0.5f * dx) / dx) + NB;
  const int sIdz = (int)((si.sz + 0.5f * dz) / dz) + NB;
  gfloat* __restrict__ Sout = gptr(const_cast<float*>(so.sbase));
  const gfloat* __restrict__ T = gptr(si.T);
  size_t index = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; index < n; index += stride) {
    const int i = (int)(index % nzz);
    const int j = (int)(index / nzz);
    const float s_iso = s_at(si, nzz, i, j);
    float s_new = s_iso;
    if ((i >= NB) && (i < nzz - NB) && (j >= NB) && (j < nxx - NB) && !((i == sIdz) && (j == sIdx))) {
      const float dTz = 0.5f * (T[(i + 1) + (size_t)j * nzz] - T[(i - 1) + (size_t)j * nzz]) / dz;
      const float dTx = 0.5f * (T[i + (size_t)(j + 1) * nzz] - T[i + (size_t)(j - 1) * nzz]) / dx;
      const float norm = sqrtf(dTx * dTx + dTz * dTz);
      const float p0 = dTx / norm;
      const float p1 = dTz / norm;
      const float c11 = decode_c(st.c[0][index], st.mn[0], st.mx[0]);
      const float c13 = decode_c(st.c[1][index], st.mn[1], st.mx[1]);
      const float c15 = decode_c(st.c[2][index], st.mn[2], st.mx[2]);
      const float c33 = decode_c(st.c[3][index], st.mn[3], st.mx[3]);
      const float c35 = decode_c(st.c[4][index], st.mn[4], st.mx[4]);
      const float c55 = decode_c(st.c[5][index], st.mn[5], st.mx[5]);
      const float Ro = c33 * s_iso * s_iso;
      const float C00 = c11 / Ro / Ro, C01 = c13 / Ro / Ro, C02 = c15 / Ro / Ro;
      const float C11 = c33 / Ro / Ro, C12 = c35 / Ro / Ro, C22 = c55 / Ro / Ro;
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
    Sout[index] = s_new;
  }
}

// ---------------------------------------------------------------------------
// Host-side launchers.
void launch_fill(float* p, size_t n, float v, hipStream_t st) {
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, st, p, n, v);
}

void launch_zpad(const float* s, float* szp, int nz, int nx, hipStream_t st) {
  const size_t n = (size_t)(nz + 2 * NB) * nx;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(k_zpad, dim3(blocks), dim3(256), 0, st, s, szp, nz, nx);
}

void launch_check_positive(const float* s, size_t n, int* bad, hipStream_t st) {
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_check_positive, dim3(blocks), dim3(256), 0, st, s, n, bad);
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

template <int SA, int SB, bool EQH>
static void tiles_diag(const ShotDev* shots, const int4* tiles, int n, int nzz, const EikConsts& k,
                       hipStream_t st) {
  hipLaunchKernelGGL((k_sweep_tiles<SA, SB, EQH>), dim3(n), dim3(64), 0, st, shots, tiles, nzz, k);
}

int launch_sweeps_tiled(const ShotDev* shots, const int4* tiles, const std::vector<int>& diag_off,
                        int nzz, const EikConsts& k, bool eqh, hipStream_t st) {
  int launches = 0;
  const int D = (int)diag_off.size() - 1;
  for (int sweep = 0; sweep < 4; ++sweep) {
    for (int d = 0; d < D; ++d) {
      const int n = diag_off[d + 1] - diag_off[d];
      if (n <= 0) continue;
      const int4* tl = tiles + diag_off[d];
      switch (sweep * 2 + (eqh ? 1 : 0)) {
        case 0: tiles_diag<1, 1, false>(shots, tl, n, nzz, k, st); break;
        case 1: tiles_diag<1, 1, true>(shots, tl, n, nzz, k, st); break;
        case 2: tiles_diag<-1, 1, false>(shots, tl, n, nzz, k, st); break;
        case 3: tiles_diag<-1, 1, true>(shots, tl, n, nzz, k, st); break;
        case 4: tiles_diag<1, -1, false>(shots, tl, n, nzz, k, st); break;
        case 5: tiles_diag<1, -1, true>(shots, tl, n, nzz, k, st); break;
        case 6: tiles_diag<-1, -1, false>(shots, tl, n, nzz, k, st); break;
        default: tiles_diag<-1, -1, true>(shots, tl, n, nzz, k, st); break;
      }
      ++launches;
    }
  }
  return launches;
}

void launch_picks(const ShotDev* shots, const int* rec_shot, const float* rx, const float* rz, float* out,
                  int nrec, int nzz, float dx, float dz, hipStream_t st) {
  if (nrec <= 0) return;
  hipLaunchKernelGGL(k_picks, dim3((nrec + 255) / 256), dim3(256), 0, st, shots, rec_shot, rx, rz, out, nrec,
                     nzz, dx, dz);
}

void launch_quasi_slowness(const ShotDev* iso, ShotDev* qp, int nshots, int nzz, int nxx, float dx, float dz,
                           const StiffDev& stf, hipStream_t st) {
  const size_t n = (size_t)nzz * nxx;
  const int bx = (int)std::min<size_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(k_quasi_slowness, dim3(bx, nshots), dim3(256), 0, st, iso, qp, nzz, dx, dz, stf);
}

}  // namespace seisfat
