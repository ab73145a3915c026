// seisfat_eik_solve_noalloc: the per-shot solve for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops).
//
// The variant of seisfat_eik_solve whose work buffers (shot record, tile
// queue, ticket counters, epoch, error flag, boundary exchange words) live in
// one fixed device array per GPU, `scratch`, instead of a batch's hipMalloc'd
// buffers; the caller passes the model handle (input), the traveltime grid and
// the picks (outputs) as device pointers.  It runs the same kernels as the
// batched persistent schedule with one launch per sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "eik_handle.h"

using namespace seisfat;

// Layout (bytes): the control block (ticket counters, epoch, error flag) at 0,
// the shot record at 256, the tile queue at 512, the exchange words after it
// (256-aligned).  Each GPU the library runs on holds one copy until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: host code uses only the per-device address that
// hipGetSymbolAddress returns.
constexpr size_t SCRATCH_BYTES = size_t(8) << 20;
static __device__ __align__(256) unsigned long long scratch[SCRATCH_BYTES / 8];
static_assert(CTL_WORDS * sizeof(int) <= 256, "control block exceeds its scratch slot");
static_assert(sizeof(ShotDev) <= 256, "shot record exceeds its scratch slot");

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
char* g_scratch[SCRATCH_MAX_DEVICES] = {};

// Once per device, with that device current: the device address of scratch.
char* scratch_on(int device) {
  if (device < 0 || device >= SCRATCH_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch[device] = static_cast<char*>(p);
  }
  return g_scratch[device];
}

// Set-up of a one-shot persistent solve (four per-sweep launches): the shot
// record, the tile queue in tile-diagonal order (one shot: tile (r, bb) on
// diagonal 2r + bb, highest priority, the order the batches' queue has), the
// ticket counters, a new epoch, the error flag, and every exchange word reset
// to the all-ones tag that no run uses.  Everything a previous call left in
// the array is overwritten before it is read.
__global__ void k_single_setup(ShotDev sd, ShotDev* __restrict__ d_shot, int4* __restrict__ tiles, int nstrips,
                               int nblocks, int* __restrict__ ctl, unsigned long long* __restrict__ words,
                               size_t nwords) {
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = tid; q < nwords; q += stride) words[q] = ~0ull;
  if (tid == 0) {
    *d_shot = sd;
    int n = 0;
    const int D = 2 * (nstrips - 1) + nblocks;
    for (int d = 0; d < D; ++d)
      for (int r = 0; r < nstrips; ++r) {
        const int bb = d - 2 * r;
        if (bb < 0) break;
        if (bb < nblocks) tiles[n++] = make_int4(0, r, bb, 3);
      }
    for (int c = 0; c < NUM_COUNTERS; ++c) ctl[c] = 0;
    unsigned* ep = reinterpret_cast<unsigned*>(ctl + CTL_EPOCH);
    *ep = (*ep + 1) % 2097151u;
    ctl[CTL_ERR] = 0;
    ctl[CTL_SKIP] = 0;
  }
}

__global__ void k_copy_int(const int* __restrict__ src, int* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = *src;
}

}  // namespace

extern "C" int seisfat_eik_solve_noalloc(seisfat_eik_t h, float sx, float sz, float* d_T, const float* d_rx,
                                         const float* d_rz, int nrec, float* d_picks, int* d_err, void* stream) {
  if (!h) return set_error(SEISFAT_EINVAL, "null handle");
  if (!d_T) return set_error(SEISFAT_EINVAL, "null traveltime buffer");
  if (nrec < 0 || (nrec > 0 && (!d_rx || !d_rz || !d_picks)))
    return set_error(SEISFAT_EINVAL, "receiver and picks buffers required when nrec > 0");
  if (!(sx >= 0.0f && sx <= (float)(h->nx - 1) * h->dx && sz >= 0.0f && sz <= (float)(h->nz - 1) * h->dz))
    return set_error(SEISFAT_EINVAL, "source outside the model");
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != h->device)
    return set_error(SEISFAT_ESTATE, "call with the handle's device current");
  char* base = scratch_on(h->device);
  if (!base) return set_error(SEISFAT_EHIP, "hipGetSymbolAddress of the solve's device state failed");
  const int nzz = h->nzz, nxx = h->nxx;
  const int nstrips = (nzz - 2 + 63) / 64;
  const int nblocks = (nxx - 2 + 63 + 63) / 64;  // 64-column parallelograms reach 63 columns past their block
  const int ldh = (nxx - 2 + 63) / 64 * 64;
  const size_t ntiles = (size_t)nstrips * nblocks;
  const size_t tiles_bytes = (ntiles * sizeof(int4) + 255) / 256 * 256;
  const size_t nwords = (size_t)nstrips * ldh + (size_t)nstrips * 2 * 128;
  if (512 + tiles_bytes + nwords * sizeof(unsigned long long) > SCRATCH_BYTES)
    return set_error(SEISFAT_EINVAL, "model too large for the allocation-free solve: its tile queue and "
                                     "exchange words exceed the fixed 8 MiB of device state");
  if (!h->has_s) return set_error(SEISFAT_ESTATE, "slowness not set (seisfat_eik_set_slowness)");
  int* ctl = reinterpret_cast<int*>(base);
  ShotDev* d_shot = reinterpret_cast<ShotDev*>(base + 256);
  int4* tiles = reinterpret_cast<int4*>(base + 512);
  unsigned long long* words = reinterpret_cast<unsigned long long*>(base + 512 + tiles_bytes);
  ShotDev sd{};
  sd.T = d_T;
  sd.sbase = h->d_spad;  // the whole model: its padding as uploaded
  sd.sofs = 0;
  sd.smax = nxx - 1;
  sd.nxx = nxx;
  sd.sx = sx;
  sd.sz = sz;
  sd.x0 = 0;
  sd.w = h->nx;
  sd.xh = words;
  sd.xc = words + (size_t)nstrips * ldh;
  sd.ldh = ldh;
  hipStream_t st = (hipStream_t)stream;
  const PersistRun pr{ctl, reinterpret_cast<unsigned*>(ctl + CTL_EPOCH), ctl + CTL_ERR};
  const int blocks = (int)std::min<size_t>((nwords + 255) / 256 + 1, 1024);
  hipLaunchKernelGGL(k_single_setup, dim3(blocks), dim3(256), 0, st, sd, d_shot, tiles, nstrips, nblocks, ctl,
                     words, nwords);
  // the boundary ring of the caller's grid must read 1e6 (sweeps never write it)
  launch_fill(d_T, (size_t)nzz * nxx, TINF, st);
  launch_sweeps_persist(d_shot, tiles, (int)ntiles, nstrips, nzz, h->kc, h->eqh, h->skew, true, pr, 0, st);
  launch_picks(d_shot, nullptr, d_rx, d_rz, d_picks, nrec, nzz, h->dx, h->dz, st, pr.err);
  if (d_err) hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(64), 0, st, pr.err, d_err);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(SEISFAT_EHIP, hipGetErrorString(e));
  return SEISFAT_OK;
}
