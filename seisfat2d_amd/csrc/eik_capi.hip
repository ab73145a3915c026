// C ABI of the eikonal solver: model state, validation, batch plans, HIP graphs.
//
// Drop-in for the device side of SeisFAT2D's Modeling classes
// (src/modeling/modeling.cuh:17-79): the reference keeps d_S/d_T per Modeling
// object and solves one shot per time_propagation(); here a handle owns the
// padded model (the reference's d_S, plus its skewed copies and compressed
// stiffness), and a batch owns the per-shot traveltime grids plus, for the
// tiled schedule, HIP graphs of the whole solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "../../include/seisfat2d_eik.h"
#include "eik_adjoint.h"
#include "eik_handle.h"
#include "eik_kernels.h"

using namespace seisfat;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(e_ == hipErrorOutOfMemory ? SEISFAT_ENOMEM : SEISFAT_EHIP,              \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                     \
  } while (0)

template <class T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) return SEISFAT_OK;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(SEISFAT_ENOMEM, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                    " bytes: " + hipGetErrorString(e));
  }
  return SEISFAT_OK;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace

int seisfat::set_error(int code, const char* msg) { return fail(code, msg); }



namespace {

int check_handle(seisfat_eik_t h) {
  if (!h) return fail(SEISFAT_EINVAL, "null handle");
  return SEISFAT_OK;
}

// Operating range of the local solver's normal-range division and square
// root (see eik_device.h): the traveltime increments slowness*spacing and the
// gradient scale slowness/spacing stay far from fp32 under- and overflow.
bool slowness_ok(float v, float hmin, float hmax) {
  if (!(v > 0.0f) || !std::isfinite(v)) return false;
  const double s = v;
  return s * hmin >= 1e-12 && s * hmax <= 1e12 && s / hmax >= 1e-12 && s / hmin <= 1e12;
}

EikConsts make_consts(float dx, float dz) {
  EikConsts k;
  k.dx = dx;
  k.dz = dz;
  k.dz2i = 1.0f / (dz * dz);
  k.dx2i = 1.0f / (dx * dx);
  k.hyp = std::sqrt(dx * dx + dz * dz);
  k.sum = k.dz2i + k.dx2i;
  k.prod = k.dz2i * k.dx2i;
  k.sum4 = 4.0f * k.sum;  // exact (power-of-two scaling)
  k.fastdiv = 0;
  // zh = RN(1/b); 1 - b*zh is exact (one fma), so zl = RN((1 - b*zh) / b)
  // = RN(1/b - zh)
  auto two_word = [](float b, float& zh, float& zl) {
    zh = 1.0f / b;
    zl = std::fma(-b, zh, 1.0f) / b;
  };
  two_word(k.sum, k.zh_sum, k.zl_sum);
  two_word(k.hyp, k.zh_hyp, k.zl_hyp);
  two_word(k.dz, k.zh_dz, k.zl_dz);
  two_word(k.dx, k.zh_dx, k.zl_dx);
  const float a = TINF * k.dz2i + TINF * k.dx2i;  // RN(RN(.) + RN(.)): no contraction in this build
  k.tinf_fixed = (a / k.sum) >= TINF;
  return k;
}

void free_batch(seisfat_batch_s* b) {
  if (!b) return;
  for (auto& g : b->gexec)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& g : b->graph)
    if (g) (void)hipGraphDestroy(g);
  for (auto& g : b->conv_exec)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& g : b->conv_graph)
    if (g) (void)hipGraphDestroy(g);
  dfree(b->d_conv);
  for (auto& ev : b->ev_sweep)
    if (ev) (void)hipEventDestroy(ev);
  if (b->ev_done) (void)hipEventDestroy(b->ev_done);
  dfree(b->d_iso);
  dfree(b->d_qp);
  dfree(b->d_T);
  dfree(b->d_swork);
  dfree(b->d_rx);
  dfree(b->d_rz);
  dfree(b->d_recshot);
  dfree(b->d_picks);
  dfree(b->d_tiles);
  dfree(b->d_xch);
  dfree(b->d_ctl);
  dfree(b->d_iso_b);
  dfree(b->d_qp_b);
  dfree(b->d_ptiles);
  dfree(b->d_qoff);
  AdjState& a = b->adj;
  dfree(a.d_tab);
  dfree(a.d_T64);
  dfree(a.d_lamA);
  dfree(a.d_lamB);
  dfree(a.d_coef);
  dfree(a.d_code);
  dfree(a.d_gc);
  dfree(a.d_gshot);
  dfree(a.d_seed_node);
  dfree(a.d_seed_ptr);
  dfree(a.d_seed_rec);
  dfree(a.d_seed_w);
  dfree(a.d_picks64);
  dfree(a.d_walk_ctl);
  dfree(a.d_walk_x);
  dfree(a.d_w);
  dfree(a.d_grad);
  delete b;
}

// Enqueue one full solve of the batch (all shots) on stream st.  mark(0) is
// called before the first sweep launch and mark(1) after the last one.
using Mark = std::function<int(int)>;
int enqueue_solve(seisfat_batch_s* b, hipStream_t st, const Mark& mark) {
  seisfat_eik_s* h = b->h;
  const int nzz = h->nzz;
  int launches = 0, sweeps = 0;
  int solve = 0;
  auto one_solve = [&](ShotDev* d_tab, ShotDev* d_tab_b, const std::vector<ShotDev>& tab,
                       const SkewView* skew) -> int {
    if (b->schedule == SEISFAT_SCHED_PLAIN) {
      // time_set + time_init as separate launches, as the reference does; the
      // tiled sweeps generate the initial field inside their first sweep.
      launch_fill(b->d_T, b->tsize, TINF, st);
      launch_init(d_tab, b->nshots, nzz, h->dx, h->dz, st);
      launches += 2;
    }
    if (sweeps == 0)
      if (int rc = mark(0)) return rc;
    int n = 0;
    if (b->schedule == SEISFAT_SCHED_TILED && b->pairs && d_tab_b) {
      n = launch_sweeps_pairs(d_tab, d_tab_b, b->d_ptiles, b->d_qoff, b->nstrips, nzz, b->kc, h->eqh, skew, b->pr,
                              solve, st);
    } else if (b->schedule == SEISFAT_SCHED_TILED) {
      n = launch_sweeps_persist(d_tab, b->d_tiles, b->ntiles, b->nstrips, nzz, b->kc, h->eqh, skew, true, b->pr,
                                solve, st);
    } else if (b->schedule == SEISFAT_SCHED_TILED_DIAG) {
      n = launch_sweeps_tiled(d_tab, b->d_tiles, b->diag_off, nzz, b->kc, h->eqh, skew, true, st);
    } else {
      for (int k = 0; k < b->nshots; ++k) n += launch_sweeps_plain(tab[k], nzz, h->kc, h->eqh, st);
    }
    launches += n;
    sweeps += n;
    ++solve;
    HIPCHK(hipGetLastError());
    return SEISFAT_OK;
  };
  if (b->schedule == SEISFAT_SCHED_TILED) {
    launch_begin_run(b->pr, st);
    ++launches;
  }
  int rc = one_solve(b->d_iso, b->d_iso_b, b->iso, h->skew);
  if (rc) return rc;
  if (b->type == SEISFAT_ANI) {
    StiffDev sf;
    for (int c = 0; c < 6; ++c) {
      sf.c[c] = h->d_c[c];
      sf.mn[c] = h->cmin[c];
      sf.mx[c] = h->cmax[c];
    }
    AniTab tb;
    const size_t plane = (size_t)nzz * h->nxx;
    for (int c = 0; c < 7; ++c) tb.v[c] = h->d_anitab + c * plane;
    // the model as a full-width shot sees it (all anisotropic shots are)
    ShotDev model = b->iso[0];
    launch_ani_table(model, nzz, h->nxx, sf, tb, st);
    launch_quasi_slowness(b->d_iso, b->d_qp, b->nshots, nzz, h->nxx, h->dx, h->dz, tb, st);
    launches += 2;
    rc = one_solve(b->d_qp, b->d_qp_b, b->qp, nullptr);  // per-shot qP slowness: staged in LDS
    if (rc) return rc;
  }
  if ((rc = mark(1))) return rc;
  // Picks read the final traveltimes (the T pointers are shared by both tables).
  launch_picks(b->d_iso, b->d_recshot, b->d_rx, b->d_rz, b->d_picks, b->nrec, nzz, h->dx, h->dz, st);
  if (b->nrec > 0) ++launches;
  HIPCHK(hipGetLastError());
  b->launches = launches;
  b->sweep_launches = sweeps;
  return SEISFAT_OK;
}

// Tiles of every shot, grouped by tile diagonal d = 2r + bb (one launch each).
void build_tiles(seisfat_batch_s* b, std::vector<int4>& tiles) {
  const int NR = b->h->nzz - 2;
  const int NRS = (NR + 63) / 64;
  std::vector<int> NBs(b->nshots);
  int Dmax = 0;
  for (int k = 0; k < b->nshots; ++k) {
    const int NC = b->iso[k].nxx - 2;
    NBs[k] = (NC + 63 + 63) / 64;
    Dmax = std::max(Dmax, 2 * (NRS - 1) + NBs[k]);
  }
  // Scheduling priority of a shot's tiles from its slack: a shot with fewer
  // tile diagonals than the widest one can fall behind by that many tile
  // times without delaying the sweep, so the widest shots' tiles (the
  // critical path) issue first on a shared SIMD.
  std::vector<int> prio(b->nshots);
  for (int k = 0; k < b->nshots; ++k) {
    const int slack = Dmax - (2 * (NRS - 1) + NBs[k]);
    prio[k] = !b->opts.priority ? 0 : slack <= 6 ? 3 : slack <= 16 ? 2 : slack <= 30 ? 1 : 0;
  }
  b->diag_off.assign(1, 0);
  tiles.clear();
  for (int d = 0; d < Dmax; ++d) {
    for (int r = 0; r < NRS; ++r) {
      const int bb = d - 2 * r;
      if (bb < 0) break;
      for (int k = 0; k < b->nshots; ++k)
        if (bb < NBs[k]) tiles.push_back(make_int4(k, r, bb, prio[k]));
    }
    b->diag_off.push_back((int)tiles.size());
  }
}

// Tiles of both sweeps of a fused pair (rows down, then rows up; same column
// direction), shots spread over the 8 XCDs by tile count, each XCD's tiles in
// order of their level in the dependency graph (first sweep: diagonal 2r+bb;
// second sweep: one more than the latest of its own predecessors and of the
// first-sweep tiles under it), so every tile follows all tiles it waits for.
// The first-sweep tiles under second-sweep tile (r, bb) follow the same
// ranges as the kernel's wait (tile_process, XMODE 2).
void build_pair_tiles(seisfat_batch_s* b, std::vector<int4>& pt, std::vector<int>& qoff) {
  const int NR = b->h->nzz - 2;
  const int NRS = (NR + 63) / 64;
  const int n = b->nshots;
  std::vector<int> NBs(n), prio(n), xcd(n), order(n);
  int Dmax = 0;
  for (int k = 0; k < n; ++k) {
    NBs[k] = (b->iso[k].nxx - 2 + 63 + 63) / 64;
    Dmax = std::max(Dmax, 2 * (NRS - 1) + NBs[k]);
  }
  for (int k = 0; k < n; ++k) {
    const int slack = Dmax - (2 * (NRS - 1) + NBs[k]);
    prio[k] = !b->opts.priority ? 0 : slack <= 6 ? 3 : slack <= 16 ? 2 : slack <= 30 ? 1 : 0;
    order[k] = k;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return NBs[a] > NBs[c]; });
  long long load[8] = {};
  for (int k : order) {
    const int x = (int)(std::min_element(load, load + 8) - load);
    xcd[k] = x;
    load[x] += NBs[k];
  }
  struct E {
    int lvl, which, k, r, bb;
  };
  std::vector<std::vector<E>> per(8);
  for (int k = 0; k < n; ++k) {
    const int nb = NBs[k];
    std::vector<int> Lb((size_t)NRS * nb, 0);
    for (int r = 0; r < NRS; ++r)
      for (int bb = 0; bb < nb; ++bb) {
        int L = 0;
        if (bb > 0) L = std::max(L, Lb[(size_t)r * nb + bb - 1] + 1);
        if (r > 0) L = std::max(L, Lb[(size_t)(r - 1) * nb + std::min(bb + 1, nb - 1)] + 1);
        const int c0 = 64 * bb;
        const int kk1_hi = std::min(64 * r + 63, NR - 1);
        const int kk0_lo = NR - 1 - kk1_hi, kk0_hi = NR - 1 - 64 * r;
        for (int r0 = kk0_lo >> 6; r0 <= (kk0_hi >> 6); ++r0) {
          const int klo = NR - 1 - 64 * r - std::min(kk0_hi, 64 * r0 + 63);
          const int khi = NR - 1 - 64 * r - std::max(kk0_lo, 64 * r0);
          const int C = NR - 1 - 64 * (r + r0);
          const int blo = std::max((c0 - 2 * khi + C) >> 6, 0);
          const int bhi = std::min((c0 + 63 - 2 * klo + C) >> 6, nb - 1);
          if (blo <= bhi) L = std::max(L, 2 * r0 + bhi + 1);
        }
        Lb[(size_t)r * nb + bb] = L;
        per[xcd[k]].push_back(E{2 * r + bb, 0, k, r, bb});
        per[xcd[k]].push_back(E{L, 1, k, r, bb});
      }
  }
  pt.clear();
  qoff.assign(1, 0);
  for (int x = 0; x < 8; ++x) {
    std::sort(per[x].begin(), per[x].end(), [](const E& a, const E& c) {
      if (a.lvl != c.lvl) return a.lvl < c.lvl;
      if (a.which != c.which) return a.which < c.which;
      if (a.k != c.k) return a.k < c.k;
      if (a.r != c.r) return a.r < c.r;
      return a.bb < c.bb;
    });
    for (const E& e : per[x]) pt.push_back(make_int4(e.k, e.r, e.bb, prio[e.k] | (e.which << 2)));
    qoff.push_back((int)pt.size());
  }
}

// Whether a launch's workgroups go round robin over 8 XCCs (checked once).
int check_xcc(seisfat_eik_s* h) {
  if (h->xcc_ok >= 0) return SEISFAT_OK;
  // SEISFAT_SWEEP_PAIRS=0 (read when a handle's first tiled batch is
  // created): one persistent launch per sweep instead of the fused pairs
  if (const char* env = std::getenv("SEISFAT_SWEEP_PAIRS")) {
    if (env[0] == '0') {
      h->xcc_ok = 0;
      return SEISFAT_OK;
    }
  }
  const int n = 4096;
  int* d = nullptr;
  if (int rc = dalloc(&d, (size_t)n)) return rc;
  launch_xcc_ids(d, n, h->stream);
  std::vector<int> ids(n);
  hipError_t e = hipMemcpyAsync(ids.data(), d, n * sizeof(int), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  dfree(d);
  if (e != hipSuccess) return fail(SEISFAT_EHIP, hipGetErrorString(e));
  // round robin over 8 XCCs from wherever the dispatcher stands: every XCC
  // then receives workgroups of any launch of at least 8 (each workgroup
  // reads its own XCC id, so the starting point does not matter)
  h->xcc_ok = ids[0] >= 0 && ids[0] < 8;
  for (int i = 0; i < n; ++i)
    if (ids[i] != (ids[0] + i) % 8) h->xcc_ok = 0;
  return SEISFAT_OK;
}

// Checks a batch's shots and receivers against the model; returns, per
// receiver, its shot and its position relative to the shot's grid origin.
int validate_shots(const seisfat_eik_s* h, int type, int nshots, const seisfat_shot* shots, int nrec_total,
                   const float* rx, const float* rz, std::vector<int>& recshot, std::vector<float>& rxr,
                   std::vector<float>& rzr) {
  if (nrec_total < 0 || (nrec_total > 0 && (!rx || !rz))) return fail(SEISFAT_EINVAL, "bad receiver arrays");
  const float dx = h->dx, dz = h->dz;
  recshot.assign(nrec_total, -1);
  rxr.assign(nrec_total, 0.0f);
  rzr.assign(nrec_total, 0.0f);
  for (int k = 0; k < nshots; ++k) {
    const seisfat_shot& s = shots[k];
    const std::string tag = "shot " + std::to_string(k) + ": ";
    if (s.width < 1 || s.x0 < 0 || (long long)s.x0 + s.width > h->nx)
      return fail(SEISFAT_EINVAL, tag + "window [x0, x0+width) outside the model");
    if (type == SEISFAT_ANI && (s.x0 != 0 || s.width != h->nx))
      return fail(SEISFAT_EINVAL, tag + "anisotropic shots must cover the whole model");
    const float sxr = s.sx - (float)s.x0 * dx;
    if (!(sxr >= 0.0f && sxr <= (float)(s.width - 1) * dx && s.sz >= 0.0f && s.sz <= (float)(h->nz - 1) * dz))
      return fail(SEISFAT_EINVAL, tag + "source outside its window");
    if (s.nrec < 0 || s.rec_offset < 0 || (long long)s.rec_offset + s.nrec > nrec_total)
      return fail(SEISFAT_EINVAL, tag + "receiver range outside the receiver arrays");
    for (int r = s.rec_offset; r < s.rec_offset + s.nrec; ++r) {
      if (recshot[r] != -1) return fail(SEISFAT_EINVAL, "receiver " + std::to_string(r) + " assigned to two shots");
      const float xr = rx[r] - (float)s.x0 * dx;
      if (!(xr >= 0.0f && xr <= (float)(s.width - 1) * dx && rz[r] >= 0.0f && rz[r] <= (float)(h->nz - 1) * dz))
        return fail(SEISFAT_EINVAL, tag + "receiver " + std::to_string(r) + " outside the shot window");
      recshot[r] = k;
      rxr[r] = xr;
      rzr[r] = rz[r];
    }
  }
  for (int r = 0; r < nrec_total; ++r)
    if (recshot[r] < 0) return fail(SEISFAT_EINVAL, "receiver " + std::to_string(r) + " belongs to no shot");

  return SEISFAT_OK;
}

int create_batch(seisfat_eik_t h, int type, int schedule, int nshots, const seisfat_shot* shots,
                 int nrec_total, const float* rx, const float* rz, bool use_graph, seisfat_batch_s** out,
                 const seisfat_sched_opts* opts = nullptr) {
  *out = nullptr;
  const seisfat_sched_opts o = opts ? *opts : seisfat_sched_opts{1, 0, 1, 1};
  if ((o.pairs != 0 && o.pairs != 1) || (o.skip != 0 && o.skip != 1) || (o.priority != 0 && o.priority != 1) ||
      o.wpcu < 0 || o.wpcu > 64)
    return fail(SEISFAT_EINVAL, "schedule options: pairs, skip and priority take 0 or 1, wpcu 0..64");
  if (type != SEISFAT_ISO && type != SEISFAT_ANI) return fail(SEISFAT_EINVAL, "type must be ISO (0) or ANI (1)");
  if (schedule != SEISFAT_SCHED_PLAIN && schedule != SEISFAT_SCHED_TILED && schedule != SEISFAT_SCHED_TILED_DIAG)
    return fail(SEISFAT_EINVAL, "schedule must be PLAIN (0), TILED (1) or TILED_DIAG (2)");
  if (!h->has_s) return fail(SEISFAT_ESTATE, "slowness not set (seisfat_eik_set_slowness)");
  if (type == SEISFAT_ANI && !h->has_c) return fail(SEISFAT_ESTATE, "stiffness not set (seisfat_eik_set_stiffness)");
  if (nshots <= 0 || !shots) return fail(SEISFAT_EINVAL, "nshots must be positive");
  if (type == SEISFAT_ANI && (size_t)h->nzz * h->nxx >= ((size_t)1 << 31))
    return fail(SEISFAT_EINVAL, "anisotropic batches take models of fewer than 2^31 padded nodes");
  const float dx = h->dx;
  std::vector<int> recshot;
  std::vector<float> rxr, rzr;
  if (int rc = validate_shots(h, type, nshots, shots, nrec_total, rx, rz, recshot, rxr, rzr)) return rc;
  seisfat_batch_s* b = new (std::nothrow) seisfat_batch_s();
  if (!b) return fail(SEISFAT_ENOMEM, "host allocation");
  b->h = h;
  b->type = type;
  b->schedule = schedule;
  b->opts = o;
  b->kc = h->kc;
  b->kc.tinf_fixed = h->kc.tinf_fixed && o.skip;
  b->nshots = nshots;
  b->nrec = nrec_total;
  b->shots.assign(shots, shots + nshots);
  b->rxr = rxr;
  b->rzr = rzr;
  b->recshot = recshot;
  b->iso.resize(nshots);
  b->toff.resize(nshots);
  const int nzz = h->nzz;
  size_t off = 0;
  for (int k = 0; k < nshots; ++k) {
    const seisfat_shot& s = shots[k];
    b->toff[k] = off;
    ShotDev& sd = b->iso[k];
    sd.nxx = s.width + 2 * NB;
    if (s.x0 == 0 && s.width == h->nx) {
      // the whole model: its padding as uploaded (the reference's d_S)
      sd.sbase = h->d_spad;
      sd.sofs = 0;
      sd.smax = h->nxx - 1;
    } else {
      // a window: its own replicated x border around model columns [x0, x0+width)
      sd.sbase = h->d_spad + (size_t)(s.x0 + NB) * nzz;
      sd.sofs = NB;
      sd.smax = s.width - 1;
    }
    sd.sx = s.sx - (float)s.x0 * dx;
    sd.sz = s.sz;
    sd.x0 = s.x0;
    sd.w = s.width;
    off += (size_t)nzz * sd.nxx;
  }
  b->tsize = off;
  int rc;
#define HIPCHK_B(expr)                                                            \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      free_batch(b);                                                              \
      return fail(SEISFAT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    }                                                                             \
  } while (0)
#define TRY(x)        \
  do {                \
    rc = (x);         \
    if (rc) {         \
      free_batch(b);  \
      return rc;      \
    }                 \
  } while (0)
  for (auto& ev : b->ev_sweep)
    if (hipEventCreate(&ev) != hipSuccess) {
      free_batch(b);
      return fail(SEISFAT_EHIP, "hipEventCreate");
    }
  if (hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) != hipSuccess) {
    free_batch(b);
    return fail(SEISFAT_EHIP, "hipEventCreate");
  }
  TRY(dalloc(&b->d_T, b->tsize));
  for (int k = 0; k < nshots; ++k) b->iso[k].T = b->d_T + b->toff[k];
  if (type == SEISFAT_ANI) {
    TRY(dalloc(&b->d_swork, b->tsize));
    b->qp = b->iso;  // exchange pointers are copied below once assigned
    for (int k = 0; k < nshots; ++k) {
      b->qp[k].sbase = b->d_swork + b->toff[k];
      b->qp[k].sofs = 0;
      b->qp[k].smax = b->qp[k].nxx - 1;
    }
    TRY(dalloc(&b->d_qp, nshots));
  }
  TRY(dalloc(&b->d_iso, nshots));
  TRY(dalloc(&b->d_rx, nrec_total));
  TRY(dalloc(&b->d_rz, nrec_total));
  TRY(dalloc(&b->d_recshot, nrec_total));
  TRY(dalloc(&b->d_picks, nrec_total));
  std::vector<int4> tiles;
  const bool tiled = schedule != SEISFAT_SCHED_PLAIN;
  if (tiled) {
    build_tiles(b, tiles);
    b->ntiles = (int)tiles.size();
    b->nstrips = (nzz - 2 + 63) / 64;
    TRY(dalloc(&b->d_tiles, tiles.size()));
  }
  if (schedule == SEISFAT_SCHED_TILED) {
    TRY(check_xcc(h));
    b->pairs = h->xcc_ok == 1 && b->opts.pairs;
    // exchange words per shot: bottom row of each strip + double-buffered
    // carries (a second set for the fused pairs' second sweeps), then the
    // pairs' done words (one per tile)
    std::vector<size_t> xoff(nshots), doff(nshots);
    size_t words = 0, dwords = 0;
    for (int k = 0; k < nshots; ++k) {
      xoff[k] = words;
      const int ldh = ((b->iso[k].nxx - 2) + 63) / 64 * 64;
      words += (size_t)b->nstrips * ldh + (size_t)b->nstrips * 2 * 128;
      doff[k] = dwords;
      dwords += (size_t)b->nstrips * ((b->iso[k].nxx - 2 + 126) / 64);
    }
    const size_t total = b->pairs ? 2 * words + dwords : words;
    b->xch_words = total;
    TRY(dalloc(&b->d_xch, total));
    TRY(dalloc(&b->d_ctl, CTL_WORDS));
    for (int k = 0; k < nshots; ++k) {
      const int ldh = ((b->iso[k].nxx - 2) + 63) / 64 * 64;
      b->iso[k].xh = b->d_xch + xoff[k];
      b->iso[k].xc = b->d_xch + xoff[k] + (size_t)b->nstrips * ldh;
      b->iso[k].ldh = ldh;
      if (b->pairs) {
        b->iso[k].xd = b->d_xch + 2 * words + doff[k];
        b->iso[k].nb = (b->iso[k].nxx - 2 + 126) / 64;
      }
    }
    if (b->pairs) {
      b->iso_b = b->iso;
      for (int k = 0; k < nshots; ++k) {
        b->iso_b[k].xh = b->iso[k].xh + words;
        b->iso_b[k].xc = b->iso[k].xc + words;
      }
      std::vector<int4> pt;
      std::vector<int> qoff;
      build_pair_tiles(b, pt, qoff);
      TRY(dalloc(&b->d_iso_b, nshots));
      TRY(dalloc(&b->d_ptiles, pt.size()));
      TRY(dalloc(&b->d_qoff, qoff.size()));
      HIPCHK_B(hipMemcpyAsync(b->d_iso_b, b->iso_b.data(), nshots * sizeof(ShotDev), hipMemcpyHostToDevice,
                              h->stream));
      HIPCHK_B(hipMemcpyAsync(b->d_ptiles, pt.data(), pt.size() * sizeof(int4), hipMemcpyHostToDevice, h->stream));
      HIPCHK_B(hipMemcpyAsync(b->d_qoff, qoff.data(), qoff.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
      // the host copies are read back only after the stream synchronises below
      b->pair_tiles = std::move(pt);
      b->pair_qoff = std::move(qoff);
    }
    // the qP solve (anisotropic batches) runs after the isotropic one and
    // reuses its exchange and done words under its own sweep ids
    for (int k = 0; k < (int)b->qp.size(); ++k) {
      b->qp[k].xh = b->iso[k].xh;
      b->qp[k].xc = b->iso[k].xc;
      b->qp[k].ldh = b->iso[k].ldh;
      b->qp[k].xd = b->iso[k].xd;
      b->qp[k].nb = b->iso[k].nb;
    }
    if (b->pairs && !b->qp.empty()) {
      b->qp_b = b->qp;
      for (int k = 0; k < nshots; ++k) {
        b->qp_b[k].xh = b->iso_b[k].xh;
        b->qp_b[k].xc = b->iso_b[k].xc;
      }
    }
    b->pr.counters = b->d_ctl;
    b->pr.wpcu = o.wpcu;
    b->pr.epoch = (unsigned*)(b->d_ctl + CTL_EPOCH);
    b->pr.err = b->d_ctl + CTL_ERR;
    // all-ones words carry a tag no run ever uses; epoch and flags start at 0
    HIPCHK_B(hipMemsetAsync(b->d_xch, 0xff, total * sizeof(unsigned long long), h->stream));
    HIPCHK_B(hipMemsetAsync(b->d_ctl, 0, CTL_WORDS * sizeof(int), h->stream));
  }
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return SEISFAT_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    return SEISFAT_OK;
  };
  TRY(up(b->d_iso, b->iso.data(), nshots * sizeof(ShotDev)));
  if (type == SEISFAT_ANI) TRY(up(b->d_qp, b->qp.data(), nshots * sizeof(ShotDev)));
  if (!b->qp_b.empty()) {
    TRY(dalloc(&b->d_qp_b, nshots));
    TRY(up(b->d_qp_b, b->qp_b.data(), nshots * sizeof(ShotDev)));
  }
  TRY(up(b->d_rx, rxr.data(), nrec_total * sizeof(float)));
  TRY(up(b->d_rz, rzr.data(), nrec_total * sizeof(float)));
  TRY(up(b->d_recshot, recshot.data(), nrec_total * sizeof(int)));
  TRY(up(b->d_tiles, tiles.data(), tiles.size() * sizeof(int4)));
  // The boundary ring of every grid holds 1e6 for good (sweeps never write
  // it); the tiled schedule relies on this one fill instead of a per-run one.
  launch_fill(b->d_T, b->tsize, TINF, h->stream);
  {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
      free_batch(b);
      return fail(SEISFAT_EHIP, std::string("upload: ") + hipGetErrorString(e));
    }
  }
  if (use_graph && schedule != SEISFAT_SCHED_PLAIN) {
    hipStream_t st = h->stream;
    int part = 0;
    auto end_part = [&]() -> int {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamEndCapture(st, &g));
      b->graph[part] = g;
      HIPCHK(hipGraphInstantiate(&b->gexec[part], g, nullptr, nullptr, 0));
      ++part;
      return SEISFAT_OK;
    };
    auto begin_part = [&]() -> int {
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      return SEISFAT_OK;
    };
    rc = begin_part();
    if (!rc) {
      rc = enqueue_solve(b, st, [&](int) -> int {
        int r = end_part();
        return r ? r : begin_part();
      });
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      (void)hipStreamIsCapturing(st, &cs);
      if (cs == hipStreamCaptureStatusActive) {
        const int r = end_part();
        if (!rc) rc = r;
      }
    }
    if (!rc && part != 3) rc = fail(SEISFAT_EHIP, "graph capture produced an unexpected number of parts");
    if (rc) {
      free_batch(b);
      return rc;
    }
  }
  *out = b;
  return SEISFAT_OK;
#undef TRY
#undef HIPCHK_B
}

// Wait for the batch's last enqueued work (only that: other streams' work is
// not waited for) and turn a device-side failure flag into an error.
int batch_status(seisfat_batch_s* b) {
  HIPCHK(hipEventSynchronize(b->ev_done));
  hipStream_t st = b->h->stream;
  int err[2] = {0, 0};
  if (b->d_ctl) HIPCHK(hipMemcpyAsync(&err[0], b->d_ctl + CTL_ERR, sizeof(int), hipMemcpyDeviceToHost, st));
  if (b->adj.d_walk_ctl)
    HIPCHK(hipMemcpyAsync(&err[1], b->adj.d_walk_ctl + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  // Fused pairs: every XCD drains its own tile queue, so each queue needs at
  // least one of the launch's workgroups on its XCD (the round robin checked
  // when the geometry was made; a kernel running at the same time on another
  // stream could take an XCD's whole share).  A queue's ticket counter ends
  // at its length plus the workgroups that drained it, or at 0 if none did.
  int cnt[NUM_COUNTERS] = {};
  const bool queues = b->pairs && b->last_run_plain && b->d_ctl && b->pair_qoff.size() == 9;
  if (queues) HIPCHK(hipMemcpyAsync(cnt, b->d_ctl, sizeof(cnt), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (err[0]) return fail(SEISFAT_EHIP, "persistent sweep: a tile dependency did not arrive within the spin limit");
  if (err[1])
    return fail(SEISFAT_EHIP, "adjoint strip walk (fp64 trajectory or reverse sweep): a strip's upper "
                              "neighbour did not arrive within the spin limit");
  if (queues) {
    const int solves = b->type == SEISFAT_ANI ? 2 : 1;
    for (int s = 0; s < solves; ++s)
      for (int p = 0; p < 2; ++p)
        for (int x = 0; x < 8; ++x)
          if (cnt[PAIR_COUNTERS + 16 * s + 8 * p + x] < b->pair_qoff[x + 1] - b->pair_qoff[x])
            return fail(SEISFAT_EHIP, "fused sweep pairs: XCD " + std::to_string(x) +
                                          " received no workgroup, its tiles were not solved (another kernel "
                                          "running at the same time?); use schedule option pairs = 0");
  }
  return SEISFAT_OK;
}

int run_batch(seisfat_batch_s* b, hipStream_t st) {
  if (!b->h->has_s) return fail(SEISFAT_ESTATE, "slowness rejected or not set since the batch was created");
  b->last_run_plain = true;
  if (b->gexec[0]) {
    HIPCHK(hipGraphLaunch(b->gexec[0], st));
    HIPCHK(hipEventRecord(b->ev_sweep[0], st));
    HIPCHK(hipGraphLaunch(b->gexec[1], st));
    HIPCHK(hipEventRecord(b->ev_sweep[1], st));
    HIPCHK(hipGraphLaunch(b->gexec[2], st));
  } else {
    if (int rc = enqueue_solve(b, st, [&](int i) -> int {
          HIPCHK(hipEventRecord(b->ev_sweep[i], st));
          return SEISFAT_OK;
        }))
      return rc;
  }
  HIPCHK(hipEventRecord(b->ev_done, st));
  return SEISFAT_OK;
}

// Converged mode: passes of the four sweeps until one leaves every
// traveltime unchanged (at most max_passes).  Three graphs: the first pass
// (initialisation fused into its first sweep), `chunk` gated passes (each
// skipped at once after convergence), and the picks; the host reads the skip
// word after each chunk.
int run_converged(seisfat_batch_s* b, int max_passes, int chunk, int* passes, hipStream_t st) {
  b->last_run_plain = false;  // gated passes leave their queues untouched
  seisfat_eik_s* h = b->h;
  if (b->type != SEISFAT_ISO || b->schedule != SEISFAT_SCHED_TILED)
    return fail(SEISFAT_EINVAL, "converged mode covers isotropic batches of the tiled schedule");
  if (max_passes < 1 || max_passes > (1 << 20)) return fail(SEISFAT_EINVAL, "max_passes must be in [1, 2^20]");
  if (chunk < 2 || chunk > 256) return fail(SEISFAT_EINVAL, "chunk must be in [2, 256]");
  if (!h->has_s) return fail(SEISFAT_ESTATE, "slowness rejected or not set since the batch was created");
  chunk += chunk & 1;  // even: the change words alternate by pass parity
  if (!b->d_conv)
    if (int rc = dalloc(&b->d_conv, (size_t)CONV_WORDS)) return rc;
  if (b->conv_chunk != chunk || b->conv_max != max_passes) {
    for (auto& g : b->conv_exec)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto& g : b->conv_graph)
      if (g) (void)hipGraphDestroy(g);
    for (int q = 0; q < 3; ++q) {
      b->conv_exec[q] = nullptr;
      b->conv_graph[q] = nullptr;
    }
    hipStream_t cs = h->stream;
    const int nzz = h->nzz;
    for (int part = 0; part < 3; ++part) {
      HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      if (part == 0) {
        launch_conv_init(b->d_conv, max_passes, cs);
        launch_begin_run(b->pr, cs);
        launch_sweeps_persist(b->d_iso, b->d_tiles, b->ntiles, b->nstrips, nzz, b->kc, h->eqh, h->skew, true, b->pr,
                              0, cs, b->d_conv + 2, b->d_conv);
      } else if (part == 1) {
        for (int q = 0; q < chunk; ++q) {  // passes 1 + q (+ a multiple of chunk)
          launch_conv_gate(b->d_conv, cs);
          launch_begin_run(b->pr, cs);
          launch_sweeps_persist(b->d_iso, b->d_tiles, b->ntiles, b->nstrips, nzz, b->kc, h->eqh, h->skew, false,
                                b->pr, 0, cs, b->d_conv + 2 + ((1 + q) & 1), b->d_conv);
        }
      } else {
        launch_picks(b->d_iso, b->d_recshot, b->d_rx, b->d_rz, b->d_picks, b->nrec, nzz, h->dx, h->dz, cs);
        // one more node, so the graph is not empty when the batch has no receivers
        launch_fill(reinterpret_cast<float*>(b->d_conv + 6), 1, 0.0f, cs);
      }
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(cs, &g);
      if (e != hipSuccess) return fail(SEISFAT_EHIP, std::string("converged-mode capture: ") + hipGetErrorString(e));
      b->conv_graph[part] = g;
      HIPCHK(hipGraphInstantiate(&b->conv_exec[part], g, nullptr, nullptr, 0));
    }
    b->conv_chunk = chunk;
    b->conv_max = max_passes;
  }
  HIPCHK(hipGraphLaunch(b->conv_exec[0], st));
  int cw[2] = {0, 1};
  for (int launched = 1; launched < max_passes; launched += chunk) {
    HIPCHK(hipGraphLaunch(b->conv_exec[1], st));
    HIPCHK(hipMemcpyAsync(cw, b->d_conv, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (cw[0]) break;
  }
  HIPCHK(hipGraphLaunch(b->conv_exec[2], st));
  HIPCHK(hipMemcpyAsync(cw, b->d_conv, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(b->ev_done, st));
  HIPCHK(hipStreamSynchronize(st));
  if (passes) *passes = cw[1];
  return batch_status(b);
}

// ---------------------------------------------------------------------------
// Adjoint.  Layout of the per-batch state: see eik_adjoint.h.
void cubic_weights(double d, double w[4]) {
  // Catmull-Rom weights of cubic1d (src/admin/admin.cpp:121-127): the pick is
  // linear in its 16 stencil values with weights wx[pIdx] * wz[pIdz].
  w[0] = 0.5 * d * (-1.0 + d * (2.0 - d));
  w[1] = 1.0 + 0.5 * d * d * (-5.0 + 3.0 * d);
  w[2] = 0.5 * d * (1.0 + d * (4.0 - 3.0 * d));
  w[3] = 0.5 * d * d * (-1.0 + d);
}

int adj_prepare(seisfat_batch_s* b) {
  AdjState& A = b->adj;
  if (A.ready) return SEISFAT_OK;
  seisfat_eik_s* h = b->h;
  const int nzz = h->nzz, nshots = b->nshots;
  A.nstr = (nzz + 63) / 64;
  A.tab.resize(nshots);
  size_t koff = 0;
  A.max_nxx = 0;
  for (int k = 0; k < nshots; ++k) {
    AdjShot& s = A.tab[k];
    s.sd = b->iso[k];
    s.toff = b->toff[k];
    const int nxx = s.sd.nxx;
    s.NTP = (nxx + 63 + ADJ_PF - 1) / ADJ_PF * ADJ_PF;
    s.koff = koff;
    koff += (size_t)A.nstr * s.NTP * 64;
    A.max_nxx = std::max(A.max_nxx, nxx);
  }
  A.ksize = koff;
  // pick seeds: (skewed node of the last reverse sweep, receiver, weight),
  // grouped by node with receivers in order
  struct Seed {
    size_t node;
    int rec;
    float w;
  };
  std::vector<Seed> es;
  es.reserve((size_t)b->nrec * 16);
  const float dx = h->dx, dz = h->dz;
  for (int r = 0; r < b->nrec; ++r) {
    const int k = b->recshot[r];
    const AdjShot& s = A.tab[k];
    const float x = b->rxr[r], z = b->rzr[r];
    const float x0 = std::floor((x + 0.5f * dx) / dx) * dx;
    const float z0 = std::floor((z + 0.5f * dz) / dz) * dz;
    const float x1 = std::floor((x + 0.5f * dx) / dx) * dx + dx;
    const float z1 = std::floor((z + 0.5f * dz) / dz) * dz + dz;
    const float xd = (x - x0) / (x1 - x0);
    const float zd = (z - z0) / (z1 - z0);
    const int i = (int)((z + 0.5f * dz) / dz) + NB;
    const int j = (int)((x + 0.5f * dx) / dx) + NB;
    double wx[4], wz[4];
    cubic_weights(xd, wx);
    cubic_weights(zd, wz);
    for (int pIdx = 0; pIdx < 4; ++pIdx)
      for (int pIdz = 0; pIdz < 4; ++pIdz) {
        const int ni = i + pIdz - 1, nj = j + pIdx - 1;
        const size_t node =
            s.koff + adj_kidx(adj_ip(ni, adj_a(3), nzz), adj_jp(nj, adj_b(3), s.sd.nxx), s.NTP);
        es.push_back(Seed{node, r, (float)(wx[pIdx] * wz[pIdz])});
      }
  }
  std::stable_sort(es.begin(), es.end(), [](const Seed& p, const Seed& q) { return p.node < q.node; });
  std::vector<size_t> snode;
  std::vector<int> sptr, srec;
  std::vector<float> sw;
  for (size_t e = 0; e < es.size(); ++e) {
    if (e == 0 || es[e].node != es[e - 1].node) {
      snode.push_back(es[e].node);
      sptr.push_back((int)e);
    }
    srec.push_back(es[e].rec);
    sw.push_back(es[e].w);
  }
  sptr.push_back((int)es.size());
  A.nseed = (int)snode.size();

  int rc;
  const size_t tsize = b->tsize;
  if ((rc = dalloc(&A.d_tab, (size_t)nshots)) || (rc = dalloc(&A.d_T64, 5 * tsize)) ||
      (rc = dalloc(&A.d_lamA, A.ksize)) || (rc = dalloc(&A.d_lamB, A.ksize)) || (rc = dalloc(&A.d_coef, A.ksize)) ||
      (rc = dalloc(&A.d_code, A.ksize)) || (rc = dalloc(&A.d_gc, A.ksize)) || (rc = dalloc(&A.d_gshot, tsize)) ||
      (rc = dalloc(&A.d_seed_node, snode.size())) ||
      (rc = dalloc(&A.d_seed_ptr, sptr.size())) || (rc = dalloc(&A.d_seed_rec, srec.size())) ||
      (rc = dalloc(&A.d_seed_w, sw.size())) || (rc = dalloc(&A.d_w, (size_t)b->nrec)) || (rc = dalloc(&A.d_picks64, (size_t)b->nrec)) ||
      (rc = dalloc(&A.d_grad, (size_t)h->nz * h->nx)) || (rc = dalloc(&A.d_walk_ctl, 2)))
    return rc;
  A.walk_ldx = 2 * A.max_nxx;
  if ((rc = dalloc(&A.d_walk_x, (size_t)nshots * A.nstr * A.walk_ldx))) return rc;
  // exchange words start with tag 0, below every launch's tag (>= 1)
  HIPCHK(hipMemsetAsync(A.d_walk_x, 0, (size_t)nshots * A.nstr * A.walk_ldx * sizeof(unsigned long long), h->stream));
  HIPCHK(hipMemsetAsync(A.d_walk_ctl, 0, 2 * sizeof(int), h->stream));
  hipStream_t st = h->stream;
  HIPCHK(hipMemcpyAsync(A.d_tab, A.tab.data(), nshots * sizeof(AdjShot), hipMemcpyHostToDevice, st));
  if (!snode.empty()) {
    HIPCHK(hipMemcpyAsync(A.d_seed_node, snode.data(), snode.size() * sizeof(size_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(A.d_seed_ptr, sptr.data(), sptr.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(A.d_seed_rec, srec.data(), srec.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(A.d_seed_w, sw.data(), sw.size() * sizeof(float), hipMemcpyHostToDevice, st));
  }
  // Entries of the skewed arrays outside every grid are never written and
  // must read as zero.
  HIPCHK(hipMemsetAsync(A.d_lamA, 0, A.ksize * sizeof(float), st));
  HIPCHK(hipMemsetAsync(A.d_lamB, 0, A.ksize * sizeof(float), st));
  HIPCHK(hipMemsetAsync(A.d_coef, 0, A.ksize * sizeof(float4), st));
  HIPCHK(hipMemsetAsync(A.d_code, 0, A.ksize, st));
  HIPCHK(hipMemsetAsync(A.d_gc, 0, A.ksize * sizeof(float), st));
  // ring nodes of every trajectory grid, and T0 off the source stencil, stay 1e6
  launch_fill_d(A.d_T64, 5 * tsize, 1e6, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  A.ready = true;
  return SEISFAT_OK;
}

// The batch's forward solve, then the fp64 trajectory the adjoint linearises
// (T0 = initial field, T_{s+1} = sweep s of T_s).
int run_record(seisfat_batch_s* b, hipStream_t st) {
  if (b->type != SEISFAT_ISO) return fail(SEISFAT_EINVAL, "the adjoint covers isotropic batches");

  if (int rc = adj_prepare(b)) return rc;
  AdjState& A = b->adj;
  A.recorded = false;
  if (int rc = run_batch(b, st)) return rc;
  seisfat_eik_s* h = b->h;
  const size_t ts = b->tsize;
  launch_fwd64_init(A.d_tab, b->nshots, h->nzz, h->dx, h->dz, A.d_T64, st);
  for (int s = 0; s < 4; ++s)
    launch_fwd64_sweep(s, A.d_tab, b->nshots, h->nzz, A.nstr, h->dx, h->dz, A.d_T64 + s * ts,
                       A.d_T64 + (s + 1) * ts, A.d_walk_ctl, A.d_walk_x, A.walk_ldx,
                       A.walk_tag = A.walk_tag == 0xffffffffu ? 1 : A.walk_tag + 1, st);
  launch_picks64(A.d_tab, b->d_recshot, b->d_rx, b->d_rz, b->nrec, h->nzz, h->dx, h->dz, A.d_T64 + 4 * ts,
                 A.d_picks64, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->ev_done, st));
  A.recorded = true;
  A.version = h->version;
  A.pad_replicated = h->pad_replicated;
  return SEISFAT_OK;
}

int enqueue_vjp(seisfat_batch_s* b, const float* d_w, double* d_grad, hipStream_t st) {
  AdjState& A = b->adj;
  seisfat_eik_s* h = b->h;
  if (!A.recorded) return fail(SEISFAT_ESTATE, "no recorded run (seisfat_batch_run_record)");
  if (A.version != h->version) return fail(SEISFAT_ESTATE, "the slowness changed since the recorded run");
  if (!d_grad) return fail(SEISFAT_EINVAL, "null gradient pointer");
  if (b->nrec > 0 && !d_w) return fail(SEISFAT_EINVAL, "null cotangent pointer");
  const int nzz = h->nzz, ns = b->nshots;
  launch_fill(A.d_lamB, A.ksize, 0.0f, st);
  launch_adj_seed(A.d_seed_node, A.d_seed_ptr, A.d_seed_rec, A.d_seed_w, A.nseed, d_w, A.d_lamB, st);
  launch_fill(A.d_gshot, b->tsize, 0.0f, st);
  const size_t ts = b->tsize;
  for (int s = 3; s >= 0; --s) {
    launch_adj_coef(s, A.d_tab, ns, A.nstr, A.max_nxx, nzz, h->dx, h->dz, A.d_T64 + (s + 1) * ts,
                    A.d_T64 + s * ts, s == 3 ? nullptr : A.d_lamA, A.d_lamB, A.d_coef, A.d_code, st);
    launch_adj_sweep(A.d_tab, ns, A.nstr, A.d_lamB, A.d_coef, A.d_lamA, A.d_gc, A.d_walk_ctl, A.d_walk_x,
                     A.walk_ldx, A.walk_tag = A.walk_tag == 0xffffffffu ? 1 : A.walk_tag + 1, st);
    launch_adj_fold(s, A.d_tab, ns, A.nstr, A.max_nxx, nzz, A.d_gc, A.d_code, A.d_gshot, st);
  }
  launch_adj_source(A.d_tab, ns, nzz, h->dx, h->dz, A.d_lamA, A.d_gshot, st);
  launch_adj_reduce(A.d_tab, ns, h->nz, h->nx, nzz, A.d_gshot, d_grad, !A.pad_replicated, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(b->ev_done, st));
  return SEISFAT_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* seisfat_last_error(void) { return g_err.c_str(); }

int seisfat_eik_create(int nz, int nx, float dz, float dx, seisfat_eik_t* out) {
  if (!out) return fail(SEISFAT_EINVAL, "null output pointer");
  *out = nullptr;
  if (nz < 2 || nx < 2) return fail(SEISFAT_EINVAL, "model needs at least 2x2 samples");
  if ((long long)(nz + 6) * (nx + 6) > (1LL << 30)) return fail(SEISFAT_EINVAL, "model too large (> 2^30 padded nodes)");
  if (!(dz >= 1e-6f && dz <= 1e6f) || !(dx >= 1e-6f && dx <= 1e6f))
    return fail(SEISFAT_EINVAL, "spacings must lie in [1e-6, 1e6]");
  seisfat_eik_s* h = new (std::nothrow) seisfat_eik_s();
  if (!h) return fail(SEISFAT_ENOMEM, "host allocation");
  h->nz = nz;
  h->nx = nx;
  h->nzz = nz + 2 * NB;
  h->nxx = nx + 2 * NB;
  h->dz = dz;
  h->dx = dx;
  h->kc = make_consts(dx, dz);
  h->eqh = (dx == dz);
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(SEISFAT_EHIP, std::string("device/stream: ") + hipGetErrorString(e));
  }
  const int NRS = (h->nzz - 2 + 63) / 64;
  h->KP = 64 * NRS + 64;
  h->NS = h->nxx + 2 * SK_OFF;
  int rc = dalloc(&h->d_spad, (size_t)h->nzz * h->nxx);
  if (!rc) rc = dalloc(&h->d_flag, 1);
  for (int o = 0; o < 4 && !rc; ++o) {
    rc = dalloc(&h->d_sk[o], (size_t)h->NS * h->KP);
    h->skew[o] = SkewView{h->d_sk[o], h->KP};
  }
  if (rc) {
    for (auto& p : h->d_sk) dfree(p);
    dfree(h->d_flag);
    dfree(h->d_spad);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return rc;
  }
  // refined reciprocals of the local solver's divisors, computed on the device
  {
    // divisors, their two-word reciprocals, refined reciprocals, check flags
    const EikConsts& c = h->kc;
    const float divs[12] = {c.sum,    c.hyp,    c.dz,    c.dx,    c.zh_sum, c.zh_hyp,
                            c.zh_dz,  c.zh_dx,  c.zl_sum, c.zl_hyp, c.zl_dz,  c.zl_dx};
    float rec[4] = {};
    int bad[4] = {1, 1, 1, 1};
    float* d = nullptr;
    hipError_t e2 = hipMalloc((void**)&d, 20 * sizeof(float));
    if (e2 == hipSuccess) e2 = hipMemcpy(d, divs, sizeof(divs), hipMemcpyHostToDevice);
    if (e2 == hipSuccess) e2 = hipMemset(d + 16, 0, 4 * sizeof(int));
    if (e2 == hipSuccess) {
      launch_recip(d, d + 12, 4, h->stream);
      launch_fastdiv_check(d, d + 4, d + 8, 4, (int*)(d + 16), h->stream);
      e2 = hipStreamSynchronize(h->stream);
    }
    if (e2 == hipSuccess) e2 = hipMemcpy(rec, d + 12, sizeof(rec), hipMemcpyDeviceToHost);
    if (e2 == hipSuccess) e2 = hipMemcpy(bad, d + 16, sizeof(bad), hipMemcpyDeviceToHost);
    if (d) (void)hipFree(d);
    h->kc.fastdiv = (bad[0] | bad[1] | bad[2] | bad[3]) == 0;
    if (e2 != hipSuccess) {
      seisfat_eik_destroy(h);
      return fail(SEISFAT_EHIP, std::string("reciprocal setup: ") + hipGetErrorString(e2));
    }
    h->kc.r_sum = rec[0];
    h->kc.r_hyp = rec[1];
    h->kc.r_dz = rec[2];
    h->kc.r_dx = rec[3];
  }
  *out = h;
  return SEISFAT_OK;
}

int seisfat_eik_destroy(seisfat_eik_t h) {
  if (!h) return SEISFAT_OK;
  (void)hipStreamSynchronize(h->stream);
  free_batch(h->single);
  dfree(h->d_spad);
  dfree(h->d_flag);
  for (auto& p : h->d_sk) dfree(p);
  for (auto& c : h->d_c) dfree(c);
  dfree(h->d_anitab);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return SEISFAT_OK;
}

int seisfat_eik_set_slowness_device(seisfat_eik_t h, const float* d_s, void* stream) {
  if (int rc = check_handle(h)) return rc;
  if (!d_s) return fail(SEISFAT_EINVAL, "null slowness pointer");
  hipStream_t st = (hipStream_t)stream;
  ++h->version;  // recorded adjoint runs of the previous model become stale
  HIPCHK(hipMemsetAsync(h->d_flag, 0, sizeof(int), st));
  launch_check_positive(d_s, (size_t)h->nz * h->nx, std::min(h->dx, h->dz), std::max(h->dx, h->dz), h->d_flag, st);
  launch_expand(d_s, h->d_spad, h->nz, h->nx, st);
  for (int o = 0; o < 4; ++o) launch_skew_s(h->d_spad, h->d_sk[o], h->nzz, h->nxx, h->KP, h->NS, o, st);
  HIPCHK(hipGetLastError());
  int bad = 0;
  HIPCHK(hipMemcpyAsync(&bad, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (bad) {
    h->has_s = false;
    return fail(SEISFAT_EINVAL, "slowness must be positive and finite, with slowness*spacing and "
                                "slowness/spacing in [1e-12, 1e12], everywhere");
  }
  h->has_s = true;
  h->pad_replicated = true;
  return SEISFAT_OK;
}

int seisfat_eik_set_slowness(seisfat_eik_t h, const float* s) {
  if (int rc = check_handle(h)) return rc;
  if (!s) return fail(SEISFAT_EINVAL, "null slowness pointer");
  const size_t n = (size_t)h->nz * h->nx;
  const float hmin = std::min(h->dx, h->dz), hmax = std::max(h->dx, h->dz);
  for (size_t k = 0; k < n; ++k)
    if (!slowness_ok(s[k], hmin, hmax)) {
      h->has_s = false;
      return fail(SEISFAT_EINVAL, "slowness must be positive and finite, with slowness*spacing and "
                                  "slowness/spacing in [1e-12, 1e12] (index " + std::to_string(k) + ")");
    }
  // runs of this handle's batches still in flight on any stream read the model
  HIPCHK(hipDeviceSynchronize());
  float* d_s = nullptr;
  if (int rc = dalloc(&d_s, n)) return rc;
  hipError_t e = hipMemcpyAsync(d_s, s, n * sizeof(float), hipMemcpyHostToDevice, h->stream);
  int rc = SEISFAT_OK;
  if (e != hipSuccess)
    rc = fail(SEISFAT_EHIP, std::string("upload slowness: ") + hipGetErrorString(e));
  else
    rc = seisfat_eik_set_slowness_device(h, d_s, h->stream);
  (void)hipStreamSynchronize(h->stream);
  dfree(d_s);
  return rc;
}

int seisfat_eik_set_slowness_padded(seisfat_eik_t h, const float* S_padded) {
  if (int rc = check_handle(h)) return rc;
  if (!S_padded) return fail(SEISFAT_EINVAL, "null slowness pointer");
  const int nz = h->nz, nx = h->nx, nzz = h->nzz, nxx = h->nxx;
  const size_t n = (size_t)nzz * nxx;
  const float hmin = std::min(h->dx, h->dz), hmax = std::max(h->dx, h->dz);
  // the padding is read by the solves (the reference's kernels read d_S
  // whole), so every node must be in the operating range
  for (size_t k = 0; k < n; ++k)
    if (!slowness_ok(S_padded[k], hmin, hmax)) {
      h->has_s = false;
      return fail(SEISFAT_EINVAL, "slowness must be positive and finite, with slowness*spacing and "
                                  "slowness/spacing in [1e-12, 1e12] (padded index " + std::to_string(k) + ")");
    }
  bool repl = true;
  for (int j = 0; j < nxx && repl; ++j) {
    const int jm = std::min(std::max(j, NB), nx + NB - 1);
    for (int i = 0; i < nzz; ++i) {
      const int im = std::min(std::max(i, NB), nz + NB - 1);
      if (S_padded[i + (size_t)j * nzz] != S_padded[im + (size_t)jm * nzz]) {
        repl = false;
        break;
      }
    }
  }
  // runs of this handle's batches still in flight on any stream read the model
  HIPCHK(hipDeviceSynchronize());
  ++h->version;  // recorded adjoint runs of the previous model become stale
  hipStream_t st = h->stream;
  HIPCHK(hipMemcpyAsync(h->d_spad, S_padded, n * sizeof(float), hipMemcpyHostToDevice, st));
  for (int o = 0; o < 4; ++o) launch_skew_s(h->d_spad, h->d_sk[o], nzz, nxx, h->KP, h->NS, o, st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  h->has_s = true;
  h->pad_replicated = repl;
  return SEISFAT_OK;
}

const float* seisfat_eik_device_S(seisfat_eik_t h) {
  if (!h || !h->has_s) {
    g_err = h ? "slowness not set" : "null handle";
    return nullptr;
  }
  return h->d_spad;
}

int seisfat_eik_set_stiffness(seisfat_eik_t h, const float* const c[6]) {
  if (int rc = check_handle(h)) return rc;
  if (!c) return fail(SEISFAT_EINVAL, "null stiffness array");
  const int nz = h->nz, nx = h->nx, nzz = h->nzz, nxx = h->nxx;
  const size_t n = (size_t)nzz * nxx;
  std::vector<float> pad(n);
  std::vector<uint16_t> q(n);
  for (int e = 0; e < 6; ++e) {
    if (!c[e]) return fail(SEISFAT_EINVAL, "null stiffness component " + std::to_string(e));
    for (size_t k = 0; k < (size_t)nz * nx; ++k)
      if (!std::isfinite(c[e][k])) return fail(SEISFAT_EINVAL, "non-finite stiffness component " + std::to_string(e));
    // expand_boundary (src/modeling/modeling.cu:204-232)
    for (int j = 0; j < nxx; ++j) {
      const int jm = std::min(std::max(j - NB, 0), nx - 1);
      for (int i = 0; i < nzz; ++i) {
        const int im = std::min(std::max(i - NB, 0), nz - 1);
        pad[i + (size_t)j * nzz] = c[e][im + (size_t)jm * nz];
      }
    }
    // compression with exact min/max (src/modeling/modeling.cu:264-279)
    float mx = -1e20f, mn = 1e20f;
    for (size_t k = 0; k < n; ++k) {
      mn = std::min(pad[k], mn);
      mx = std::max(pad[k], mx);
    }
    // A constant component (every VTI model's C15 and C35) would make the
    // reference's expression 0/0; code 1 decodes to mn + 0 * (mx - mn) = mn,
    // the value, exactly.
    for (size_t k = 0; k < n; ++k)
      q[k] = (mx == mn) ? uint16_t(1) : static_cast<uint16_t>(1.0f + (65535 - 1) * (pad[k] - mn) / (mx - mn));
    h->cmin[e] = mn;
    h->cmax[e] = mx;
    if (!h->d_c[e])
      if (int rc = dalloc(&h->d_c[e], n)) return rc;
    HIPCHK(hipMemcpyAsync(h->d_c[e], q.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  if (!h->d_anitab)
    if (int rc = dalloc(&h->d_anitab, 7 * n)) return rc;
  h->has_c = true;
  return SEISFAT_OK;
}

int seisfat_eik_solve(seisfat_eik_t h, int type, int schedule, float sx, float sz, float* T_padded,
                      const float* rx, const float* rz, int nrec, float* picks) {
  if (int rc = check_handle(h)) return rc;
  if (nrec < 0 || (nrec > 0 && !picks)) return fail(SEISFAT_EINVAL, "picks buffer required when nrec > 0");
  const seisfat_shot s{sx, sz, 0, h->nx, 0, nrec};
  seisfat_batch_s* b = h->single;
  int rc;
  if (b && b->type == type && b->schedule == schedule && b->nrec == nrec && !b->adj.ready) {
    // Same shape as the previous call: only the source and the receivers
    // change.  The kernels (and the graphs that replay them) read both from
    // device memory, so updating the shot tables and receiver arrays is all
    // a new shot needs: no allocation, fill or capture per shot.
    std::vector<int> recshot;
    std::vector<float> rxr, rzr;
    if ((rc = validate_shots(h, type, 1, &s, nrec, rx, rz, recshot, rxr, rzr))) return rc;
    const float sxr = sx;  // the whole model: grid origin = model origin
    for (auto* tab : {&b->iso, &b->iso_b, &b->qp, &b->qp_b})
      for (ShotDev& sd : *tab) {
        sd.sx = sxr;
        sd.sz = sz;
      }
    b->shots[0] = s;
    b->rxr = rxr;
    b->rzr = rzr;
    hipStream_t st = h->stream;
    auto up = [&](void* dst, const void* src, size_t bytes) -> int {
      if (dst && bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
      return SEISFAT_OK;
    };
    if ((rc = up(b->d_iso, b->iso.data(), sizeof(ShotDev))) || (rc = up(b->d_iso_b, b->iso_b.data(), sizeof(ShotDev))) ||
        (rc = up(b->d_qp, b->qp.data(), sizeof(ShotDev))) || (rc = up(b->d_qp_b, b->qp_b.data(), sizeof(ShotDev))) ||
        (rc = up(b->d_rx, b->rxr.data(), nrec * sizeof(float))) || (rc = up(b->d_rz, b->rzr.data(), nrec * sizeof(float))))
      return rc;
  } else {
    free_batch(h->single);
    h->single = nullptr;
    if ((rc = create_batch(h, type, schedule, 1, &s, nrec, rx, rz, /*use_graph=*/true, &b))) return rc;
    h->single = b;
  }
  if ((rc = run_batch(b, h->stream))) return rc;
  if ((rc = batch_status(b))) return rc;
  hipStream_t st = h->stream;
  if (T_padded) HIPCHK(hipMemcpyAsync(T_padded, b->d_T, b->tsize * sizeof(float), hipMemcpyDeviceToHost, st));
  if (nrec > 0) HIPCHK(hipMemcpyAsync(picks, b->d_picks, nrec * sizeof(float), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return SEISFAT_OK;
}

int seisfat_eik_fast_division(seisfat_eik_t h) {
  if (int rc = check_handle(h)) return rc;
  return h->kc.fastdiv ? 1 : 0;
}

const float* seisfat_eik_device_T(seisfat_eik_t h) {
  if (!h || !h->single) {
    g_err = "no solve yet";
    return nullptr;
  }
  return h->single->d_T;
}

int seisfat_eik_copy_T(seisfat_eik_t h, float* T_padded) {
  if (int rc = check_handle(h)) return rc;
  if (!h->single) return fail(SEISFAT_ESTATE, "no solve yet");
  if (!T_padded) return fail(SEISFAT_EINVAL, "null output buffer");
  return seisfat_batch_copy_T(h->single, 0, T_padded);
}

int seisfat_batch_create(seisfat_eik_t h, int type, int schedule, int nshots, const seisfat_shot* shots,
                         int nrec_total, const float* rx, const float* rz, seisfat_batch_t* out) {
  if (!out) return fail(SEISFAT_EINVAL, "null output pointer");
  *out = nullptr;
  if (int rc = check_handle(h)) return rc;
  return create_batch(h, type, schedule, nshots, shots, nrec_total, rx, rz, /*use_graph=*/true, out);
}

int seisfat_batch_create_opts(seisfat_eik_t h, int type, int schedule, int nshots, const seisfat_shot* shots,
                              int nrec_total, const float* rx, const float* rz, const seisfat_sched_opts* opts,
                              seisfat_batch_t* out) {
  if (!out) return fail(SEISFAT_EINVAL, "null output pointer");
  *out = nullptr;
  if (int rc = check_handle(h)) return rc;
  return create_batch(h, type, schedule, nshots, shots, nrec_total, rx, rz, /*use_graph=*/true, out, opts);
}

int seisfat_batch_get_opts(seisfat_batch_t b, seisfat_sched_opts* opts) {
  if (!b || !opts) return fail(SEISFAT_EINVAL, "null batch or output");
  *opts = b->opts;
  opts->pairs = b->pairs ? 1 : 0;
  return SEISFAT_OK;
}

int seisfat_batch_run(seisfat_batch_t b, void* stream) {
  if (!b) return fail(SEISFAT_EINVAL, "null batch");
  return run_batch(b, (hipStream_t)stream);
}

int seisfat_batch_run_converged(seisfat_batch_t b, int max_passes, int chunk, int* passes, void* stream) {
  if (!b) return fail(SEISFAT_EINVAL, "null batch");
  return run_converged(b, max_passes, chunk, passes, (hipStream_t)stream);
}

int seisfat_batch_picks(seisfat_batch_t b, float* picks) {
  if (!b) return fail(SEISFAT_EINVAL, "null batch");
  if (b->nrec > 0 && !picks) return fail(SEISFAT_EINVAL, "null picks buffer");
  if (int rc = batch_status(b)) return rc;
  if (b->nrec > 0) {
    HIPCHK(hipMemcpyAsync(picks, b->d_picks, b->nrec * sizeof(float), hipMemcpyDeviceToHost, b->h->stream));
    HIPCHK(hipStreamSynchronize(b->h->stream));
  }
  return SEISFAT_OK;
}

const float* seisfat_batch_device_picks(seisfat_batch_t b) { return b ? b->d_picks : nullptr; }

const float* seisfat_batch_device_T(seisfat_batch_t b, int shot, int* nzz, int* nxx) {
  if (!b || shot < 0 || shot >= b->nshots) {
    g_err = "bad batch or shot index";
    return nullptr;
  }
  if (nzz) *nzz = b->h->nzz;
  if (nxx) *nxx = b->iso[shot].nxx;
  return b->d_T + b->toff[shot];
}

int seisfat_batch_copy_T(seisfat_batch_t b, int shot, float* T_padded) {
  if (!b || shot < 0 || shot >= b->nshots || !T_padded) return fail(SEISFAT_EINVAL, "bad batch, shot or buffer");
  if (int rc = batch_status(b)) return rc;
  const size_t n = (size_t)b->h->nzz * b->iso[shot].nxx;
  HIPCHK(hipMemcpyAsync(T_padded, b->d_T + b->toff[shot], n * sizeof(float), hipMemcpyDeviceToHost, b->h->stream));
  HIPCHK(hipStreamSynchronize(b->h->stream));
  return SEISFAT_OK;
}

int seisfat_batch_num_launches(seisfat_batch_t b) { return b ? b->launches : -1; }

int seisfat_batch_num_sweep_launches(seisfat_batch_t b) { return b ? b->sweep_launches : -1; }

int seisfat_batch_sweep_ms(seisfat_batch_t b, float* ms) {
  if (!b || !ms) return fail(SEISFAT_EINVAL, "null batch or output");
  HIPCHK(hipEventSynchronize(b->ev_sweep[1]));
  HIPCHK(hipEventElapsedTime(ms, b->ev_sweep[0], b->ev_sweep[1]));
  return SEISFAT_OK;
}

int seisfat_batch_status(seisfat_batch_t b) {
  if (!b) return fail(SEISFAT_EINVAL, "null batch");
  return batch_status(b);
}

int seisfat_batch_tile_counts(seisfat_batch_t b, long long* total, long long* skipped) {
  if (!b || !total || !skipped) return fail(SEISFAT_EINVAL, "null batch or output");
  if (!b->d_ctl) return fail(SEISFAT_EINVAL, "tile counts exist for the persistent tiled schedule only");
  const long long solves = b->type == SEISFAT_ANI ? 2 : 1;
  *total = solves * (b->pairs ? 2LL * (long long)b->pair_tiles.size() : 4LL * b->ntiles);
  int n = 0;
  HIPCHK(hipEventSynchronize(b->ev_done));
  HIPCHK(hipMemcpyAsync(&n, b->d_ctl + CTL_SKIP, sizeof(int), hipMemcpyDeviceToHost, b->h->stream));
  HIPCHK(hipStreamSynchronize(b->h->stream));
  *skipped = n;
  return SEISFAT_OK;
}

int seisfat_batch_run_record(seisfat_batch_t b, void* stream) {
  if (!b) return fail(SEISFAT_EINVAL, "null batch");
  return run_record(b, (hipStream_t)stream);
}

int seisfat_batch_vjp_device(seisfat_batch_t b, const float* d_w, double* d_grad, void* stream) {
  if (!b) return fail(SEISFAT_EINVAL, "null batch");
  return enqueue_vjp(b, d_w, d_grad, (hipStream_t)stream);
}

int seisfat_batch_misfit_gradient(seisfat_batch_t b, const float* obs, double* J, double* grad, float* picks) {
  if (!b) return fail(SEISFAT_EINVAL, "null batch");
  if (!grad) return fail(SEISFAT_EINVAL, "null gradient buffer");
  if (b->nrec > 0 && !obs) return fail(SEISFAT_EINVAL, "null observed picks");
  seisfat_eik_s* h = b->h;
  if (int rc = run_record(b, h->stream)) return rc;
  if (int rc = batch_status(b)) return rc;
  std::vector<float> pk(b->nrec), w(b->nrec);
  std::vector<double> pk64(b->nrec);
  if (b->nrec > 0) {
    HIPCHK(hipMemcpy(pk.data(), b->d_picks, b->nrec * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(pk64.data(), b->adj.d_picks64, b->nrec * sizeof(double), hipMemcpyDeviceToHost));
  }
  double j = 0.0;
  for (int r = 0; r < b->nrec; ++r) {
    const double res = pk64[r] - (double)obs[r];
    j += 0.5 * res * res;
    w[r] = (float)res;
  }
  AdjState& A = b->adj;
  if (b->nrec > 0) HIPCHK(hipMemcpyAsync(A.d_w, w.data(), b->nrec * sizeof(float), hipMemcpyHostToDevice, h->stream));
  if (int rc = enqueue_vjp(b, A.d_w, A.d_grad, h->stream)) return rc;
  HIPCHK(hipMemcpyAsync(grad, A.d_grad, (size_t)h->nz * h->nx * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (J) *J = j;
  if (picks && b->nrec > 0) std::memcpy(picks, pk.data(), b->nrec * sizeof(float));
  return SEISFAT_OK;
}

int seisfat_batch_vjp(seisfat_batch_t b, const float* w, double* grad) {
  if (!b) return fail(SEISFAT_EINVAL, "null batch");
  if (!grad) return fail(SEISFAT_EINVAL, "null gradient buffer");
  if (b->nrec > 0 && !w) return fail(SEISFAT_EINVAL, "null cotangents");
  seisfat_eik_s* h = b->h;
  AdjState& A = b->adj;
  if (!A.recorded) return fail(SEISFAT_ESTATE, "no recorded run (seisfat_batch_run_record)");
  hipStream_t st = h->stream;
  // the recorded run may have gone to another stream: wait for it first
  HIPCHK(hipStreamWaitEvent(st, b->ev_done, 0));
  if (b->nrec > 0) HIPCHK(hipMemcpyAsync(A.d_w, w, b->nrec * sizeof(float), hipMemcpyHostToDevice, st));
  if (int rc = enqueue_vjp(b, A.d_w, A.d_grad, st)) return rc;
  HIPCHK(hipMemcpyAsync(grad, A.d_grad, (size_t)h->nz * h->nx * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return batch_status(b);
}

int seisfat_batch_destroy(seisfat_batch_t b) {
  if (!b) return SEISFAT_OK;
  (void)hipStreamSynchronize(b->h->stream);
  free_batch(b);
  return SEISFAT_OK;
}

}  // extern "C"
