// Kirchhoff depth migration operators on MI355X: the image- and angle-domain
// kernels of src/migration/migration.cu:398-756 and the wavelet
// convolutions of :173-261.
//
// The reference launches one kernel per trace (source, receiver pair), one
// thread per padded node, each evaluating the node's weight from 9-point
// finite differences of both traveltime tables and adding into the image
// (adjoint) or the trace (forward) with atomicAdd.  Here:
//   - the per-table part of the weight (first and second derivatives, the
//     Hessian determinant, the gradient norm, the incidence angle and its
//     cosine) is evaluated once per table (k_kdm_prep), with the reference's
//     operations, and every trace reads it (7 words per node and table);
//   - image-domain adjoint: one thread per image point loops over the traces
//     in the migration's order and keeps its sum in a register, so the image
//     receives the same FP32 sequence of additions as under the reference's
//     launch order, without atomics;
//   - angle-domain adjoint: the same loop, adding into the (depth, angle,
//     CMP) bins with atomics, as the reference (their order varies);
//   - forward: one workgroup per trace accumulates the trace in an LDS
//     histogram of its nt samples (LDS atomics) and adds it to the trace once;
//   - convolution: the FFT product of the reference is a linear
//     (cross-)correlation with the wavelet, summed directly in double.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/seisfat2d_kdm.h"
#include "eik_handle.h"

namespace seisfat {

constexpr float KDM_EPS = 1e-6f;  // EPS of src/migration/migration.cuh:7
constexpr int KQ = 7;             // per-table quantities: T, dT/dx, dT/dz, |det H| + EPS, |grad T| + EPS, angle, |cos|

struct KdmK {
  int nz, nx, nzz, npts, nt, nang;
  float dx, dz, dt, da;
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
// coordinates), each the expression the reference kernels evaluate.
__global__ __launch_bounds__(256) void k_kdm_prep(const float* __restrict__ tables, int ntab, KdmK K,
                                                  float* __restrict__ q) {
  const size_t n = (size_t)ntab * K.npts;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t tsize = (size_t)K.nzz * (K.nx + 6);
  for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < n; g += stride) {
    const int tab = (int)(g / K.npts);
    const int p = (int)(g - (size_t)tab * K.npts);
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
    float* __restrict__ o = q + (size_t)tab * KQ * K.npts + p;
    o[0] = c;
    o[(size_t)1 * K.npts] = d_x;
    o[(size_t)2 * K.npts] = d_z;
    o[(size_t)3 * K.npts] = det;
    o[(size_t)4 * K.npts] = norm;
    o[(size_t)5 * K.npts] = acosf(fminf(1.0f, fmaxf(-1.0f, dot)));
    o[(size_t)6 * K.npts] = fabsf(dot);
  }
}

struct Q7 {
  float t, dx, dz, det, norm, theta, cosv;
};
__device__ __forceinline__ Q7 q_load(const float* __restrict__ q, int tab, int p, int npts) {
  const float* __restrict__ b = q + (size_t)tab * KQ * npts + p;
  return Q7{b[0], b[(size_t)npts], b[(size_t)2 * npts], b[(size_t)3 * npts], b[(size_t)4 * npts],
            b[(size_t)5 * npts], b[(size_t)6 * npts]};
}

// Kirchhoff weight of one node for one trace (the product the reference
// kernels form): Hessian amplitude, obliquity, spreading, reflection-angle
// factor and the aperture taper about the trace's CMP.
__device__ __forceinline__ float kdm_weight(const Q7& s, const Q7& r, float S, float x, float sigma, float cmp,
                                            const KdmK& K) {
  const float jterm = sqrtf(s.det * r.det) * S / (s.norm * r.norm);
  const float obliq = sqrtf(fmaxf(0.0f, s.cosv) * fmaxf(0.0f, r.cosv));
  const float rs = fmaxf(s.t / S, KDM_EPS), rr = fmaxf(r.t / S, KDM_EPS);
  const float g = 1.0f / sqrtf(rs * rr);
  const float rf = 1.0f + 0.2f * cosf(0.5f * (s.theta + r.theta));
  const float a = (x - cmp) / (sigma + KDM_EPS);
  const float taper = expf(-0.5f * a * a);
  return K.scale * jterm * obliq * g * rf * taper;
}

// Reflection-angle bin of the node for the trace (the reference's aId).
__device__ __forceinline__ int kdm_angle_bin(const Q7& s, const Q7& r, const KdmK& K) {
  const float ra = 0.5f * acosf((s.dx * r.dx + s.dz * r.dz) / (s.norm * r.norm));
  const float ang = (float)((double)(180.0f * ra) / M_PI);
  return f2i_rd(ang / K.da);
}

// Image-domain adjoint: thread per image point, traces in order.
__global__ __launch_bounds__(256) void k_kdm_adjoint_image(const float* __restrict__ q, const float* __restrict__ S,
                                                           const int* __restrict__ tsrc, const int* __restrict__ trec,
                                                           const float* __restrict__ tcmp, int ntr, int nsrc,
                                                           const float* __restrict__ data, float* __restrict__ model,
                                                           KdmK K) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= K.npts) return;
  const int i = p % K.nz, j = p / K.nz;
  const float z = (float)i * K.dz, x = (float)j * K.dx;
  const float s_node = S[(i + 3) + (size_t)(j + 3) * K.nzz];
  const float sigma = K.tan_ap * z;
  float acc = model[p];
  int cur = -1;
  Q7 qs{};
  for (int t = 0; t < ntr; ++t) {
    const int s = tsrc[t];
    if (s != cur) {
      qs = q_load(q, s, p, K.npts);
      cur = s;
    }
    const Q7 qr = q_load(q, nsrc + trec[t], p, K.npts);
    const int it = f2i_rd((qs.t + qr.t) / K.dt);
    if ((unsigned)it < (unsigned)K.nt)
      acc += kdm_weight(qs, qr, s_node, x, sigma, tcmp[t], K) * data[(size_t)t * K.nt + it];
  }
  model[p] = acc;
}

// Angle-domain adjoint: thread per image point, traces in order, bins by atomics.
__global__ __launch_bounds__(256) void k_kdm_adjoint_angle(const float* __restrict__ q, const float* __restrict__ S,
                                                           const int* __restrict__ tsrc, const int* __restrict__ trec,
                                                           const float* __restrict__ tcmp,
                                                           const int* __restrict__ tcid, int ntr, int nsrc,
                                                           const float* __restrict__ data, float* __restrict__ model,
                                                           KdmK K) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= K.npts) return;
  const int i = p % K.nz, j = p / K.nz;
  const float z = (float)i * K.dz, x = (float)j * K.dx;
  const float s_node = S[(i + 3) + (size_t)(j + 3) * K.nzz];
  const float sigma = K.tan_ap * z;
  int cur = -1;
  Q7 qs{};
  for (int t = 0; t < ntr; ++t) {
    const int s = tsrc[t];
    if (s != cur) {
      qs = q_load(q, s, p, K.npts);
      cur = s;
    }
    const Q7 qr = q_load(q, nsrc + trec[t], p, K.npts);
    const int it = f2i_rd((qs.t + qr.t) / K.dt);
    if ((unsigned)it < (unsigned)K.nt) {
      const float w = kdm_weight(qs, qr, s_node, x, sigma, tcmp[t], K);
      const int a = kdm_angle_bin(qs, qr, K);
      if (a >= 0 && a < K.nang)
        atomicAdd(model + i + (size_t)a * K.nz + (size_t)tcid[t] * K.nang * K.nz, w * data[(size_t)t * K.nt + it]);
    }
  }
}

// Forward (image or angle domain): one workgroup per trace, the trace's nt
// samples in LDS.
template <bool ANGLE>
__global__ __launch_bounds__(256) void k_kdm_forward(const float* __restrict__ q, const float* __restrict__ S,
                                                     const int* __restrict__ tsrc, const int* __restrict__ trec,
                                                     const float* __restrict__ tcmp, const int* __restrict__ tcid,
                                                     int nsrc, const float* __restrict__ model,
                                                     float* __restrict__ data, KdmK K) {
  extern __shared__ float hist[];
  const int t = blockIdx.x;
  for (int k = threadIdx.x; k < K.nt; k += blockDim.x) hist[k] = 0.0f;
  __syncthreads();
  const int s = tsrc[t], r = nsrc + trec[t];
  const float cmp = tcmp[t];
  const size_t moff = ANGLE ? (size_t)tcid[t] * K.nang * K.nz : 0;
  for (int p = threadIdx.x; p < K.npts; p += blockDim.x) {
    const Q7 qs = q_load(q, s, p, K.npts);
    const Q7 qr = q_load(q, r, p, K.npts);
    const int it = f2i_rd((qs.t + qr.t) / K.dt);
    if ((unsigned)it >= (unsigned)K.nt) continue;
    const int i = p % K.nz, j = p / K.nz;
    const float z = (float)i * K.dz, x = (float)j * K.dx;
    const float w = kdm_weight(qs, qr, S[(i + 3) + (size_t)(j + 3) * K.nzz], x, K.tan_ap * z, cmp, K);
    if (ANGLE) {
      const int a = kdm_angle_bin(qs, qr, K);
      if (a >= 0 && a < K.nang) atomicAdd(hist + it, w * model[moff + i + (size_t)a * K.nz]);
    } else {
      atomicAdd(hist + it, w * model[p]);
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

}  // namespace seisfat

using namespace seisfat;

struct seisfat_kdm_s {
  seisfat_eik_s* h = nullptr;
  seisfat_kdm_params p{};
  int nsrc = 0, nrec = 0, ntr = 0;
  KdmK K{};
  float* d_q = nullptr;    // per-table quantities, (nsrc + nrec) x 7 x npts
  int* d_src = nullptr;    // trace tables
  int* d_rec = nullptr;
  float* d_cmp = nullptr;
  int* d_cid = nullptr;
  bool has_tables = false;
};

namespace {

int kfail(int code, const std::string& m) { return seisfat::set_error(code, m.c_str()); }

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

void free_traces(seisfat_kdm_s* k) {
  kfree(k->d_src);
  kfree(k->d_rec);
  kfree(k->d_cmp);
  kfree(k->d_cid);
  k->ntr = 0;
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
  K.nt = p->nt;
  K.nang = p->domain == SEISFAT_KDM_ANGLE ? p->nang : 0;
  K.dx = h->dx;
  K.dz = h->dz;
  K.dt = p->dt;
  K.da = p->domain == SEISFAT_KDM_ANGLE ? p->da : 1.0f;
  // sigma = tanf(aperture * pi / 180) * z, the angle rounded to float as the
  // reference's float argument takes it; 1 / (2 pi) rounded from double
  K.tan_ap = std::tan((float)((double)p->aperture * M_PI / 180.0));
  K.scale = (float)(1.0 / (2.0 * M_PI));
  const hipError_t e = hipMalloc((void**)&k->d_q, (size_t)(nsrc + nrec) * KQ * npts * sizeof(float));
  if (e != hipSuccess) {
    delete k;
    return kfail(SEISFAT_ENOMEM, std::string("hipMalloc of the table quantities: ") + hipGetErrorString(e));
  }
  *out = k;
  return SEISFAT_OK;
}

int seisfat_kdm_set_tables(seisfat_kdm_t k, const float* d_src, const float* d_rec, void* stream) {
  if (!k || !d_src || !d_rec) return kfail(SEISFAT_EINVAL, "null operator or table pointer");
  hipStream_t st = (hipStream_t)stream;
  const size_t nq = (size_t)k->nsrc * k->K.npts;
  const int blocks = (int)std::min<size_t>((nq + 255) / 256, 16384);
  hipLaunchKernelGGL(k_kdm_prep, dim3(blocks), dim3(256), 0, st, d_src, k->nsrc, k->K, k->d_q);
  const size_t nr = (size_t)k->nrec * k->K.npts;
  const int blocks_r = (int)std::min<size_t>((nr + 255) / 256, 16384);
  hipLaunchKernelGGL(k_kdm_prep, dim3(blocks_r), dim3(256), 0, st, d_rec, k->nrec, k->K,
                     k->d_q + (size_t)k->nsrc * KQ * k->K.npts);
  KHIP(hipGetLastError());
  k->has_tables = true;
  return SEISFAT_OK;
}

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
  if (ntr == 0) return SEISFAT_OK;
  KHIP(hipMalloc((void**)&k->d_src, ntr * sizeof(int)));
  KHIP(hipMalloc((void**)&k->d_rec, ntr * sizeof(int)));
  KHIP(hipMalloc((void**)&k->d_cmp, ntr * sizeof(float)));
  KHIP(hipMalloc((void**)&k->d_cid, ntr * sizeof(int)));
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
  if (!k || !d_data || !d_model) return kfail(SEISFAT_EINVAL, "null operator or buffer");
  if (!k->has_tables) return kfail(SEISFAT_ESTATE, "tables not set (seisfat_kdm_set_tables)");
  if (k->ntr == 0) return SEISFAT_OK;
  hipStream_t st = (hipStream_t)stream;
  const int blocks = (k->K.npts + 255) / 256;
  if (k->p.domain == SEISFAT_KDM_IMAGE)
    hipLaunchKernelGGL(k_kdm_adjoint_image, dim3(blocks), dim3(256), 0, st, k->d_q, k->h->d_spad, k->d_src,
                       k->d_rec, k->d_cmp, k->ntr, k->nsrc, d_data, d_model, k->K);
  else
    hipLaunchKernelGGL(k_kdm_adjoint_angle, dim3(blocks), dim3(256), 0, st, k->d_q, k->h->d_spad, k->d_src,
                       k->d_rec, k->d_cmp, k->d_cid, k->ntr, k->nsrc, d_data, d_model, k->K);
  KHIP(hipGetLastError());
  return SEISFAT_OK;
}

int seisfat_kdm_forward(seisfat_kdm_t k, const float* d_model, float* d_data, void* stream) {
  if (!k || !d_data || !d_model) return kfail(SEISFAT_EINVAL, "null operator or buffer");
  if (!k->has_tables) return kfail(SEISFAT_ESTATE, "tables not set (seisfat_kdm_set_tables)");
  if (k->ntr == 0) return SEISFAT_OK;
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = (size_t)k->K.nt * sizeof(float);
  if (k->p.domain == SEISFAT_KDM_IMAGE)
    hipLaunchKernelGGL(k_kdm_forward<false>, dim3(k->ntr), dim3(256), lds, st, k->d_q, k->h->d_spad, k->d_src,
                       k->d_rec, k->d_cmp, k->d_cid, k->nsrc, d_model, d_data, k->K);
  else
    hipLaunchKernelGGL(k_kdm_forward<true>, dim3(k->ntr), dim3(256), lds, st, k->d_q, k->h->d_spad, k->d_src,
                       k->d_rec, k->d_cmp, k->d_cid, k->nsrc, d_model, d_data, k->K);
  KHIP(hipGetLastError());
  return SEISFAT_OK;
}

int seisfat_kdm_destroy(seisfat_kdm_t k) {
  if (!k) return SEISFAT_OK;
  (void)hipDeviceSynchronize();
  free_traces(k);
  kfree(k->d_q);
  delete k;
  return SEISFAT_OK;
}

int seisfat_trace_convolve(int adjoint, int ntr, int nt, const float* wavelet, int nw, const float* d_in,
                           float* d_out, void* stream) {
  if (ntr < 0 || nt <= 0 || nw <= 0 || !wavelet || (ntr > 0 && (!d_in || !d_out)) || d_in == d_out)
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
  const int blocks = (int)((n + 255) / 256);
  if (adjoint)
    hipLaunchKernelGGL(k_trace_conv<true>, dim3(blocks), dim3(256), 0, st, d_in, d_out, ntr, nt, d_w, nw, shift,
                       (double)nfft);
  else
    hipLaunchKernelGGL(k_trace_conv<false>, dim3(blocks), dim3(256), 0, st, d_in, d_out, ntr, nt, d_w, nw, shift,
                       (double)nfft);
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
