/*
 * seisfat2d_kdm.h — C ABI of the MI355X Kirchhoff depth migration operators.
 *
 * Drop-in for the kernels and trace processing of SeisFAT2D's migration
 * module (paths relative to the SeisFAT2D repository): the image- and
 * angle-domain operators of src/migration/migration.cu:398-756 (IDKDM, ADKDM
 * and their least-squares variants call them per trace) and the wavelet
 * convolutions of src/migration/migration.cu:173-261.  They read the padded
 * traveltime tables the eikonal solver writes (eikonal_src_<k>.bin,
 * eikonal_rec_<k>.bin; here resident device arrays) and the model's padded
 * slowness (the reference's Modeling::d_S, the eikonal handle's model).
 * Conventions as in seisfat2d_eik.h: status 0 or a negative SEISFAT_E* code,
 * float32 grids z-fastest, void* streams are hipStream_t.
 */
#ifndef SEISFAT2D_KDM_H
#define SEISFAT2D_KDM_H

#include "seisfat2d_eik.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SEISFAT_KDM_IMAGE 0 /* image domain: model nz x nx (src/migration/IDKDM.cu:3-24) */
#define SEISFAT_KDM_ANGLE 1 /* angle domain: model nz x nang x ncmp (src/migration/ADKDM.cu:3-24) */

typedef struct seisfat_kdm_s* seisfat_kdm_t;

/* Parameters of src/migration/migration.cu:3-35 (time_samples, time_spacing,
 * mig_aperture in degrees, mig_angle_spacing) and the gather sizes of
 * Migration::set_gathers (:75-99): nang angle bins, ncmp CMP bins. */
typedef struct {
  int32_t domain;
  int32_t nt;
  float dt;
  float aperture;
  float da;
  int32_t nang;
  int32_t ncmp;
} seisfat_kdm_params;

/* Operators on the model of eikonal handle h (its padded slowness is the S
 * the weights read), for nsrc source and nrec receiver tables.  Replaces the
 * device state of Migration::set_parameters (src/migration/migration.cu:37-52). */
int seisfat_kdm_create(seisfat_eik_t h, const seisfat_kdm_params* p, int nsrc, int nrec,
                       seisfat_kdm_t* out);
/* The traveltime tables: d_src (nsrc) and d_rec (nrec) consecutive padded
 * (nz+6) x (nx+6) device grids, as set_src_travel_times / set_rec_travel_times
 * write them (src/migration/migration.cu:116-171).  Each table's derivative
 * quantities are evaluated here once, with the handle's current slowness (the
 * reference re-evaluates them for every trace, with the same operations); a
 * later slowness upload to the handle makes the tables stale: the operators
 * then fail with SEISFAT_ESTATE until every table is set again. */
int seisfat_kdm_set_tables(seisfat_kdm_t k, const float* d_src, const float* d_rec, void* stream);
/* One table at a time: table `index` of the sources (which = 0) or receivers
 * (which = 1) from a padded grid in device memory, e.g. seisfat_batch_device_T
 * of a batched solve, so tables need not be gathered into one array.  The
 * operators run once every table has been set (by either call). */
int seisfat_kdm_set_table(seisfat_kdm_t k, int which, int index, const float* d_table, void* stream);
/* The traces, in the order the migration loops visit them
 * (src/migration/KDM.cu:14-53): trace t pairs source table src[t] with
 * receiver table rec[t], at CMP position cmp[t] (metres) and CMP bin
 * cmp_id[t] (angle domain).  Host arrays. */
int seisfat_kdm_set_traces(seisfat_kdm_t k, int ntr, const int* src, const int* rec, const float* cmp,
                           const int* cmp_id);
/* model += L^T data: perform_adjoint for every trace in order
 * (image_domain_adjoint_kernel / angle_domain_adjoint_kernel,
 * src/migration/migration.cu:398-483,572-663).  d_data: ntr x nt device
 * array (each trace after the adjoint convolution); d_model: the domain's
 * model (device).  The image domain keeps partial images in a buffer of the
 * operator (sized by seisfat_kdm_set_traces, at most 1 GiB): calls of one
 * operator run one at a time (one stream, or synchronised between streams). */
int seisfat_kdm_adjoint(seisfat_kdm_t k, const float* d_data, float* d_model, void* stream);
/* data += L model: perform_forward for every trace
 * (image_domain_forward_kernel / angle_domain_forward_kernel,
 * src/migration/migration.cu:485-570,665-756). */
int seisfat_kdm_forward(seisfat_kdm_t k, const float* d_model, float* d_data, void* stream);
int seisfat_kdm_destroy(seisfat_kdm_t k);

/* Wavelet convolution of ntr traces of nt samples: Migration::forward_convolution
 * (adjoint = 0) or adjoint_convolution (adjoint = 1), src/migration/migration.cu:197-261,
 * with wavelet (host, nw samples).  The reference multiplies spectra with FFTW
 * (unnormalised, size nfft = the power of two >= nt + nw - 1) and shifts by
 * nw/2 + nw/10 samples; here the same linear (cross-)correlation is summed
 * directly in double precision, times nfft.  d_in and d_out: ntr x nt device
 * arrays (distinct). */
int seisfat_trace_convolve(int adjoint, int ntr, int nt, const float* wavelet, int nw, const float* d_in,
                           float* d_out, void* stream);

/* The Ricker wavelet of Migration::set_wavelet (src/migration/migration.cu:54-73)
 * for time spacing dt (s) and maximum frequency fmax (Hz), evaluated on the
 * host as the reference does: returns its length nw = 2 * (int)((t0 - dt/2) / dt),
 * t0 = 2 sqrt(pi) / fmax, and writes its first min(nw, cap) samples to out. */
int seisfat_ricker(float dt, float fmax, float* out, int cap);

#ifdef __cplusplus
}
#endif

#endif
