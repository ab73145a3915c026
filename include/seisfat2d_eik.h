/*
 * seisfat2d_eik.h — C ABI of the MI355X fast-sweeping eikonal solver.
 *
 * Drop-in boundary for SeisFAT2D's eikonal modeling path.  Every entry point
 * names the reference interface it replaces (paths relative to the SeisFAT2D
 * repository).  Conventions:
 *   - status: 0 on success, a negative SEISFAT_E* code on failure; the message
 *     of the last failure on the calling thread is seisfat_last_error().
 *   - grids are float32, z-fastest (index = i + j*n_rows), as the reference
 *     stores them (src/modeling/modeling.cu:209-212).  "Padded" grids carry the
 *     reference's nb = 3 replicated border: (nz+6) x (nx+6) nodes.
 *   - positions are in metres from the model origin, x along columns and z
 *     along rows, as in the reference's SPS/RPS files.
 *   - a handle is used by one host thread at a time; buffers are caller owned.
 *   - "stream" arguments are hipStream_t values passed as void* (NULL = HIP's
 *     default stream); work is ordered on that stream only.
 */
#ifndef SEISFAT2D_EIK_H
#define SEISFAT2D_EIK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEISFAT_OK 0
#define SEISFAT_EINVAL (-1) /* invalid argument (sizes, positions, slowness) */
#define SEISFAT_EHIP (-2)   /* HIP runtime error */
#define SEISFAT_ENOMEM (-3) /* device or host allocation failed */
#define SEISFAT_ESTATE (-4) /* call out of order (e.g. solve before set_slowness) */
#define SEISFAT_EIO (-5)    /* file I/O error */

/* Model type: the reference's plugin index (src/modeling_main.cpp:6-15). */
#define SEISFAT_ISO 0 /* Eikonal_ISO (src/modeling/eikonal_iso.cu) */
#define SEISFAT_ANI 1 /* Eikonal_ANI (src/modeling/eikonal_ani.cu) */

/* Sweep schedule of a solve.  Both compute the reference's single pass of four
 * sweeps (src/modeling/modeling.cu:123-159) and give bit-identical traveltimes. */
#define SEISFAT_SCHED_PLAIN 0      /* one launch per anti-diagonal: the reference's schedule */
#define SEISFAT_SCHED_TILED 1      /* 64x64 skewed LDS tiles, one persistent launch per sweep */
#define SEISFAT_SCHED_TILED_DIAG 2 /* the same tiles, one launch per tile diagonal */

typedef struct seisfat_eik_s* seisfat_eik_t;
typedef struct seisfat_batch_s* seisfat_batch_t;

/* One shot of a batch.  The shot is solved on the model columns
 * [x0, x0+width) as if that window were the whole model file (nb = 3
 * replicated border around the window); sx is still measured from the model
 * origin.  Receivers rec_offset .. rec_offset+nrec-1 of the batch's receiver
 * arrays are picked from the shot's traveltimes. */
typedef struct {
  float sx, sz;
  int32_t x0, width;
  int32_t rec_offset, nrec;
} seisfat_shot;

/* Last error message of the calling thread ("" if none). */
const char* seisfat_last_error(void);

/* Model geometry: Modeling::set_parameters (src/modeling/modeling.cu:3-36)
 * reading x_samples, z_samples, x_spacing, z_spacing.  Allocates the model on
 * the current HIP device. */
int seisfat_eik_create(int nz, int nx, float dz, float dx, seisfat_eik_t* out);
int seisfat_eik_destroy(seisfat_eik_t h);

/* Slowness model, nz*nx unpadded, strictly positive and finite.
 * Replaces Modeling::set_properties + expand_boundary + copy_slowness_to_device
 * (src/modeling/modeling.cu:47-60,204-232,281-284).  The host variant copies
 * from host memory; the device variant from a device pointer on `stream`. */
int seisfat_eik_set_slowness(seisfat_eik_t h, const float* s);
int seisfat_eik_set_slowness_device(seisfat_eik_t h, const float* d_s, void* stream);

/* Padded slowness, (nz+6)*(nx+6) values uploaded verbatim from host memory:
 * Modeling::copy_slowness_to_device (src/modeling/modeling.cu:281-284), which
 * copies the caller's whole padded S.  A shot on the whole model (every
 * seisfat_eik_solve, and batch shots with x0 = 0, width = nx) reads this
 * padding as given, as the reference's kernels read d_S; so a tomography loop
 * that edits only the interior of S (src/inversion/tomography_iso.cpp:77-92)
 * keeps the starting model's border, as in the reference.  Window shots read
 * the border rows as uploaded and replicate their own edge columns.  Every
 * value, padding included, must be positive and finite with slowness*spacing
 * and slowness/spacing in [1e-12, 1e12].  The adjoint entry points
 * differentiate with respect to the unpadded model: when the border is the
 * replication of the interior (always after seisfat_eik_set_slowness) the
 * border's derivative folds onto the edge cells it copies; otherwise the
 * border is a constant and only interior cells (and a window's own
 * replicated columns) carry derivative. */
int seisfat_eik_set_slowness_padded(seisfat_eik_t h, const float* S_padded);

/* Device pointer of the padded slowness, (nz+6)*(nx+6) floats z-fastest: the
 * reference's public Modeling::d_S (src/modeling/modeling.cuh:171).  Read
 * only: change it through seisfat_eik_set_slowness*.  NULL before a slowness
 * is set.  Valid until the next slowness upload or destroy. */
const float* seisfat_eik_device_S(seisfat_eik_t h);

/* Anisotropic stiffness C11, C13, C15, C33, C35, C55 (each nz*nx unpadded, in
 * the reference's units).  Replaces Eikonal_ANI::set_stiffness_element
 * (src/modeling/eikonal_ani.cu:3-16): pad, then uint16 compression with exact
 * min/max (src/modeling/modeling.cu:264-279). */
int seisfat_eik_set_stiffness(seisfat_eik_t h, const float* const c[6]);

/* One shot on the whole model: Eikonal_ISO/ANI::time_propagation
 * (src/modeling/eikonal_iso.cu:9-13, src/modeling/eikonal_ani.cu:33-45)
 * followed by Modeling::compute_seismogram (src/modeling/modeling.cu:161-194).
 * T_padded (nullable, host) receives the padded (nz+6)*(nx+6) traveltimes;
 * picks (nullable when nrec = 0, host) receive the Catmull-Rom receiver
 * times.  Synchronous on the handle's own stream.  The first call (and any
 * call that changes type, schedule or receiver count) builds a one-shot
 * batch; later calls reuse it, uploading only the source position and the
 * receivers (no device allocation, fill or graph capture per shot). */
int seisfat_eik_solve(seisfat_eik_t h, int type, int schedule, float sx, float sz,
                      float* T_padded, const float* rx, const float* rz, int nrec,
                      float* picks);

/* The per-shot solve of seisfat_eik_solve for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Isotropic, on
 * the whole model (its padding as uploaded), device pointers, asynchronous on
 * `stream`: d_T receives the padded (nz+6)*(nx+6) traveltimes, d_picks the
 * nrec Catmull-Rom picks at (d_rx, d_rz), and d_err (nullable) 0, or nonzero
 * when a sweep's tile dependency timed out (1) or a receiver lay outside the
 * model (2, its pick NaN).  It allocates nothing: its work buffers live in a
 * fixed 8 MiB device array per GPU, set up on the first call on that device;
 * models whose tile queue and exchange words exceed it (about 60 M padded
 * nodes, fewer for models narrower than a few hundred columns) are refused
 * with SEISFAT_EINVAL.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (from two
 * streams or threads) race.  Same traveltimes and picks as seisfat_eik_solve,
 * bit for bit. */
int seisfat_eik_solve_noalloc(seisfat_eik_t h, float sx, float sz, float* d_T, const float* d_rx,
                              const float* d_rz, int nrec, float* d_picks, int* d_err, void* stream);

/* Device pointer of the padded traveltimes of the last seisfat_eik_solve (the
 * reference's public Modeling::d_T, read by migration,
 * src/migration/migration.cu:130,167).  Valid until the next solve/destroy. */
const float* seisfat_eik_device_T(seisfat_eik_t h);
/* Copy the padded traveltimes of the last seisfat_eik_solve to host memory
 * ((nz+6)*(nx+6) floats): the D2H copy of Modeling::compute_seismogram
 * (src/modeling/modeling.cu:165) that leaves the public Modeling::T valid. */
int seisfat_eik_copy_T(seisfat_eik_t h, float* T_padded);
/* 1 when the batched tiled sweeps of this geometry divide with the
 * two-instruction known-divisor sequence (its four divisors passed the
 * exhaustive check at create), 0 when they use the five-instruction one; the
 * traveltimes are bit-identical either way.  Diagnostic, no reference
 * counterpart. */
int seisfat_eik_fast_division(seisfat_eik_t h);

/* Batched solve of many shots (no reference counterpart: the reference loops
 * shots one at a time, src/modeling_main.cpp:21-29).  create() validates the
 * shots, allocates every shot's traveltime grid and, for the tiled schedule,
 * records the whole solve (initialisation, all sweeps, picks) as one HIP
 * graph.  rx/rz are host arrays of nrec_total receivers. */
int seisfat_batch_create(seisfat_eik_t h, int type, int schedule, int nshots,
                         const seisfat_shot* shots, int nrec_total, const float* rx,
                         const float* rz, seisfat_batch_t* out);
/* Schedule options of a batch of the persistent tiled schedule (no reference
 * counterpart).  Every combination gives bit-identical traveltimes; they
 * trade launch shape against load balance, and are what the package's
 * schedule tuner (seisfat2d_amd.tune) searches per workload.
 * seisfat_batch_create uses the defaults. */
typedef struct seisfat_sched_opts {
  int pairs;    /* 1 (default): both sweeps of each pair of sweeps in one launch where the device runs
                   workgroups round robin over 8 XCDs; 0: one persistent launch per sweep */
  int wpcu;     /* persistent workgroups per CU; 0 (default): as many as fit */
  int skip;     /* 1 (default): tiles whose every input is the not-reached value 1e6 are not run where
                   the geometry provably keeps them at 1e6; 0: every tile runs */
  int priority; /* 1 (default): wave priorities from each shot's scheduling slack; 0: all equal */
} seisfat_sched_opts;
/* seisfat_batch_create with schedule options (null opts: the defaults). */
int seisfat_batch_create_opts(seisfat_eik_t h, int type, int schedule, int nshots,
                              const seisfat_shot* shots, int nrec_total, const float* rx,
                              const float* rz, const seisfat_sched_opts* opts, seisfat_batch_t* out);
/* The options a batch runs with (pairs reads 0 where the fused pairs are
 * unavailable). */
int seisfat_batch_get_opts(seisfat_batch_t b, seisfat_sched_opts* opts);
/* Run the batch asynchronously on `stream` (graph replay for the tiled schedule). */
int seisfat_batch_run(seisfat_batch_t b, void* stream);
/* Wait for the batch and copy all nrec_total picks to host memory. */
int seisfat_batch_picks(seisfat_batch_t b, float* picks);
/* Device pointer of the picks (nrec_total floats). */
const float* seisfat_batch_device_picks(seisfat_batch_t b);
/* Device pointer of shot k's padded traveltimes; *nzz, *nxx receive its shape. */
const float* seisfat_batch_device_T(seisfat_batch_t b, int shot, int* nzz, int* nxx);
/* Copy shot k's padded traveltimes to host memory (nzz*nxx floats). */
int seisfat_batch_copy_T(seisfat_batch_t b, int shot, float* T_padded);
/* Number of HIP kernel launches of one run, and of them the sweep-kernel
 * launches (diagnostics). */
int seisfat_batch_num_launches(seisfat_batch_t b);
int seisfat_batch_num_sweep_launches(seisfat_batch_t b);
/* Device time in ms of the sweep phase of the last completed run (from HIP
 * events recorded on the run's stream around the sweep launches). */
int seisfat_batch_sweep_ms(seisfat_batch_t b, float* ms);
/* Tiles of the last completed run (or converged-mode pass) of the persistent
 * tiled schedule: *total = the tiles of all its sweeps, *skipped = those whose
 * every input (old traveltimes, upwind row, carry-in) was the "not reached"
 * value 1e6, which the geometry provably keeps at 1e6, so the kernel did not
 * run their steps (results unchanged).  Diagnostics with no reference
 * counterpart; other schedules report SEISFAT_EINVAL. */
int seisfat_batch_tile_counts(seisfat_batch_t b, long long* total, long long* skipped);
/* Converged mode (no reference counterpart: the reference runs exactly one
 * pass of four sweeps, src/modeling/modeling.cu:130-158): passes repeat until
 * one leaves every traveltime of every shot unchanged, at most max_passes
 * (passes receives the count, the unchanged pass included, as the oracle's
 * converged mode counts them).  Each tile votes whether it lowered a
 * traveltime (one wave vote, one global word per pass); passes run as HIP
 * graphs of `chunk` passes (rounded up to even) whose passes after
 * convergence exit at once, with a host check between chunks; then the
 * picks.  Isotropic batches of the tiled schedule (one persistent launch per
 * sweep).  Synchronous on `stream`. */
int seisfat_batch_run_converged(seisfat_batch_t b, int max_passes, int chunk, int* passes, void* stream);
/* Wait for the batch's last enqueued work (run, recorded run or
 * vector-Jacobian product; work of other streams is not waited for) and
 * report a failure the device flagged during it (a persistent sweep whose
 * tile dependencies did not arrive in time).  seisfat_batch_picks and
 * seisfat_batch_copy_T wait the same way. */
int seisfat_batch_status(seisfat_batch_t b);
int seisfat_batch_destroy(seisfat_batch_t b);

/* Adjoint-state traveltime gradient (isotropic batches).  Replaces the
 * ray-traced sensitivity of the reference's tomography
 * (Inversion::gradient_ray_tracing, src/inversion/inversion.cpp:95-177, and
 * Tomography_ISO::set_sensitivity_matrix, src/inversion/tomography_iso.cpp:
 * 15-68) with the exact derivative of the single-pass solve.
 *
 * run_record: the batch's forward solve (same traveltimes and picks as
 * seisfat_batch_run) followed by an fp64 solve of the same single-pass scheme
 * whose trajectory (initial field and the state after each sweep) the
 * adjoint linearises; asynchronous on `stream`.  The first call allocates the
 * adjoint state (about 20 floats per padded node per shot).  The calls on one
 * batch (run, run_record, vjp_device) share its device state: issue them on
 * one stream, or order them with events; concurrent calls on one batch race. */
int seisfat_batch_run_record(seisfat_batch_t b, void* stream);
/* Vector-Jacobian product of the picks of the recorded run:
 *   d_grad[iz + ix*nz] = sum_r d_w[r] * d pick_r / d s[iz + ix*nz]
 * over the unpadded model (nz*nx doubles, overwritten; fp32 local derivatives,
 * fp64 accumulation over shots).  d_w: nrec_total floats.  Device pointers,
 * asynchronous on `stream`.  SEISFAT_ESTATE if there is no recorded run or the
 * slowness changed since it. */
int seisfat_batch_vjp_device(seisfat_batch_t b, const float* d_w, double* d_grad, void* stream);
/* seisfat_batch_vjp_device with host buffers (w: nrec_total floats, grad:
 * nz*nx doubles), synchronous: the products of the reference's tomography
 * with its ray matrix, G^T w (src/inversion/inversion.cpp:277-278), for the
 * exact Jacobian of the recorded solve; w = 1 gives the column sums of G, the
 * ray coverage (src/inversion/tomography_iso.cpp:34-38). */
int seisfat_batch_vjp(seisfat_batch_t b, const float* w, double* grad);
/* Misfit J = 1/2 sum_r (pick_r - obs_r)^2 and its gradient dJ/ds (nz*nx
 * doubles) at the handle's current slowness: run_record, residuals and the
 * vector-Jacobian product, host buffers, synchronous.  J and the residuals use
 * the picks of the recorded fp64 trajectory (the map the gradient is exact
 * for); picks (nullable) receives the batch's fp32 picks. */
int seisfat_batch_misfit_gradient(seisfat_batch_t b, const float* obs, double* J, double* grad, float* picks);

#ifdef __cplusplus
}
#endif
#endif /* SEISFAT2D_EIK_H */
