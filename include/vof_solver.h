/*
 * vof_solver.h -- C ABI of the MI355X (gfx950) SimpleVOFSolver time step.
 *
 * Drop-in for the reference's solver interface: the same struct layouts and
 * entry-point names as the reference's include/ headers, compiled once per
 * dimensionality because the structs are NDIMS-sized:
 *   libvofhip2d.so  (-DNDIMS=2)   libvofhip3d.so  (-DNDIMS=3)
 *
 * Differences a caller must know:
 *  - every array_t.data is a DEVICE pointer (hipMalloc'd by *_init) with the
 *    reference's index map (include/array_macros/ of the reference;
 *    tools/define_arrays.py:201-229 for 3D).  Use vof_array_to_host /
 *    vof_array_from_host to move whole padded arrays.
 *  - domain_t.info, an sdecomp_info_t* in the reference, is an opaque solver
 *    context here (grid metrics on the device, scratch buffers, HIP stream).
 *  - one process, one GPU: no MPI, no domain decomposition (mysizes ==
 *    glsizes, offsets == 0).
 *  - return codes: 0 = success, nonzero = failure; vof_last_error() gives
 *    the message (the reference prints to stdout/stderr and returns 1).
 *  - integrate() and the fluid_* / interface_* steps swap array_t.data
 *    pointers of the RK source terms exactly like the reference
 *    (predict/main.c:9-29, interface/update/main.c:28-45); integrate() may
 *    also swap interface->vof.data with an internal buffer.  Do not cache
 *    data pointers across calls.
 *  - all work is queued on the context's HIP stream; entry points that hand
 *    values to the host (fluid_decide_dt, integrate, *_save,
 *    vof_array_to_host) synchronise it.
 */
#ifndef VOF_SOLVER_H
#define VOF_SOLVER_H

#include <stddef.h>
#include <stdbool.h>

#if !defined(NDIMS)
#error "define NDIMS as 2 or 3 before including vof_solver.h"
#endif
#if NDIMS != 2 && NDIMS != 3
#error "NDIMS must be 2 or 3"
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- structs ------------------------------------------------------------ */

/* replaces include/domain.h:19-28 (3D adds dz after dy) */
typedef struct {
  void * info;                 /* opaque solver context (was sdecomp_info_t *) */
  size_t glsizes[NDIMS];
  size_t mysizes[NDIMS];
  size_t offsets[NDIMS];
  double lengths[NDIMS];
  double * xf, * xc;           /* host copies, as the reference keeps them */
  double * dxf, * dxc;         /* host copies */
  double dy;
#if NDIMS == 3
  double dz;
#endif
} domain_t;

/* replaces include/array.h:10-23; data is a device pointer */
typedef struct {
  size_t size;
  int (* nadds)[2];
  size_t datasize;
  void * data;
} array_t;

/* replaces include/fluid.h:20-31 (3D adds uz after uy, srcuz after srcuy) */
typedef struct {
  array_t ux;
  array_t uy;
#if NDIMS == 3
  array_t uz;
#endif
  array_t p;
  array_t psi;
  array_t t;
  array_t srcux[3];
  array_t srcuy[3];
#if NDIMS == 3
  array_t srcuz[3];
#endif
  array_t srct[3];
  double Ra, Pr;
  double m_dif, t_dif;
} fluid_t;

/* replaces include/interface.h:13-24 (3D adds ifrcz, flxz) */
typedef double vector_t[NDIMS];
typedef double normal_t[NDIMS + 1];
typedef struct {
  array_t vof;
  array_t ifrcx;
  array_t ifrcy;
#if NDIMS == 3
  array_t ifrcz;
#endif
  array_t dvof;
  array_t normal;
  array_t curv;
  array_t flxx;
  array_t flxy;
#if NDIMS == 3
  array_t flxz;
#endif
  array_t src[2];
  double tension;
} interface_t;

/* ---- domain (replaces include/domain.h:30-50) ---------------------------- */

/* src/domain.c:170-209: loads glsizes/lengths/xf/xc NPY, derives dxf, dxc,
 * dy (dz), and creates the device context */
int domain_init(const char dirname_ic[], domain_t * domain);
/* src/domain.c:52-73 */
int domain_save(const char dirname[], const domain_t * domain);
/* src/domain.c:75-98 */
int domain_check_x_grid_is_uniform(const domain_t * domain, bool * x_grid_is_uniform);

/* ---- fluid (replaces include/fluid_solver.h:10-99) ----------------------- */

/* src/fluid/init.c:72-96: allocates device arrays, loads ux uy (uz) p t,
 * applies boundary conditions, reads Pr and Ra (environment, as
 * config.get_double; vof_set_param overrides) */
int fluid_init(const char dirname_ic[], const domain_t * domain, fluid_t * fluid);
/* src/fluid/save.c:8-27 */
int fluid_save(const char dirname[], const domain_t * domain, const fluid_t * fluid);
/* src/fluid/decide_dt.c:104-145 (reads coef_dt_adv, coef_dt_dif) */
int fluid_decide_dt(const domain_t * domain, const fluid_t * fluid, double * dt);
/* src/fluid/predict/main.c:62-79 */
int fluid_predict_field(const domain_t * domain, const size_t rkstep, const double dt,
                        fluid_t * fluid, const interface_t * interface);
/* src/fluid/compute_potential/dft.c:406-468 */
int fluid_compute_potential_dft(const domain_t * domain, const size_t rkstep,
                                const double dt, fluid_t * fluid);
/* src/fluid/compute_potential/dct.c:373-421; the same discrete Poisson
 * problem, solved by the same FFT-in-y(z) / Thomas-in-x kernels */
int fluid_compute_potential_dct(const domain_t * domain, const size_t rkstep,
                                const double dt, fluid_t * fluid);
/* src/fluid/correct_velocity/main.c:15-27 */
int fluid_correct_velocity(const domain_t * domain, const size_t rkstep,
                           const double dt, fluid_t * fluid);
/* src/fluid/update_pressure.c:87-109 */
int fluid_update_pressure(const domain_t * domain, const size_t rkstep,
                          const double dt, fluid_t * fluid);
/* src/fluid/boundary/{ux,uy,p,psi,t}.c */
int fluid_update_boundaries_ux(const domain_t * domain, array_t * ux);
int fluid_update_boundaries_uy(const domain_t * domain, array_t * uy);
#if NDIMS == 3
int fluid_update_boundaries_uz(const domain_t * domain, array_t * uz);
#endif
int fluid_update_boundaries_p(const domain_t * domain, array_t * p);
int fluid_update_boundaries_psi(const domain_t * domain, array_t * psi);
int fluid_update_boundaries_t(const domain_t * domain, array_t * t);

/* ---- interface (replaces include/interface_solver.h:8-41) ---------------- */

/* src/interface/init.c:55-71 (reads We) */
int interface_init(const char dirname_ic[], const domain_t * domain, interface_t * interface);
/* src/interface/save.c:7-22 */
int interface_save(const char dirname[], const domain_t * domain, const interface_t * interface);
/* src/interface/curvature_tensor.c:178-186 */
int interface_compute_curvature_tensor(const domain_t * domain, interface_t * interface);
/* src/interface/force.c:69-76 */
int interface_compute_force(const domain_t * domain, interface_t * interface);
/* src/interface/update/main.c:112-126 */
int interface_update_vof(const domain_t * domain, const size_t rkstep, const double dt,
                         const fluid_t * fluid, interface_t * interface);
/* src/interface/boundary.c:27-39 */
int interface_update_boundaries_vof(const domain_t * domain, array_t * vof);

/* ---- integrator (replaces include/integrate.h:5-10) ---------------------- */

/* src/integrate.c:11-68: one RK3 time step; dt is returned on the host */
int integrate(const domain_t * domain, fluid_t * fluid, interface_t * interface, double * dt);

/* ---- tri-diagonal solver (replaces include/tdm.h:6-45) ------------------- */

/* src/tdm.c:147-324.  l, c, u are HOST arrays owned by the info object
 * (filled by the caller through get_l/get_c/get_u, as in the reference);
 * solve() takes a DEVICE pointer to nrhs contiguous lines of `size` values
 * (double, or interleaved (re, im) pairs when is_complex), runs the GPU
 * Thomas / Sherman-Morrison kernels on the object's own stream (blocking
 * w.r.t. the null stream) and returns when the answers are in place, bit for
 * bit the reference's.  construct() also rejects sizes the reference's loops
 * would read out of bounds with: size < 2, or size < 3 when periodic. */
typedef struct tdm_info_t_ tdm_info_t;
typedef struct {
  int (* const construct)(const int size, const int nrhs, const bool is_periodic,
                          const bool is_complex, tdm_info_t ** info);
  int (* const get_l)(const tdm_info_t * info, double ** l);
  int (* const get_c)(const tdm_info_t * info, double ** c);
  int (* const get_u)(const tdm_info_t * info, double ** u);
  int (* const get_size)(const tdm_info_t * info, int * size);
  int (* const get_nrhs)(const tdm_info_t * info, int * nrhs);
  int (* const solve)(const tdm_info_t * info, void * data);
  int (* const destruct)(tdm_info_t * info);
} tdm_t;
extern const tdm_t tdm;
/* tdm.solve without device allocation.  No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  l, c, u: DEVICE arrays of `size`
 * coefficients; data: as tdm.solve (device, nrhs lines, real or interleaved
 * complex), solved in place with the same kernels, so the answers are those of
 * tdm.solve bit for bit.  The factor table lives in a fixed device array of
 * the library (one per GPU, held until the process ends): size must be at
 * most 16384, larger sizes fail with an error.  stream: a hipStream_t (NULL =
 * the null stream); the call enqueues and returns without waiting.  Calls
 * that overlap on one GPU race on that table. */
int tdm_solve_noalloc(int size, int nrhs, bool is_periodic, bool is_complex, const double * l,
                      const double * c, const double * u, void * data, void * stream);

/* ---- linear system (replaces include/linear_system.h:37-46) -------------- */

/* src/linear_system.c:28-85.  x1pncl is a device buffer of prod(glsizes)
 * doubles; tdm_x is the non-periodic x solver (implicit[0]).  Implicit
 * treatment in y/z is not provided (the reference's default
 * param/implicit.c:4,7 is explicit) and returns an error. */
typedef struct {
  bool is_initialised;
  bool implicit[NDIMS];
  double * x1pncl;
  double * y1pncl;
  size_t x1pncl_mysizes[NDIMS];
  size_t y1pncl_mysizes[NDIMS];
  tdm_info_t * tdm_x;
  tdm_info_t * tdm_y;
  void * transposer_x1_to_y1;
  void * transposer_y1_to_x1;
} linear_system_t;
int linear_system_init(const void * info, const bool implicit[NDIMS],
                       const size_t glsizes[NDIMS], linear_system_t * linear_system);
int linear_system_finalise(linear_system_t * linear_system);

/* ---- extensions (no counterpart in the reference) ------------------------ */

/* last error message of this thread ("" if none) */
const char * vof_last_error(void);
/* override a run-time parameter instead of the environment variable the
 * reference reads: "Ra", "Pr", "We", "coef_dt_adv", "coef_dt_dif", or one
 * of the reference's compile-time constants (src/param/ files):
 * "add_buoyancy" (0/1), "uy_xm", "uy_xp", "uz_xm", "uz_xp", "t_xm", "t_xp".
 * Must be called after domain_init and before fluid_init / interface_init
 * for Ra, Pr, We. */
int vof_set_param(domain_t * domain, const char name[], double value);
int vof_get_param(const domain_t * domain, const char name[], double * value);
/* whole padded arrays (datasize bytes) between device and host */
int vof_array_to_host(const domain_t * domain, const array_t * array, void * host);
int vof_array_from_host(const domain_t * domain, array_t * array, const void * host);
/* wait for all queued work of this domain */
int vof_synchronize(const domain_t * domain);
/* device buffers for callers of tdm.solve and linear_system_init (whose
 * data are device pointers here) that hold no HIP runtime of their own.
 * vof_device_copy is synchronous; kind 1 = host to device, 2 = device to
 * host, 3 = device to device. */
int vof_device_alloc(void ** ptr, size_t bytes);
int vof_device_free(void * ptr);
int vof_device_copy(void * dst, const void * src, size_t bytes, int kind);
/* free device memory (the reference never frees) */
int interface_finalise(const domain_t * domain, interface_t * interface);
int fluid_finalise(const domain_t * domain, fluid_t * fluid);
int domain_finalise(domain_t * domain);
/* select the per-step kernel schedule of integrate():
 * 0 = one launch sequence per reference function (the per-function entry
 *     points), 1 = fused kernels (default).  Both give the same fields; the
 *     fused schedule does not refresh the intermediate arrays dvof, flxx,
 *     flxy, flxz, ifrcx, ifrcy, ifrcz and srcg, does not store the last RK
 *     stage's explicit sources (srcux[0], srcuy[0], srcuz[0], srct[0],
 *     interface src[0]: the next step's first stage has beta = 0 and never
 *     reads them) and skips the beta-weighted source reads of the first
 *     stage (their weight is 0). */
int vof_set_schedule(domain_t * domain, int schedule);
/* stages fused by schedule 1 (bit mask, default 63): 1 curvature
 * tensor, 2 VOF update, 4 sources with the surface force, 8 boundary images
 * of the predictor, 16 boundary images of the potential, 32 correction with
 * pressure, 64 the whole predictor in one kernel (sources, x solve and
 * update with images; needs 2, 4 and 8, and x lines that fit the kernel's
 * LDS: nx <= 2048).  With 64 the predictor writes the new ux, uy, uz, t into
 * second buffers and swaps them in: like the src arrays, the data pointers
 * of these fields change between calls. */
int vof_set_fusion(domain_t * domain, int mask);
/* device diagnostics (logging/divergence.c:37-55, logging/interface.c:29-51,
 * logging/energy.c:43-65, logging/momentum.c:39-53): out = {max|div u|,
 * vof min, vof max, vof volume fraction, kinetic energy x, y[, z], thermal
 * energy, momentum x, y[, z], interface measure sum |grad vof| dV}
 * (at most 12 values) */
int vof_diagnostics(const domain_t * domain, const fluid_t * fluid,
                    const interface_t * interface, double * out);

/* Nusselt numbers (logging/nusselt/main.c:18-50): out = {wall heat flux,
 * kinetic energy injection, kinetic energy dissipation, thermal
 * dissipation}; 3D adds the z terms and normalises by t_dif * ly * lz */
int vof_nusselt(const domain_t * domain, const fluid_t * fluid, double out[4]);

/* ---- statistics (replaces include/statistics.h:7-30) ---------------------
 * init reads stat_rate and stat_after from the environment; collect adds
 * one sample of ux, ux^2, uy, uy^2, t, t^2, ux t (3D: uz, uz^2) on the
 * device; output writes output/stat/step%010zu with the y (and z) sums as
 * 1D NPY profiles, num, m_dif, t_dif and the domain */
typedef struct {
  int (* const init)(const domain_t * domain, const double time);
  int (* const collect)(const domain_t * domain, const fluid_t * fluid);
  int (* const output)(const domain_t * domain, const size_t step);
  double (* const get_next_time)(void);
} statistics_t;
extern const statistics_t statistics;
/* statistics.output into any directory */
int vof_statistics_output_to(const domain_t * domain, const char dirname[]);

/* Poisson precision: refinements < 0 (default) is the plain FP64 solver;
 * refinements >= 0 transforms in FP32 (FP32 spectrum, FP64 x eliminations)
 * and then applies that many FP64 residual corrections psi += M(r - A psi) */
int vof_set_poisson_mixed(domain_t * domain, int refinements);

/* ---- batches: many independent cases advanced together -----------------
 * The cases are initialised one by one with the entry points above (own
 * sizes, parameters and grids; fused schedule) and then stepped together:
 * each kernel of the step runs as one launch covering every case's
 * work-groups, with results bit-identical to integrate() on each case.  The
 * step is asynchronous on the batch's stream: vof_batch_synchronize before
 * touching a case through any other entry point.  A batch of 8 cases or more
 * is run as 2 batches of alternate cases (4 from 16 cases on: case c in
 * part c % 4), each on its own stream, stepped in turn by
 * vof_batch_integrate: the host's work of one part (reading back its dt,
 * recording its first stage) then overlaps the other parts' kernels.  No
 * counterpart in the reference (one case per process there). */
typedef struct vof_batch vof_batch_t;
int vof_batch_create(int ncases, const domain_t * const domains[], fluid_t * const fluids[],
                     interface_t * const interfaces[], vof_batch_t ** batch);
/* one time step of every case; dts[c] receives case c's dt (may be NULL) */
int vof_batch_integrate(vof_batch_t * batch, double dts[]);
/* hipGraph replay of the batched step (on = 1, the default): each phase of
 * the step (dt, then the three RK stages), with its argument-table and
 * Thomas-table copies and its merged launches, is captured into a graph
 * once and replayed while its launch layout is unchanged (re-captured when
 * it changes, e.g. after tuning); on = 0 issues the launches directly.
 * Results are the same either way. */
int vof_batch_set_graphs(vof_batch_t * batch, int on);
/* graphs captured and graph launches issued so far (both parts of a split
 * batch together) */
int vof_batch_graph_counts(const vof_batch_t * batch, size_t * captures, size_t * launches);
int vof_batch_synchronize(vof_batch_t * batch);
int vof_batch_destroy(vof_batch_t * batch);

/* one time step with a given dt instead of fluid_decide_dt (fused schedule):
 * the frozen dt sequence of the adjoint and its finite-difference checks */
int vof_integrate_fixed_dt(const domain_t * domain, fluid_t * fluid, interface_t * interface, double dt);
/* discrete decisions of the 2D fused step, for finite-difference checks of
 * the adjoint: mode 1 records, RK stage by RK stage from the next step on,
 * the upwind side and single-phase test of every face flux, the interface
 * test and Newton update count of every cell's intercept, and which corner
 * gradients lie below the adjoint's floor (with their normals); mode 2
 * imposes the recorded decisions on the following stages, again from the
 * first recorded one (so a perturbed run follows the map the adjoint
 * differentiates); 0 stops.  2D only (3D: an error for modes 1 and 2).  No
 * counterpart in the reference. */
int vof_set_freeze(domain_t * domain, int mode);

#if NDIMS == 2
/* ---- adjoint (2D): gradient of a loss after nsteps w.r.t. the initial
 * velocity.  Loss = droplet centroid x (sum xc vof dV / sum vof dV) +
 * interface length (sum |grad vof| dV, central differences, with |grad vof|
 * regularised as sqrt(|g|^2 + g0^2) - g0, g0 = 1e-6 / (2 dx)).  The map differentiated is the step with its
 * discrete decisions held (the ones vof_set_freeze records) and the normals
 * of flat corners (|grad vof| DXC below the floor, "adj_floor_exp") held
 * constant.  The forward
 * runs dts[0..nsteps-1] from the current state, keeping a checkpoint every
 * `every` steps; the reverse sweep recomputes each segment.  grad_ux /
 * grad_uy are host arrays of the padded ux / uy shapes (values on the
 * velocity DOFs, zero elsewhere).  The solver is left at its initial state.
 * No counterpart in the reference. */
int vof_adjoint_gradient(const domain_t * domain, fluid_t * fluid, interface_t * interface, int nsteps,
                         const double * dts, int every, double * loss, double * grad_ux, double * grad_uy);
/* the same loss of the current state */
int vof_adjoint_loss(const domain_t * domain, const interface_t * interface, double * loss);
#endif

/* autotuner: picks "fft_radix", "fft_elems", "iface_planes" (3D),
 * "pred_rows", "rhs_rows", "iface_overlap", "thomas_rows" and "lines_lpb" of
 * vof_set_tuning by timing `steps` steps of
 * this case per candidate (coordinate descent in that order; a candidate
 * must be 1 % faster to replace the current value).  The case's device state
 * is copied first and restored after every trial, so it is left as it was,
 * with the fastest parameters set.  report (may be NULL) receives one
 * "name value ms_per_step" line per trial and a "best name value ms" line
 * per parameter.  Fused schedule only.  No counterpart in the reference. */
int vof_autotune(domain_t * domain, fluid_t * fluid, interface_t * interface, int steps, char * report,
                 size_t len);
/* the same over a batch: every candidate is applied to all cases and timed
 * over `steps` calls of vof_batch_integrate */
int vof_batch_autotune(vof_batch_t * batch, int steps, char * report, size_t len);

/* tile parameters (results do not depend on them): "fft_elems" = 1024 or
 * 2048 complex values per LDS buffer of the FFT kernels, "lines_lpb" = lines
 * per work-group of the predictor's x solver (0 = automatic); "xcd_bands" =
 * 1 walks the 3D stencil tiles band by band per XCD; "rhs_rows" = 1, 2, 4,
 * 8 (default) or 16 rows of one column per thread of the source kernel;
 * "pred_rows" = 1, 2, 4 or 8 (default) rows per thread of the one-kernel
 * predictor (fewer where its x lines would not fit 64 KiB of LDS);
 * "iface_planes" = 0 (blocks), 4, 8 (default) or 16 z planes marched per
 * work-group by the 3D interface kernel;
 * "iface_overlap" = 1 runs the interface kernel of RK stages 2 and 3 on a
 * second stream from the end of the previous stage's source kernel (beside
 * its x solve, Poisson solve and correction), 2 from the start of its
 * Thomas pass, 0 on one stream (default: 1 from 2^20 cells on, else 0);
 * "thomas_rows" = 8 (default), 16 or 32 rows of the spectrum prefetched per
 * lane by the Poisson solver's sequential x eliminations;
 * "thomas_part" = 1 (default) solves the Poisson solver's x systems by
 * partitions (segments of 2 .. 8 rows, one lane each, and a reduced system
 * of the segments' end rows) where 4 <= nx <= 256, 0 in sequence (this one
 * changes results by rounding only, a few 1e-14 relative to each line);
 * "fft_radix" = 16 (default: 256-point transforms on the register
 * radix-16 path) or 4 (every length on the mixed-radix 4/2/3/5 path; this
 * one changes results by rounding only, about 1e-16 relative);
 * "adj_floor_exp" = e sets the adjoint's corner-gradient floor to 10^-e
 * (default 2: corners with |grad vof| DXC < 1e-2 are held constant)
 * (this one changes adjoint results) */
int vof_set_tuning(domain_t * domain, const char name[], int value);
/* step.npy / time.npy of an IC or save directory (the reference's main.c
 * reads and writes them with fileio.r_serial / w_serial) */
int vof_read_step_time(const char dirname[], size_t * step, double * time);
int vof_write_step_time(const char dirname[], size_t step, double time);
/* per-kernel timing with HIP events recorded on the context's stream around
 * every launch (on = 1 starts, on = 0 stops); the report lists
 * "kernel launches total_ms" lines */
int vof_profile(domain_t * domain, int on);
int vof_profile_reset(domain_t * domain);
int vof_profile_report(domain_t * domain, char * buf, size_t len);
/* copy of n doubles repeated reps times on the device: mean read+write
 * bandwidth in GB/s (calibrates hardware byte counters, kernel k_copy) */
int vof_copy_probe(domain_t * domain, size_t n, int reps, double * gbps);

#ifdef __cplusplus
}
#endif

#endif /* VOF_SOLVER_H */
