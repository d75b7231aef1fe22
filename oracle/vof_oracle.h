/*
 * vof_oracle.h -- CPU restatement of one SimpleVOFSolver time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product path (simplevofsolver_amd/) never does.
 *
 * Built twice from one source: -DNDIMS=2 (liboracle2d.so) follows the
 * reference's 2D code line by line; -DNDIMS=3 (liboracle3d.so) is the 3D
 * extrapolation (the reference snapshot has no 3D code).
 *
 * Array layouts are the reference's index maps (include/array_macros/,
 * tools/define_arrays.py:201-229): x fastest, then y, then z, each array with
 * its own halo padding.  All values are FP64.
 */
#ifndef VOF_ORACLE_H
#define VOF_ORACLE_H

#include <stddef.h>

typedef struct orc orc_t;

/* physical / numerical parameters (reference defaults in comments) */
typedef struct {
  double Ra, Pr, We;          /* config.get_double: fluid/init.c:90-93, interface/init.c:66-68 */
  double coef_dt_adv;         /* decide_dt.c:109 */
  double coef_dt_dif;         /* decide_dt.c:110 */
  int add_buoyancy;           /* param/buoyancy.c:3 (true) */
  double uy_xm, uy_xp;        /* param/boundary-condition.c:3-4 (0, 0) */
  double uz_xm, uz_xp;        /* 3D extrapolation (0, 0) */
  double t_xm, t_xp;          /* param/boundary-condition.c:5-6 (+0.5, -0.5) */
} orc_params_t;

/* lifecycle: domain_init (src/domain.c:170-209) from glsizes/lengths/xf/xc */
orc_t * orc_create(const size_t * glsizes, const double * lengths,
                   const double * xf, const double * xc);
void orc_destroy(orc_t * o);
int orc_ndims(void);
void orc_set_params(orc_t * o, const orc_params_t * prm);
void orc_get_params(const orc_t * o, orc_params_t * prm);

/* field access by name ("ux","uy","uz","p","psi","t","vof","ifrcx","ifrcy",
 * "ifrcz","dvof","normal","curv","flxx","flxy","flxz","src_a","src_b",
 * "srcux_a","srcux_b","srcux_g", ... ) -- whole padded arrays */
size_t orc_field_count(const orc_t * o, const char * name);
int orc_get_field(const orc_t * o, const char * name, double * out);
int orc_set_field(orc_t * o, const char * name, const double * in);
/* IC load as fluid_init / interface_init do: NPY-shaped interior-y(z) slabs
 * with x padding, then boundary conditions (fluid/init.c:80-88) */
int orc_load_ic(orc_t * o, const char * name, const double * slab);
int orc_dump_slab(const orc_t * o, const char * name, double * slab);
/* scalars */
double orc_m_dif(const orc_t * o);
double orc_t_dif(const orc_t * o);
double orc_tension(const orc_t * o);
int orc_x_grid_is_uniform(const orc_t * o);

/* solver entry points (names follow include/{interface,fluid}_solver.h of the reference) */
int orc_integrate(orc_t * o, double * dt);
int orc_decide_dt(orc_t * o, double * dt);
int orc_interface_compute_curvature_tensor(orc_t * o);
int orc_interface_compute_force(orc_t * o);
int orc_interface_update_vof(orc_t * o, int rkstep, double dt);
int orc_fluid_predict_field(orc_t * o, int rkstep, double dt);
int orc_fluid_compute_potential(orc_t * o, int rkstep, double dt, int use_dct);
int orc_fluid_correct_velocity(orc_t * o, int rkstep, double dt);
int orc_fluid_update_pressure(orc_t * o, int rkstep, double dt);
int orc_update_boundaries(orc_t * o, const char * name);

/* numerics exposed for pinning tests */
int orc_tdm_solve(int n, int nrhs, int periodic, const double * l,
                  const double * c, const double * u, double * q);
int orc_tdm_solve_complex(int n, int nrhs, int periodic, const double * l,
                          const double * c, const double * u, double * q_reim);
int orc_fft(int n, int sign, double * reim);
int orc_rfft(int n, const double * in, double * out_reim);
int orc_irfft(int n, const double * in_reim, double * out);
int orc_redft10(int n, const double * in, double * out);
int orc_redft01(int n, const double * in, double * out);
double orc_compute_intercept(double vof, const double * normal);
double orc_indicator(const double * normal_with_d, const double * x);

/* diagnostics (logging/divergence.c, logging/interface.c, logging/energy.c) */
double orc_max_divergence(const orc_t * o);
void orc_vof_stats(const orc_t * o, double * vmin, double * vmax, double * volume);
void orc_energy(const orc_t * o, double * out /* NDIMS + 1 */);
void orc_momentum(const orc_t * o, double * out /* NDIMS */);
double orc_interface_area(const orc_t * o);
void orc_nusselt(const orc_t * o, double * out /* 4 */);

/* twiddle table of the transforms (0 octant-reduced, the default; 1 direct
 * cos / sin): the same DFT up to rounding */
void orc_set_fft_variant(int v);

/* statistics accumulators (statistics.c) */
typedef struct orc_stats orc_stats_t;
orc_stats_t * orc_stats_create(const orc_t * o);
void orc_stats_destroy(orc_stats_t * s);
void orc_stats_collect(orc_stats_t * s, const orc_t * o);
size_t orc_stats_num(const orc_stats_t * s);
int orc_stats_reduce(const orc_stats_t * s, const char * name, double * out);

#endif
