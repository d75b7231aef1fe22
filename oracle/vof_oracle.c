/*
 * vof_oracle.c -- CPU restatement of one SimpleVOFSolver time step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
 * The product path never links or calls this file.
 *
 * Every function cites the reference file:line it restates.  Floating-point
 * operations follow the reference's evaluation order (compiled with
 * -std=c99, i.e. no FMA contraction, like the reference's Makefile:2).
 * NDIMS=2 is a restatement of the reference; NDIMS=3 is the extrapolation
 * documented beside each z term (no 3D reference code exists).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include "vof_oracle.h"

#if !defined(NDIMS)
#error "NDIMS must be 2 or 3"
#endif

typedef double vector_t[NDIMS];
typedef double normal_t[NDIMS + 1];

struct orc {
  int isize, jsize, ksize;           /* interior cells (ksize = 1 in 2D) */
  double lengths[3];
  double * xf, * xc, * dxf, * dxc;   /* src/domain.c:21-44,103-143 */
  double dy, dz;
  int x_uniform;
  orc_params_t prm;
  double m_dif, t_dif, tension;
  /* fluid_t (include/fluid.h:20-31) */
  double * ux, * uy, * uz, * p, * psi, * t;
  double * srcux[3], * srcuy[3], * srcuz[3], * srct[3];
  /* interface_t (include/interface.h:13-24) */
  double * vof, * ifrcx, * ifrcy, * ifrcz, * curv, * flxx, * flxy, * flxz;
  vector_t * dvof;
  normal_t * normal;
  double * src[2];
  /* solver scratch (the reference keeps these in function statics) */
  double * lapx_ux, * lapx_c;        /* [nx-1][3], [nx][3]: predict/ux.c:35-62, uy.c:35-62 */
  double * x1pncl;                   /* linear_system.c:63 */
  double * pbuf0, * pbuf1;           /* compute_potential/{dct,dft}.c allocate_buffers */
};

/* ---------------------------------------------------------------------- */
/* index maps: include/array_macros/ (2D) and tools/define_arrays.py:201-229 */
#if NDIMS == 2
#define ARR(a, I, J, K, i0, n0, j0, n1, k0) \
  (a)[(size_t)((I) - (i0)) + (size_t)(n0) * (size_t)((J) - (j0))]
#else
#define ARR(a, I, J, K, i0, n0, j0, n1, k0) \
  (a)[(size_t)((I) - (i0)) + (size_t)(n0) * ((size_t)((J) - (j0)) + (size_t)(n1) * (size_t)((K) - (k0)))]
#endif
#define UX(I, J, K)     ARR(ux,     I, J, K,  1, isize + 1,  0, jsize + 2,  0)
#define UY(I, J, K)     ARR(uy,     I, J, K,  0, isize + 2,  0, jsize + 2,  0)
#define UZ(I, J, K)     ARR(uz,     I, J, K,  0, isize + 2,  0, jsize + 2,  0)
#define P(I, J, K)      ARR(p,      I, J, K,  0, isize + 2,  0, jsize + 2,  0)
#define PSI(I, J, K)    ARR(psi,    I, J, K,  0, isize + 2,  0, jsize + 2,  0)
#define T(I, J, K)      ARR(t,      I, J, K,  0, isize + 2,  0, jsize + 2,  0)
#define CURV(I, J, K)   ARR(curv,   I, J, K,  0, isize + 2,  0, jsize + 2,  0)
#define VOF(I, J, K)    ARR(vof,    I, J, K,  0, isize + 2, -1, jsize + 4, -1)
#define DVOF(I, J, K)   ARR(dvof,   I, J, K,  1, isize + 1,  0, jsize + 3,  0)
#define NORMAL(I, J, K) ARR(normal, I, J, K,  1, isize,      0, jsize + 2,  0)
#define FLXX(I, J, K)   ARR(flxx,   I, J, K,  1, isize + 1,  1, jsize,      1)
#define FLXY(I, J, K)   ARR(flxy,   I, J, K,  1, isize,      1, jsize + 1,  1)
#define FLXZ(I, J, K)   ARR(flxz,   I, J, K,  1, isize,      1, jsize,      1)
#define SRC(I, J, K)    ARR(src,    I, J, K,  1, isize,      1, jsize,      1)
#define IFRCX(I, J, K)  ARR(ifrcx,  I, J, K,  2, isize - 1,  1, jsize,      1)
#define IFRCY(I, J, K)  ARR(ifrcy,  I, J, K,  1, isize,      1, jsize,      1)
#define IFRCZ(I, J, K)  ARR(ifrcz,  I, J, K,  1, isize,      1, jsize,      1)
#define DXF(I) (dxf[(I) - 1])
#define DXC(I) (dxc[(I) - 1])
#define XF(I)  (xf[(I) - 1])
#define XC(I)  (xc[(I)])

#if NDIMS == 2
#define KLOOP(lo, hi) for (int k = 1; k <= 1; k++)
#else
#define KLOOP(lo, hi) for (int k = (lo); k <= (hi); k++)
#endif

#define SIZES \
  const int isize = o->isize; const int jsize = o->jsize; \
  const int ksize = o->ksize; (void)ksize; (void)jsize

/* sizes of each padded array: {x width, y extent, z extent, components} */
typedef struct { int n0, n1, n2, ncomp; int nhy; } shape_t;

static shape_t shape_of(const orc_t * o, const char * name, int * ok) {
  const int nx = o->isize, ny = o->jsize;
  const int nz = NDIMS == 3 ? o->ksize : 1;
  const int pz1 = NDIMS == 3 ? 2 : 0;  /* z padding for 1-halo arrays */
  shape_t s = {0, 0, 0, 1, 0};
  *ok = 1;
  if (!strcmp(name, "ux")) { s.n0 = nx + 1; s.n1 = ny + 2; s.n2 = nz + pz1; s.nhy = 1; }
  else if (!strcmp(name, "uy") || !strcmp(name, "uz") || !strcmp(name, "p") ||
           !strcmp(name, "psi") || !strcmp(name, "t") || !strcmp(name, "curv")) {
    s.n0 = nx + 2; s.n1 = ny + 2; s.n2 = nz + pz1; s.nhy = 1;
  }
  else if (!strcmp(name, "vof")) { s.n0 = nx + 2; s.n1 = ny + 4; s.n2 = NDIMS == 3 ? nz + 4 : 1; s.nhy = 2; }
  else if (!strcmp(name, "dvof")) { s.n0 = nx + 1; s.n1 = ny + 3; s.n2 = NDIMS == 3 ? nz + 3 : 1; s.ncomp = NDIMS; }
  else if (!strcmp(name, "normal")) { s.n0 = nx; s.n1 = ny + 2; s.n2 = nz + pz1; s.ncomp = NDIMS + 1; }
  else if (!strcmp(name, "flxx")) { s.n0 = nx + 1; s.n1 = ny; s.n2 = nz; }
  else if (!strcmp(name, "flxy")) { s.n0 = nx; s.n1 = ny + 1; s.n2 = nz; }
  else if (!strcmp(name, "flxz")) { s.n0 = nx; s.n1 = ny; s.n2 = NDIMS == 3 ? nz + 1 : 1; }
  else if (!strcmp(name, "ifrcx") || !strncmp(name, "srcux", 5)) { s.n0 = nx - 1; s.n1 = ny; s.n2 = nz; }
  else if (!strcmp(name, "ifrcy") || !strcmp(name, "ifrcz") || !strncmp(name, "srcuy", 5) ||
           !strncmp(name, "srcuz", 5) || !strncmp(name, "srct", 4) || !strncmp(name, "src_", 4)) {
    s.n0 = nx; s.n1 = ny; s.n2 = nz;
  }
  else { *ok = 0; }
  if (NDIMS == 2 && (!strcmp(name, "uz") || !strcmp(name, "ifrcz") || !strcmp(name, "flxz") ||
                     !strncmp(name, "srcuz", 5))) {
    *ok = 0;
  }
  return s;
}

static double * field_ptr(const orc_t * o, const char * name) {
  if (!strcmp(name, "ux")) return o->ux;
  if (!strcmp(name, "uy")) return o->uy;
  if (!strcmp(name, "uz")) return o->uz;
  if (!strcmp(name, "p")) return o->p;
  if (!strcmp(name, "psi")) return o->psi;
  if (!strcmp(name, "t")) return o->t;
  if (!strcmp(name, "vof")) return o->vof;
  if (!strcmp(name, "curv")) return o->curv;
  if (!strcmp(name, "dvof")) return (double *)o->dvof;
  if (!strcmp(name, "normal")) return (double *)o->normal;
  if (!strcmp(name, "flxx")) return o->flxx;
  if (!strcmp(name, "flxy")) return o->flxy;
  if (!strcmp(name, "flxz")) return o->flxz;
  if (!strcmp(name, "ifrcx")) return o->ifrcx;
  if (!strcmp(name, "ifrcy")) return o->ifrcy;
  if (!strcmp(name, "ifrcz")) return o->ifrcz;
  if (!strcmp(name, "src_a")) return o->src[0];
  if (!strcmp(name, "src_b")) return o->src[1];
  static const char * const tags[3] = {"_a", "_b", "_g"};
  for (int n = 0; n < 3; n++) {
    char buf[16];
    snprintf(buf, sizeof buf, "srcux%s", tags[n]); if (!strcmp(name, buf)) return o->srcux[n];
    snprintf(buf, sizeof buf, "srcuy%s", tags[n]); if (!strcmp(name, buf)) return o->srcuy[n];
    snprintf(buf, sizeof buf, "srcuz%s", tags[n]); if (!strcmp(name, buf)) return o->srcuz[n];
    snprintf(buf, sizeof buf, "srct%s", tags[n]);  if (!strcmp(name, buf)) return o->srct[n];
  }
  return NULL;
}

size_t orc_field_count(const orc_t * o, const char * name) {
  int ok = 0;
  shape_t s = shape_of(o, name, &ok);
  if (!ok) return 0;
  return (size_t)s.n0 * s.n1 * s.n2 * s.ncomp;
}

int orc_get_field(const orc_t * o, const char * name, double * out) {
  const double * a = field_ptr(o, name);
  const size_t n = orc_field_count(o, name);
  if (!a || !n) return 1;
  memcpy(out, a, n * sizeof(double));
  return 0;
}

int orc_set_field(orc_t * o, const char * name, const double * in) {
  double * a = field_ptr(o, name);
  const size_t n = orc_field_count(o, name);
  if (!a || !n) return 1;
  memcpy(a, in, n * sizeof(double));
  return 0;
}

int orc_ndims(void) { return NDIMS; }

/* ---------------------------------------------------------------------- */
/* runge_kutta.c:8-15 */
static const double rkcoefs[3][3] = {
  {+32. / 60., 0. / 60., +32. / 60. + 0. / 60.},
  {+25. / 60., -17. / 60., +25. / 60. + -17. / 60.},
  {+45. / 60., -25. / 60., +45. / 60. + -25. / 60.},
};
enum { rk_a = 0, rk_b = 1, rk_g = 2 };

/* interface/common.c:4-14 */
#define NGAUSS 2
static const double gauss_ps[NGAUSS] = {-0.2886751345948129, +0.2886751345948129};
static const double gauss_ws[NGAUSS] = {+0.5, +0.5};
static const double vofbeta = 1.;
static const double vofmin = 1.e-8;

static const double g_pi = 3.14159265358979324;  /* compute_potential/dct.c:20 */

/* ---------------------------------------------------------------------- */
/* halo exchange, single process: src/halo.c:18-76 restricted to the periodic
 * neighbour being this rank.  Rows are copied with their x padding; in 3D the
 * y exchange runs over interior z planes, then whole padded x-y planes are
 * exchanged in z, so edges and corners hold their periodic images. */
static void halo_periodic(double * a, int n0, int ny, int nz, int nh) {
  const size_t row = (size_t)n0;
#if NDIMS == 2
  (void)nz;
  /* send to positive, receive from negative: rows [ny, ny+nh) -> [0, nh) */
  memcpy(a, a + row * ny, sizeof(double) * row * nh);
  /* send to negative, receive from positive: rows [nh, 2nh) -> [ny+nh, ny+2nh) */
  memcpy(a + row * (ny + nh), a + row * nh, sizeof(double) * row * nh);
#else
  const size_t plane = row * (size_t)(ny + 2 * nh);
  for (int kk = nh; kk < nz + nh; kk++) {
    double * b = a + plane * kk;
    memcpy(b, b + row * ny, sizeof(double) * row * nh);
    memcpy(b + row * (ny + nh), b + row * nh, sizeof(double) * row * nh);
  }
  memcpy(a, a + plane * nz, sizeof(double) * plane * nh);
  memcpy(a + plane * (nz + nh), a + plane * nh, sizeof(double) * plane * nh);
#endif
}

/* fluid/boundary/ux.c:6-39 */
static void bc_ux(orc_t * o) {
  SIZES; double * ux = o->ux;
  halo_periodic(ux, isize + 1, jsize, ksize, 1);
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) {
    UX(1, j, k) = 0.;
    UX(isize + 1, j, k) = 0.;
  }
}

/* fluid/boundary/uy.c:7-40 */
static void bc_uy(orc_t * o) {
  SIZES; double * uy = o->uy;
  halo_periodic(uy, isize + 2, jsize, ksize, 1);
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) {
    UY(0, j, k) = o->prm.uy_xm;
    UY(isize + 1, j, k) = o->prm.uy_xp;
  }
}

#if NDIMS == 3
/* 3D extrapolation of fluid/boundary/uy.c for uz */
static void bc_uz(orc_t * o) {
  SIZES; double * uz = o->uz;
  halo_periodic(uz, isize + 2, jsize, ksize, 1);
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) {
    UZ(0, j, k) = o->prm.uz_xm;
    UZ(isize + 1, j, k) = o->prm.uz_xp;
  }
}
#endif

/* fluid/boundary/p.c:6-39 and psi.c:6-39 (Neumann) */
static void bc_neumann(orc_t * o, double * p) {
  SIZES;
  halo_periodic(p, isize + 2, jsize, ksize, 1);
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) {
    P(0, j, k) = P(1, j, k);
    P(isize + 1, j, k) = P(isize, j, k);
  }
}

/* fluid/boundary/t.c:7-40 */
static void bc_t(orc_t * o) {
  SIZES; double * t = o->t;
  halo_periodic(t, isize + 2, jsize, ksize, 1);
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) {
    T(0, j, k) = o->prm.t_xm;
    T(isize + 1, j, k) = o->prm.t_xp;
  }
}

/* interface/boundary.c:7-39: 2 halo layers, walls zero on every padded row */
static void bc_vof(orc_t * o) {
  SIZES; double * vof = o->vof;
  halo_periodic(vof, isize + 2, jsize, ksize, 2);
  KLOOP(-1, ksize + 2) for (int j = -1; j <= jsize + 2; j++) {
    VOF(0, j, k) = 0.;
    VOF(isize + 1, j, k) = 0.;
  }
}

int orc_update_boundaries(orc_t * o, const char * name) {
  if (!strcmp(name, "ux")) bc_ux(o);
  else if (!strcmp(name, "uy")) bc_uy(o);
#if NDIMS == 3
  else if (!strcmp(name, "uz")) bc_uz(o);
#endif
  else if (!strcmp(name, "p")) bc_neumann(o, o->p);
  else if (!strcmp(name, "psi")) bc_neumann(o, o->psi);
  else if (!strcmp(name, "t")) bc_t(o);
  else if (!strcmp(name, "vof")) bc_vof(o);
  else return 1;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* domain: src/domain.c:103-143,170-209 */
orc_t * orc_create(const size_t * glsizes, const double * lengths,
                   const double * xf_in, const double * xc_in) {
  orc_t * o = calloc(1, sizeof(orc_t));
  if (!o) return NULL;
  o->isize = (int)glsizes[0];
  o->jsize = (int)glsizes[1];
  o->ksize = NDIMS == 3 ? (int)glsizes[2] : 1;
  for (int d = 0; d < NDIMS; d++) o->lengths[d] = lengths[d];
  const int isize = o->isize, jsize = o->jsize;
  o->xf = calloc(isize + 1, sizeof(double));
  o->xc = calloc(isize + 2, sizeof(double));
  memcpy(o->xf, xf_in, sizeof(double) * (isize + 1));
  memcpy(o->xc, xc_in, sizeof(double) * (isize + 2));
  o->dxf = calloc(isize, sizeof(double));
  o->dxc = calloc(isize + 1, sizeof(double));
  {
    double * dxf = o->dxf; const double * xf = o->xf;
    for (int i = 1; i <= isize; i++) DXF(i) = XF(i + 1) - XF(i);
    double * dxc = o->dxc; const double * xc = o->xc;
    for (int i = 1; i <= isize + 1; i++) DXC(i) = XC(i) - XC(i - 1);
  }
  o->dy = o->lengths[1] / jsize;
  o->dz = NDIMS == 3 ? o->lengths[2] / o->ksize : 0.;
  /* domain.c:75-98 */
  {
    double mn = +1. * DBL_MAX, mx = -1. * DBL_MAX;
    for (int i = 0; i < isize; i++) { mn = fmin(mn, o->dxf[i]); mx = fmax(mx, o->dxf[i]); }
    o->x_uniform = fabs(mx - mn) < 1.e-15;
  }
  /* defaults = reference compile-time parameters and exec.sh values */
  orc_params_t d = {1.e4, 1., 10., 0.4, 0.95, 1, 0., 0., 0., 0., +0.5, -0.5};
  orc_set_params(o, &d);
  /* allocate every array (fluid/init.c:24-43, interface/init.c:26-43) */
  static const char * const names[] = {
    "ux", "uy", "uz", "p", "psi", "t", "vof", "curv", "dvof", "normal", "flxx",
    "flxy", "flxz", "ifrcx", "ifrcy", "ifrcz",
  };
  double ** slots[] = {
    &o->ux, &o->uy, &o->uz, &o->p, &o->psi, &o->t, &o->vof, &o->curv,
    (double **)&o->dvof, (double **)&o->normal, &o->flxx, &o->flxy, &o->flxz,
    &o->ifrcx, &o->ifrcy, &o->ifrcz,
  };
  for (size_t n = 0; n < sizeof(names) / sizeof(names[0]); n++) {
    const size_t cnt = orc_field_count(o, names[n]);
    *slots[n] = cnt ? calloc(cnt, sizeof(double)) : NULL;
  }
  for (int n = 0; n < 3; n++) {
    o->srcux[n] = calloc(orc_field_count(o, "srcux_a"), sizeof(double));
    o->srcuy[n] = calloc(orc_field_count(o, "srcuy_a"), sizeof(double));
    if (NDIMS == 3) o->srcuz[n] = calloc(orc_field_count(o, "srcuy_a"), sizeof(double));
    o->srct[n] = calloc(orc_field_count(o, "srct_a"), sizeof(double));
  }
  for (int n = 0; n < 2; n++) o->src[n] = calloc(orc_field_count(o, "src_a"), sizeof(double));
  /* predict/ux.c:35-62: Laplacian of ux in x, i = 2..isize */
  o->lapx_ux = calloc(3 * (size_t)(isize - 1), sizeof(double));
  {
    const double * dxf = o->dxf, * dxc = o->dxc;
    for (int i = 2; i <= isize; i++) {
      const double l = 1. / DXF(i - 1) / DXC(i);
      const double u = 1. / DXF(i) / DXC(i);
      const double c = -l - u;
      o->lapx_ux[3 * (i - 2) + 0] = l;
      o->lapx_ux[3 * (i - 2) + 1] = c;
      o->lapx_ux[3 * (i - 2) + 2] = u;
    }
  }
  /* predict/uy.c:35-62 and t.c:33-60: cell-centred Laplacian in x */
  o->lapx_c = calloc(3 * (size_t)isize, sizeof(double));
  {
    const double * dxf = o->dxf, * dxc = o->dxc;
    for (int i = 1; i <= isize; i++) {
      const double l = 1. / DXC(i) / DXF(i);
      const double u = 1. / DXC(i + 1) / DXF(i);
      const double c = -l - u;
      o->lapx_c[3 * (i - 1) + 0] = l;
      o->lapx_c[3 * (i - 1) + 1] = c;
      o->lapx_c[3 * (i - 1) + 2] = u;
    }
  }
  const size_t ncells = (size_t)isize * jsize * o->ksize;
  o->x1pncl = calloc(ncells, sizeof(double));
  /* complex spectra need (isize) * (jsize/2+1) * ksize * 2 doubles */
  const size_t nspec = (size_t)isize * (jsize / 2 + 1) * o->ksize * 2;
  const size_t nbuf = nspec > ncells ? nspec : ncells;
  o->pbuf0 = calloc(nbuf, sizeof(double));
  o->pbuf1 = calloc(nbuf, sizeof(double));
  return o;
}

void orc_destroy(orc_t * o) {
  if (!o) return;
  free(o->xf); free(o->xc); free(o->dxf); free(o->dxc);
  free(o->ux); free(o->uy); free(o->uz); free(o->p); free(o->psi); free(o->t);
  free(o->vof); free(o->curv); free(o->dvof); free(o->normal);
  free(o->flxx); free(o->flxy); free(o->flxz);
  free(o->ifrcx); free(o->ifrcy); free(o->ifrcz);
  for (int n = 0; n < 3; n++) { free(o->srcux[n]); free(o->srcuy[n]); free(o->srcuz[n]); free(o->srct[n]); }
  free(o->src[0]); free(o->src[1]);
  free(o->lapx_ux); free(o->lapx_c); free(o->x1pncl); free(o->pbuf0); free(o->pbuf1);
  free(o);
}

/* fluid/init.c:90-93, interface/init.c:66-68 */
void orc_set_params(orc_t * o, const orc_params_t * prm) {
  o->prm = *prm;
  o->m_dif = 1. * sqrt(prm->Pr) / sqrt(prm->Ra);
  o->t_dif = 1. / sqrt(prm->Pr) / sqrt(prm->Ra);
  o->tension = 1. / prm->We;
}
void orc_get_params(const orc_t * o, orc_params_t * prm) { *prm = o->prm; }
double orc_m_dif(const orc_t * o) { return o->m_dif; }
double orc_t_dif(const orc_t * o) { return o->t_dif; }
double orc_tension(const orc_t * o) { return o->tension; }
int orc_x_grid_is_uniform(const orc_t * o) { return o->x_uniform; }

/* array.c:80-144: NPY slab (interior y/z, padded x) -> padded array, then
 * the boundary update that fluid_init / interface_init apply */
static int slab_copy(orc_t * o, const char * name, double * slab, int to_array) {
  int ok = 0;
  shape_t s = shape_of(o, name, &ok);
  double * a = field_ptr(o, name);
  if (!ok || !a || s.ncomp != 1) return 1;
  const int nh = s.nhy;
  const int ny = o->jsize;
  const int nz = NDIMS == 3 ? o->ksize : 1;
  const int nhz = NDIMS == 3 ? nh : 0;
  size_t cnt = 0;
  for (int k = 0; k < nz; k++) {
    for (int j = 0; j < ny; j++) {
      double * row = a + (size_t)s.n0 * ((size_t)(j + nh) + (size_t)s.n1 * (size_t)(k + nhz));
      if (to_array) memcpy(row, slab + cnt, sizeof(double) * s.n0);
      else memcpy(slab + cnt, row, sizeof(double) * s.n0);
      cnt += s.n0;
    }
  }
  return 0;
}

int orc_load_ic(orc_t * o, const char * name, const double * slab) {
  if (slab_copy(o, name, (double *)slab, 1)) return 1;
  return orc_update_boundaries(o, name);
}

int orc_dump_slab(const orc_t * o, const char * name, double * slab) {
  return slab_copy((orc_t *)o, name, slab, 0);
}

/* ---------------------------------------------------------------------- */
/* tri-diagonal solvers: src/tdm.c:17-106 */
static void gtsv_real(int n, const double * l, const double * c, const double * u,
                      double * v, double * q) {
  v[0] = u[0] / c[0];
  q[0] = q[0] / c[0];
  for (int i = 1; i < n - 1; i++) {
    const double val = 1. / (c[i] - l[i] * v[i - 1]);
    v[i] = val * (u[i]);
    q[i] = val * (q[i] - l[i] * q[i - 1]);
  }
  const double val = c[n - 1] - l[n - 1] * v[n - 2];
  if (fabs(val) > DBL_EPSILON) {
    q[n - 1] = 1. / val * (q[n - 1] - l[n - 1] * q[n - 2]);
  } else {
    q[n - 1] = 0.;
  }
  for (int i = n - 2; i >= 0; i--) q[i] -= v[i] * q[i + 1];
}

/* same recurrence on an interleaved (re, im) right-hand side; real-by-complex
 * products act component-wise (C99 Annex G, as the fftw_complex instance) */
static void gtsv_cplx(int n, const double * l, const double * c, const double * u,
                      double * v, double * q) {
  v[0] = u[0] / c[0];
  q[0] = q[0] / c[0];
  q[1] = q[1] / c[0];
  for (int i = 1; i < n - 1; i++) {
    const double val = 1. / (c[i] - l[i] * v[i - 1]);
    v[i] = val * (u[i]);
    q[2 * i + 0] = val * (q[2 * i + 0] - l[i] * q[2 * i - 2]);
    q[2 * i + 1] = val * (q[2 * i + 1] - l[i] * q[2 * i - 1]);
  }
  const double val = c[n - 1] - l[n - 1] * v[n - 2];
  if (fabs(val) > DBL_EPSILON) {
    const double r = 1. / val;
    q[2 * (n - 1) + 0] = r * (q[2 * (n - 1) + 0] - l[n - 1] * q[2 * (n - 2) + 0]);
    q[2 * (n - 1) + 1] = r * (q[2 * (n - 1) + 1] - l[n - 1] * q[2 * (n - 2) + 1]);
  } else {
    q[2 * (n - 1) + 0] = 0.;
    q[2 * (n - 1) + 1] = 0.;
  }
  for (int i = n - 2; i >= 0; i--) {
    q[2 * i + 0] -= v[i] * q[2 * (i + 1) + 0];
    q[2 * i + 1] -= v[i] * q[2 * (i + 1) + 1];
  }
}

/* tdm.c:66-106, real right-hand sides */
static void tdm_solve_real(int n, int nrhs, int periodic, const double * l,
                           const double * c, const double * u, double * v,
                           double * q, double * q1) {
  if (periodic) {
    for (int i = 0; i < n - 1; i++) {
      q1[i] = i == 0 ? -l[i] : i == n - 2 ? -u[i] : 0.;
    }
    gtsv_real(n - 1, l, c, u, v, q1);
    for (int j = 0; j < nrhs; j++) {
      double * q0 = q + (size_t)j * n;
      gtsv_real(n - 1, l, c, u, v, q0);
      const double num = q0[n - 1] - u[n - 1] * q0[0] - l[n - 1] * q0[n - 2];
      const double den = c[n - 1] + u[n - 1] * q1[0] + l[n - 1] * q1[n - 2];
      q0[n - 1] = fabs(den) < DBL_EPSILON ? 0. : num / den;
      for (int i = 0; i < n - 1; i++) q0[i] = q0[i] + q0[n - 1] * q1[i];
    }
  } else {
    for (int j = 0; j < nrhs; j++) gtsv_real(n, l, c, u, v, q + (size_t)j * n);
  }
}

static void tdm_solve_cplx(int n, int nrhs, int periodic, const double * l,
                           const double * c, const double * u, double * v,
                           double * q, double * q1) {
  if (periodic) {
    for (int i = 0; i < n - 1; i++) {
      q1[i] = i == 0 ? -l[i] : i == n - 2 ? -u[i] : 0.;
    }
    gtsv_real(n - 1, l, c, u, v, q1);
    for (int j = 0; j < nrhs; j++) {
      double * q0 = q + (size_t)2 * j * n;
      gtsv_cplx(n - 1, l, c, u, v, q0);
      const double den = c[n - 1] + u[n - 1] * q1[0] + l[n - 1] * q1[n - 2];
      for (int ri = 0; ri < 2; ri++) {
        const double num = q0[2 * (n - 1) + ri] - u[n - 1] * q0[ri] - l[n - 1] * q0[2 * (n - 2) + ri];
        q0[2 * (n - 1) + ri] = fabs(den) < DBL_EPSILON ? 0. : num / den;
      }
      for (int i = 0; i < n - 1; i++) {
        q0[2 * i + 0] = q0[2 * i + 0] + q0[2 * (n - 1) + 0] * q1[i];
        q0[2 * i + 1] = q0[2 * i + 1] + q0[2 * (n - 1) + 1] * q1[i];
      }
    }
  } else {
    for (int j = 0; j < nrhs; j++) gtsv_cplx(n, l, c, u, v, q + (size_t)2 * j * n);
  }
}

int orc_tdm_solve(int n, int nrhs, int periodic, const double * l,
                  const double * c, const double * u, double * q) {
  if (n <= 1 || nrhs <= 0) return 1;
  double * v = calloc(n, sizeof(double));
  double * q1 = calloc(n, sizeof(double));
  tdm_solve_real(n, nrhs, periodic, l, c, u, v, q, q1);
  free(v); free(q1);
  return 0;
}

int orc_tdm_solve_complex(int n, int nrhs, int periodic, const double * l,
                          const double * c, const double * u, double * q) {
  if (n <= 1 || nrhs <= 0) return 1;
  double * v = calloc(n, sizeof(double));
  double * q1 = calloc(n, sizeof(double));
  tdm_solve_cplx(n, nrhs, periodic, l, c, u, v, q, q1);
  free(v); free(q1);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* transforms with FFTW's unnormalised conventions (the reference calls
 * fftw_plan_many_dft_r2c/c2r and REDFT10/REDFT01: dct.c:209-220,
 * dft.c:239-250).  Mixed-radix recursive DFT with an exactly symmetric
 * twiddle table. */
typedef struct { int n; double * c, * s; } twiddle_t;

/* Twiddle table: 0 = octant-reduced, exactly symmetric (default); 1 =
 * cos / sin of 2 pi k / n directly.  Both give the same DFT up to rounding
 * (as two FFTW plans do); the tests use 1 to measure how a step amplifies
 * rounding. */
static int g_fft_variant = 0;
void orc_set_fft_variant(int v) { g_fft_variant = v ? 1 : 0; }

static void twiddle_make(twiddle_t * tw, int n) {
  tw->n = n;
  tw->c = malloc(sizeof(double) * n);
  tw->s = malloc(sizeof(double) * n);
  for (int k = 0; k < n; k++) {
    if (g_fft_variant) {
      tw->c[k] = cos(2. * g_pi * (double)k / n);
      tw->s[k] = sin(2. * g_pi * (double)k / n);
      continue;
    }
    /* reduce to the first octant so that symmetric angles give
     * bit-identical magnitudes */
    const long m8 = 8L * k;  /* angle = 2 pi k / n = (pi/4) (m8 / n) */
    double c, s;
    if (m8 % n == 0) {
      static const double cs[8][2] = {
        {1., 0.}, {0.70710678118654752440, 0.70710678118654752440}, {0., 1.},
        {-0.70710678118654752440, 0.70710678118654752440}, {-1., 0.},
        {-0.70710678118654752440, -0.70710678118654752440}, {0., -1.},
        {0.70710678118654752440, -0.70710678118654752440},
      };
      const int oct = (int)(m8 / n);
      c = cs[oct][0]; s = cs[oct][1];
    } else {
      /* k in quadrant q: angle = q pi/2 + r, 0 < r < pi/2 */
      const long kk = 4L * k;
      const int q = (int)(kk / n);
      const long rem = kk - (long)q * n;    /* r = (pi/2) rem / n */
      double cr, sr;
      if (2 * rem <= n) {
        const double r = 0.5 * g_pi * (double)rem / n;
        cr = cos(r); sr = sin(r);
      } else {
        const double r = 0.5 * g_pi * (double)(n - rem) / n;
        cr = sin(r); sr = cos(r);
      }
      switch (q) {
        case 0: c = cr; s = sr; break;
        case 1: c = -sr; s = cr; break;
        case 2: c = -cr; s = -sr; break;
        default: c = sr; s = -cr; break;
      }
    }
    tw->c[k] = c;
    tw->s[k] = s;
  }
}

static void twiddle_free(twiddle_t * tw) { free(tw->c); free(tw->s); }

/* out[k] = sum_j in[j*stride] * exp(sign * 2 pi i j k / n), n divides tw->n */
static void dft_rec(const twiddle_t * tw, int n, int stride, int sign,
                    const double * in, double * out, double * tmp) {
  if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
  int r = 2;
  while (r * r <= n && n % r) r++;
  if (n % r) r = n;
  const int m = n / r;
  const int scale = tw->n / n;
  /* r sub-transforms of length m on the decimated sequences */
  for (int q = 0; q < r; q++) {
    dft_rec(tw, m, stride * r, sign, in + 2 * (size_t)q * stride, tmp + 2 * (size_t)q * m, out);
  }
  for (int kk = 0; kk < n; kk++) {
    const int k = kk % m;
    double re = 0., im = 0.;
    for (int q = 0; q < r; q++) {
      const long idx = ((long)q * kk % n) * scale;
      const double wc = tw->c[idx];
      const double ws = sign * tw->s[idx];
      const double xr = tmp[2 * ((size_t)q * m + k) + 0];
      const double xi = tmp[2 * ((size_t)q * m + k) + 1];
      re += xr * wc - xi * ws;
      im += xr * ws + xi * wc;
    }
    out[2 * kk + 0] = re;
    out[2 * kk + 1] = im;
  }
  /* the sub-results in tmp were consumed; the caller's tmp is free again */
}

static void dft(const twiddle_t * tw, int n, int sign, double * reim, double * w1, double * w2) {
  /* w1, w2: scratch of 2n doubles each */
  memcpy(w1, reim, sizeof(double) * 2 * n);
  dft_rec(tw, n, 1, sign, w1, reim, w2);
}

int orc_fft(int n, int sign, double * reim) {
  twiddle_t tw; twiddle_make(&tw, n);
  double * w1 = malloc(sizeof(double) * 2 * n), * w2 = malloc(sizeof(double) * 2 * n);
  dft(&tw, n, sign < 0 ? -1 : +1, reim, w1, w2);
  free(w1); free(w2); twiddle_free(&tw);
  return 0;
}

/* r2c: out[k] = sum_j in[j] e^{-2 pi i jk/n}, k = 0..n/2 (FFTW forward) */
static void rfft_line(const twiddle_t * tw, int n, const double * in, double * out,
                      double * w0, double * w1, double * w2) {
  for (int j = 0; j < n; j++) { w0[2 * j] = in[j]; w0[2 * j + 1] = 0.; }
  dft(tw, n, -1, w0, w1, w2);
  memcpy(out, w0, sizeof(double) * 2 * (n / 2 + 1));
}

/* c2r: out[j] = sum_k Z[k] e^{+2 pi i jk/n} over the Hermitian extension; the
 * imaginary parts of the k = 0 and k = n/2 bins do not contribute */
static void irfft_line(const twiddle_t * tw, int n, const double * in, double * out,
                       double * w0, double * w1, double * w2) {
  const int h = n / 2 + 1;
  for (int k = 0; k < h; k++) { w0[2 * k] = in[2 * k]; w0[2 * k + 1] = in[2 * k + 1]; }
  w0[1] = 0.;
  if (n % 2 == 0) w0[2 * (n / 2) + 1] = 0.;
  for (int k = h; k < n; k++) { w0[2 * k] = in[2 * (n - k)]; w0[2 * k + 1] = -in[2 * (n - k) + 1]; }
  dft(tw, n, +1, w0, w1, w2);
  for (int j = 0; j < n; j++) out[j] = w0[2 * j];
}

int orc_rfft(int n, const double * in, double * out) {
  twiddle_t tw; twiddle_make(&tw, n);
  double * w = malloc(sizeof(double) * 6 * n);
  rfft_line(&tw, n, in, out, w, w + 2 * n, w + 4 * n);
  free(w); twiddle_free(&tw);
  return 0;
}

int orc_irfft(int n, const double * in, double * out) {
  twiddle_t tw; twiddle_make(&tw, n);
  double * w = malloc(sizeof(double) * 6 * n);
  irfft_line(&tw, n, in, out, w, w + 2 * n, w + 4 * n);
  free(w); twiddle_free(&tw);
  return 0;
}

/* REDFT10 (DCT-II): Y_k = 2 sum_j x_j cos(pi (2j+1) k / (2n)); direct sum
 * over a cosine table of period 4n */
static double * dct_table(int n) {
  double * ct = malloc(sizeof(double) * 4 * n);
  twiddle_t tw; twiddle_make(&tw, 4 * n);
  memcpy(ct, tw.c, sizeof(double) * 4 * n);
  twiddle_free(&tw);
  return ct;  /* ct[m] = cos(2 pi m / (4n)) = cos(pi m / (2n)) */
}

static void redft10_line(int n, const double * ct, const double * in, double * out) {
  for (int k = 0; k < n; k++) {
    double s = 0.;
    for (int j = 0; j < n; j++) s += in[j] * ct[((long)(2 * j + 1) * k) % (4 * n)];
    out[k] = 2. * s;
  }
}

/* REDFT01 (DCT-III): Y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (2k+1) / (2n)) */
static void redft01_line(int n, const double * ct, const double * in, double * out) {
  for (int k = 0; k < n; k++) {
    double s = 0.;
    for (int j = 1; j < n; j++) s += in[j] * ct[((long)j * (2 * k + 1)) % (4 * n)];
    out[k] = in[0] + 2. * s;
  }
}

int orc_redft10(int n, const double * in, double * out) {
  double * ct = dct_table(n); redft10_line(n, ct, in, out); free(ct); return 0;
}
int orc_redft01(int n, const double * in, double * out) {
  double * ct = dct_table(n); redft01_line(n, ct, in, out); free(ct); return 0;
}

/* ---------------------------------------------------------------------- */
/* fluid_decide_dt: src/fluid/decide_dt.c:28-145 */
int orc_decide_dt(orc_t * o, double * dt) {
  SIZES;
  const double * dxc = o->dxc;
  const double dy = o->dy;
  const double * ux = o->ux, * uy = o->uy;
  const double small = 1.e-8;
  double dt_adv = 1.;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++) {
    const double dx = DXC(i);
    const double vel = fabs(UX(i, j, k)) + small;
    dt_adv = fmin(dt_adv, dx / vel);
  }
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    const double vel = fabs(UY(i, j, k)) + small;
    dt_adv = fmin(dt_adv, dy / vel);
  }
#if NDIMS == 3
  {
    const double dz = o->dz; const double * uz = o->uz;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double vel = fabs(UZ(i, j, k)) + small;
      dt_adv = fmin(dt_adv, dz / vel);
    }
  }
#endif
  dt_adv *= o->prm.coef_dt_adv;
  /* decide_dt_dif: only the explicit directions (y, z) enter (param/implicit.c) */
  double hs[3];
  hs[0] = DBL_MAX;
  for (int i = 2; i <= isize; i++) hs[0] = fmin(hs[0], DXC(i));
  hs[1] = dy;
  hs[2] = o->dz;
  double res = dt_adv;
  const double difs[2] = {o->m_dif, o->t_dif};
  for (int f = 0; f < 2; f++) {
    for (int dim = 1; dim < NDIMS; dim++) {
      const double dt_dif = o->prm.coef_dt_dif / difs[f] * 0.5 / NDIMS * (hs[dim] * hs[dim]);
      res = fmin(res, dt_dif);
    }
  }
  *dt = res;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* interface_compute_curvature_tensor: src/interface/curvature_tensor.c */

/* curvature_tensor.c:14-48; 3D: eight-cell differences, z component added */
static void compute_gradient(orc_t * o) {
  SIZES;
  const double * dxc = o->dxc;
  const double dy = o->dy;
  const double * vof = o->vof;
  vector_t * dvof = o->dvof;
  KLOOP(0, ksize + 2) for (int j = 0; j <= jsize + 2; j++) for (int i = 1; i <= isize + 1; i++) {
    const double dx = DXC(i);
#if NDIMS == 2
    const double dvofdx = 1. / dx * (
        - VOF(i - 1, j - 1, k) + VOF(i, j - 1, k)
        - VOF(i - 1, j,     k) + VOF(i, j,     k));
    const double dvofdy = 1. / dy * (
        - VOF(i - 1, j - 1, k) - VOF(i, j - 1, k)
        + VOF(i - 1, j,     k) + VOF(i, j,     k));
    const double norm = sqrt(dvofdx * dvofdx + dvofdy * dvofdy);
    const double norminv = 1. / fmax(norm, DBL_EPSILON);
    DVOF(i, j, k)[0] = dvofdx * norminv;
    DVOF(i, j, k)[1] = dvofdy * norminv;
#else
    /* per-plane partial sums in the 2D order, so that a z-uniform field
     * reproduces the 2D corner normals exactly */
    const double dz = o->dz;
    const double sx0 = -VOF(i - 1, j - 1, k - 1) + VOF(i, j - 1, k - 1) - VOF(i - 1, j, k - 1) + VOF(i, j, k - 1);
    const double sx1 = -VOF(i - 1, j - 1, k) + VOF(i, j - 1, k) - VOF(i - 1, j, k) + VOF(i, j, k);
    const double sy0 = -VOF(i - 1, j - 1, k - 1) - VOF(i, j - 1, k - 1) + VOF(i - 1, j, k - 1) + VOF(i, j, k - 1);
    const double sy1 = -VOF(i - 1, j - 1, k) - VOF(i, j - 1, k) + VOF(i - 1, j, k) + VOF(i, j, k);
    const double sz0 = VOF(i - 1, j - 1, k - 1) + VOF(i, j - 1, k - 1) + VOF(i - 1, j, k - 1) + VOF(i, j, k - 1);
    const double sz1 = VOF(i - 1, j - 1, k) + VOF(i, j - 1, k) + VOF(i - 1, j, k) + VOF(i, j, k);
    const double dvofdx = 1. / dx * (sx0 + sx1);
    const double dvofdy = 1. / dy * (sy0 + sy1);
    const double dvofdz = 1. / dz * (-sz0 + sz1);
    const double norm = sqrt(dvofdx * dvofdx + dvofdy * dvofdy + dvofdz * dvofdz);
    const double norminv = 1. / fmax(norm, DBL_EPSILON);
    DVOF(i, j, k)[0] = dvofdx * norminv;
    DVOF(i, j, k)[1] = dvofdy * norminv;
    DVOF(i, j, k)[2] = dvofdz * norminv;
#endif
  }
}

/* curvature_tensor.c:50-92; 3D: 2x2x2 Gauss points */
static double compute_intercept(const double vof, const double * normal) {
  const int cntmax = 8;
  const double resmax = 1.e-12;
#if NDIMS == 2
  double exps[NGAUSS * NGAUSS];
  for (int jj = 0; jj < NGAUSS; jj++) for (int ii = 0; ii < NGAUSS; ii++) {
    exps[jj * NGAUSS + ii] = exp(-2. * vofbeta * (normal[0] * gauss_ps[ii] + normal[1] * gauss_ps[jj]));
  }
#else
  double exps[NGAUSS * NGAUSS * NGAUSS];
  for (int kk = 0; kk < NGAUSS; kk++) for (int jj = 0; jj < NGAUSS; jj++) for (int ii = 0; ii < NGAUSS; ii++) {
    exps[(kk * NGAUSS + jj) * NGAUSS + ii] = exp(-2. * vofbeta * (
          normal[0] * gauss_ps[ii] + normal[1] * gauss_ps[jj] + normal[2] * gauss_ps[kk]));
  }
#endif
  double val = 1. / vof - 1.;
  for (int cnt = 0; cnt < cntmax; cnt++) {
    double f0 = -1. * vof;
    double f1 = 0.;
#if NDIMS == 2
    for (int jj = 0; jj < NGAUSS; jj++) for (int ii = 0; ii < NGAUSS; ii++) {
      const double weight = gauss_ws[ii] * gauss_ws[jj];
      const double p = exps[jj * NGAUSS + ii];
      const double denom = 1. / (1. + p * val);
      f0 += weight * denom;
      f1 -= weight * p * denom * denom;
    }
#else
    /* z weights innermost: with a zero z normal the two z points coincide
     * and the sums reduce exactly to the 2D ones */
    for (int jj = 0; jj < NGAUSS; jj++) for (int ii = 0; ii < NGAUSS; ii++) {
      const double weight = gauss_ws[ii] * gauss_ws[jj];
      double s0 = 0., s1 = 0.;
      for (int kk = 0; kk < NGAUSS; kk++) {
        const double p = exps[(kk * NGAUSS + jj) * NGAUSS + ii];
        const double denom = 1. / (1. + p * val);
        s0 += gauss_ws[kk] * denom;
        s1 += gauss_ws[kk] * (weight * p * denom * denom);
      }
      f0 += weight * s0;
      f1 -= s1;
    }
#endif
    val -= f0 / f1;
    if (fabs(f0) < resmax) break;
  }
  return -0.5 / vofbeta * log(val);
}

double orc_compute_intercept(double vof, const double * normal) {
  return compute_intercept(vof, normal);
}

/* curvature_tensor.c:94-145 */
static void compute_normal(orc_t * o) {
  SIZES;
  const double * dxf = o->dxf;
  const double dy = o->dy;
  const double * vof = o->vof;
  const vector_t * dvof = o->dvof;
  normal_t * normal = o->normal;
  KLOOP(0, ksize + 1) for (int j = 0; j <= jsize + 1; j++) for (int i = 1; i <= isize; i++) {
    const double dx = DXF(i);
    const double lvof = VOF(i, j, k);
    if (lvof < vofmin || 1. - vofmin < lvof) continue;
#if NDIMS == 2
    double nx = DVOF(i, j, k)[0] + DVOF(i + 1, j, k)[0] + DVOF(i, j + 1, k)[0] + DVOF(i + 1, j + 1, k)[0];
    double ny = DVOF(i, j, k)[1] + DVOF(i + 1, j, k)[1] + DVOF(i, j + 1, k)[1] + DVOF(i + 1, j + 1, k)[1];
    nx /= dx;
    ny /= dy;
    const double norm = sqrt(nx * nx + ny * ny);
    const double norminv = 1. / fmax(norm, DBL_EPSILON);
    nx *= norminv;
    ny *= norminv;
    const double n2[2] = {nx, ny};
    const double seg = compute_intercept(lvof, n2);
    NORMAL(i, j, k)[0] = nx;
    NORMAL(i, j, k)[1] = ny;
    NORMAL(i, j, k)[2] = seg;
#else
    double n3[3];
    for (int d = 0; d < 3; d++) {
      n3[d] = (DVOF(i, j, k)[d] + DVOF(i + 1, j, k)[d] + DVOF(i, j + 1, k)[d] + DVOF(i + 1, j + 1, k)[d])
            + (DVOF(i, j, k + 1)[d] + DVOF(i + 1, j, k + 1)[d] + DVOF(i, j + 1, k + 1)[d] + DVOF(i + 1, j + 1, k + 1)[d]);
    }
    n3[0] /= dx;
    n3[1] /= dy;
    n3[2] /= o->dz;
    const double norm = sqrt(n3[0] * n3[0] + n3[1] * n3[1] + n3[2] * n3[2]);
    const double norminv = 1. / fmax(norm, DBL_EPSILON);
    n3[0] *= norminv;
    n3[1] *= norminv;
    n3[2] *= norminv;
    const double seg = compute_intercept(lvof, n3);
    NORMAL(i, j, k)[0] = n3[0];
    NORMAL(i, j, k)[1] = n3[1];
    NORMAL(i, j, k)[2] = n3[2];
    NORMAL(i, j, k)[3] = seg;
#endif
  }
}

/* curvature_tensor.c:147-176; 3D: four differences per direction, 1/4 */
static void compute_curvature(orc_t * o) {
  SIZES;
  const double * dxf = o->dxf;
  const double dy = o->dy;
  const vector_t * dvof = o->dvof;
  double * curv = o->curv;
  KLOOP(0, ksize + 1) for (int j = 0; j <= jsize + 1; j++) for (int i = 1; i <= isize; i++) {
    const double dx = DXF(i);
#if NDIMS == 2
    const double dnxdx = 1. / dx * (
        - DVOF(i, j, k)[0] + DVOF(i + 1, j, k)[0]
        - DVOF(i, j + 1, k)[0] + DVOF(i + 1, j + 1, k)[0]);
    const double dnydy = 1. / dy * (
        - DVOF(i, j, k)[1] - DVOF(i + 1, j, k)[1]
        + DVOF(i, j + 1, k)[1] + DVOF(i + 1, j + 1, k)[1]);
    CURV(i, j, k) = 0.5 * (-dnxdx - dnydy);
#else
    const double dz = o->dz;
    const double dnxdx = 1. / dx * (
        (- DVOF(i, j, k)[0] + DVOF(i + 1, j, k)[0] - DVOF(i, j + 1, k)[0] + DVOF(i + 1, j + 1, k)[0])
      + (- DVOF(i, j, k + 1)[0] + DVOF(i + 1, j, k + 1)[0] - DVOF(i, j + 1, k + 1)[0] + DVOF(i + 1, j + 1, k + 1)[0]));
    const double dnydy = 1. / dy * (
        (- DVOF(i, j, k)[1] - DVOF(i + 1, j, k)[1] + DVOF(i, j + 1, k)[1] + DVOF(i + 1, j + 1, k)[1])
      + (- DVOF(i, j, k + 1)[1] - DVOF(i + 1, j, k + 1)[1] + DVOF(i, j + 1, k + 1)[1] + DVOF(i + 1, j + 1, k + 1)[1]));
    const double dnzdz = 1. / dz * (
        - (DVOF(i, j, k)[2] + DVOF(i + 1, j, k)[2] + DVOF(i, j + 1, k)[2] + DVOF(i + 1, j + 1, k)[2])
        + (DVOF(i, j, k + 1)[2] + DVOF(i + 1, j, k + 1)[2] + DVOF(i, j + 1, k + 1)[2] + DVOF(i + 1, j + 1, k + 1)[2]));
    CURV(i, j, k) = 0.25 * (-dnxdx - dnydy - dnzdz);
#endif
  }
}

int orc_interface_compute_curvature_tensor(orc_t * o) {
  compute_gradient(o);
  compute_normal(o);
  compute_curvature(o);
  return 0;
}

/* interface_compute_force: src/interface/force.c:11-76 */
int orc_interface_compute_force(orc_t * o) {
  SIZES;
  const double * dxc = o->dxc;
  const double dy = o->dy;
  const double tension = o->tension;
  const double * vof = o->vof, * curv = o->curv;
  double * ifrcx = o->ifrcx, * ifrcy = o->ifrcy;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++) {
    const double dx = DXC(i);
    const double grad = 1. / dx * (-VOF(i - 1, j, k) + VOF(i, j, k));
    const double kappa = 0.5 * (CURV(i - 1, j, k) + CURV(i, j, k));
    IFRCX(i, j, k) = tension * grad * kappa;
  }
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    const double grad = 1. / dy * (-VOF(i, j - 1, k) + VOF(i, j, k));
    const double kappa = 0.5 * (CURV(i, j - 1, k) + CURV(i, j, k));
    IFRCY(i, j, k) = tension * grad * kappa;
  }
#if NDIMS == 3
  {
    const double dz = o->dz; double * ifrcz = o->ifrcz;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double grad = 1. / dz * (-VOF(i, j, k - 1) + VOF(i, j, k));
      const double kappa = 0.5 * (CURV(i, j, k - 1) + CURV(i, j, k));
      IFRCZ(i, j, k) = tension * grad * kappa;
    }
  }
#endif
  return 0;
}

/* ---------------------------------------------------------------------- */
/* interface_update_vof: src/interface/update/{main,flxx,flxy}.c */

/* update/main.c:16-26 */
static double indicator(const double * n, const double * x) {
#if NDIMS == 2
  return 1. / (1. + exp(-2. * vofbeta * (n[0] * x[0] + n[1] * x[1] + n[NDIMS])));
#else
  return 1. / (1. + exp(-2. * vofbeta * (n[0] * x[0] + n[1] * x[1] + n[2] * x[2] + n[NDIMS])));
#endif
}

double orc_indicator(const double * n, const double * x) { return indicator(n, x); }

/* update/flxx.c:11-44; 3D: 2x2 Gauss points on the face */
static void compute_flux_x(orc_t * o) {
  SIZES;
  const double * ux = o->ux, * vof = o->vof;
  const normal_t * normal = o->normal;
  double * flxx = o->flxx;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++) {
    const double vel = UX(i, j, k);
    const int ii = vel < 0. ? i : i - 1;
    const double x = vel < 0. ? -0.5 : +0.5;
    const double lvof = VOF(ii, j, k);
    if (lvof < vofmin || 1. - vofmin < lvof) { FLXX(i, j, k) = vel * lvof; continue; }
    double flux = 0.;
#if NDIMS == 2
    for (int jj = 0; jj < NGAUSS; jj++) {
      const double w = gauss_ws[jj];
      const double pt[2] = {x, gauss_ps[jj]};
      flux += w * indicator(NORMAL(ii, j, k), pt);
    }
#else
    for (int jj = 0; jj < NGAUSS; jj++) {
      double sz = 0.;
      for (int kk = 0; kk < NGAUSS; kk++) {
        const double pt[3] = {x, gauss_ps[jj], gauss_ps[kk]};
        sz += gauss_ws[kk] * indicator(NORMAL(ii, j, k), pt);
      }
      flux += gauss_ws[jj] * sz;
    }
#endif
    FLXX(i, j, k) = vel * flux;
  }
}

/* update/flxy.c:11-44 */
static void compute_flux_y(orc_t * o) {
  SIZES;
  const double * uy = o->uy, * vof = o->vof;
  const normal_t * normal = o->normal;
  double * flxy = o->flxy;
  KLOOP(1, ksize) for (int j = 1; j <= jsize + 1; j++) for (int i = 1; i <= isize; i++) {
    const double vel = UY(i, j, k);
    const int jj = vel < 0. ? j : j - 1;
    const double y = vel < 0. ? -0.5 : +0.5;
    const double lvof = VOF(i, jj, k);
    if (lvof < vofmin || 1. - vofmin < lvof) { FLXY(i, j, k) = vel * lvof; continue; }
    double flux = 0.;
#if NDIMS == 2
    for (int ii = 0; ii < NGAUSS; ii++) {
      const double w = gauss_ws[ii];
      const double pt[2] = {gauss_ps[ii], y};
      flux += w * indicator(NORMAL(i, jj, k), pt);
    }
#else
    for (int ii = 0; ii < NGAUSS; ii++) {
      double sz = 0.;
      for (int kk = 0; kk < NGAUSS; kk++) {
        const double pt[3] = {gauss_ps[ii], y, gauss_ps[kk]};
        sz += gauss_ws[kk] * indicator(NORMAL(i, jj, k), pt);
      }
      flux += gauss_ws[ii] * sz;
    }
#endif
    FLXY(i, j, k) = vel * flux;
  }
}

#if NDIMS == 3
/* 3D extrapolation of update/flxy.c to the z faces */
static void compute_flux_z(orc_t * o) {
  SIZES;
  const double * uz = o->uz, * vof = o->vof;
  const normal_t * normal = o->normal;
  double * flxz = o->flxz;
  KLOOP(1, ksize + 1) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    const double vel = UZ(i, j, k);
    const int kk0 = vel < 0. ? k : k - 1;
    const double z = vel < 0. ? -0.5 : +0.5;
    const double lvof = VOF(i, j, kk0);
    if (lvof < vofmin || 1. - vofmin < lvof) { FLXZ(i, j, k) = vel * lvof; continue; }
    double flux = 0.;
    for (int ii = 0; ii < NGAUSS; ii++) {
      double sy = 0.;
      for (int jj = 0; jj < NGAUSS; jj++) {
        const double pt[3] = {gauss_ps[ii], gauss_ps[jj], z};
        sy += gauss_ws[jj] * indicator(NORMAL(i, j, kk0), pt);
      }
      flux += gauss_ws[ii] * sy;
    }
    FLXZ(i, j, k) = vel * flux;
  }
}
#endif

/* update/main.c:28-45 (pointer swap carry + zero) */
static void reset_srcs2(int rkstep, double ** srca, double ** srcb, size_t n) {
  if (0 != rkstep) { double * tmp = *srca; *srca = *srcb; *srcb = tmp; }
  memset(*srca, 0, n * sizeof(double));
}

int orc_interface_update_vof(orc_t * o, int rkstep, double dt) {
  SIZES;
  const size_t nsrc = orc_field_count(o, "src_a");
  reset_srcs2(rkstep, &o->src[0], &o->src[1], nsrc);
  compute_flux_x(o);
  compute_flux_y(o);
#if NDIMS == 3
  compute_flux_z(o);
#endif
  /* update/main.c:47-78 */
  {
    const double * dxf = o->dxf;
    const double dy = o->dy;
    const double * flxx = o->flxx, * flxy = o->flxy;
    double * src = o->src[rk_a];
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double dx = DXF(i);
      SRC(i, j, k) += 1. / dx * (FLXX(i, j, k) - FLXX(i + 1, j, k));
    }
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      SRC(i, j, k) += 1. / dy * (FLXY(i, j, k) - FLXY(i, j + 1, k));
    }
#if NDIMS == 3
    const double dz = o->dz;
    const double * flxz = o->flxz;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      SRC(i, j, k) += 1. / dz * (FLXZ(i, j, k) - FLXZ(i, j, k + 1));
    }
#endif
  }
  /* update/main.c:80-110 */
  {
    double * vof = o->vof;
    {
      const double coef = rkcoefs[rkstep][rk_a];
      const double * src = o->src[rk_a];
      KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
        VOF(i, j, k) += dt * coef * SRC(i, j, k);
      }
    }
    if (0 != rkstep) {
      const double coef = rkcoefs[rkstep][rk_b];
      const double * src = o->src[rk_b];
      KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
        VOF(i, j, k) += dt * coef * SRC(i, j, k);
      }
    }
  }
  bc_vof(o);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* fluid_predict_field: src/fluid/predict/{main,ux,uy,t}.c */

static void reset_srcs3(int rkstep, double ** srca, double ** srcb, double ** srcg, size_t n) {
  if (0 != rkstep) { double * tmp = *srca; *srca = *srcb; *srcb = tmp; }
  memset(*srca, 0, n * sizeof(double));
  memset(*srcg, 0, n * sizeof(double));
}

/* compute_rhs_ux: predict/ux.c:71-249.  The loops run term by term over the
 * whole array, exactly as the reference accumulates into srca / srcg. */
static void compute_rhs_ux(orc_t * o) {
  SIZES;
  const double * dxf = o->dxf, * dxc = o->dxc;
  const double dy = o->dy;
  const double * ux = o->ux, * uy = o->uy, * p = o->p, * t = o->t;
  double * srca = o->srcux[rk_a], * srcg = o->srcux[rk_g];
  const double dif = o->m_dif;
  const double * lapx = o->lapx_ux;
  const double lapy[3] = {+1. / dy / dy, -2. / dy / dy, +1. / dy / dy};
#define BEGIN_UX \
  { size_t cnt = 0; KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++, cnt++) {
#define END_UX } }
  /* advection_x: ux.c:71-93 */
  BEGIN_UX
    const double ux_l = +0.5 * UX(i - 1, j, k) + 0.5 * UX(i, j, k);
    const double ux_u = +0.5 * UX(i, j, k) + 0.5 * UX(i + 1, j, k);
    const double l = +0.5 / DXC(i) * ux_l;
    const double u = -0.5 / DXC(i) * ux_u;
    const double c = -l - u;
    srca[cnt] += +l * UX(i - 1, j, k) + c * UX(i, j, k) + u * UX(i + 1, j, k);
  END_UX
  /* advection_y: ux.c:95-122 */
  BEGIN_UX
    const double w_xm = 0.5 * DXF(i - 1) / DXC(i);
    const double w_xp = 0.5 * DXF(i) / DXC(i);
    const double uy_l = w_xm * UY(i - 1, j, k) + w_xp * UY(i, j, k);
    const double uy_u = w_xm * UY(i - 1, j + 1, k) + w_xp * UY(i, j + 1, k);
    const double l = +0.5 / dy * uy_l;
    const double u = -0.5 / dy * uy_u;
    const double c = -l - u;
    srca[cnt] += +l * UX(i, j - 1, k) + c * UX(i, j, k) + u * UX(i, j + 1, k);
  END_UX
#if NDIMS == 3
  const double dz = o->dz;
  const double * uz = o->uz;
  const double lapz[3] = {+1. / dz / dz, -2. / dz / dz, +1. / dz / dz};
  /* advection_z: extrapolation of advection_y */
  BEGIN_UX
    const double w_xm = 0.5 * DXF(i - 1) / DXC(i);
    const double w_xp = 0.5 * DXF(i) / DXC(i);
    const double uz_l = w_xm * UZ(i - 1, j, k) + w_xp * UZ(i, j, k);
    const double uz_u = w_xm * UZ(i - 1, j, k + 1) + w_xp * UZ(i, j, k + 1);
    const double l = +0.5 / dz * uz_l;
    const double u = -0.5 / dz * uz_u;
    const double c = -l - u;
    srca[cnt] += +l * UX(i, j, k - 1) + c * UX(i, j, k) + u * UX(i, j, k + 1);
  END_UX
#endif
  /* diffusion_x (implicit -> srcg): ux.c:125-144 */
  BEGIN_UX
    srcg[cnt] += dif * (lapx[3 * (i - 2) + 0] * UX(i - 1, j, k) + lapx[3 * (i - 2) + 1] * UX(i, j, k)
                        + lapx[3 * (i - 2) + 2] * UX(i + 1, j, k));
  END_UX
  /* diffusion_y (explicit -> srca): ux.c:146-165 */
  BEGIN_UX
    srca[cnt] += dif * (lapy[0] * UX(i, j - 1, k) + lapy[1] * UX(i, j, k) + lapy[2] * UX(i, j + 1, k));
  END_UX
#if NDIMS == 3
  BEGIN_UX
    srca[cnt] += dif * (lapz[0] * UX(i, j, k - 1) + lapz[1] * UX(i, j, k) + lapz[2] * UX(i, j, k + 1));
  END_UX
#endif
  /* pressure (srcg): ux.c:168-183 */
  BEGIN_UX
    srcg[cnt] -= 1. / DXC(i) * (-P(i - 1, j, k) + P(i, j, k));
  END_UX
  /* buoyancy: ux.c:185-201 */
  if (o->prm.add_buoyancy) {
    BEGIN_UX
      srca[cnt] += +0.5 * T(i - 1, j, k) + 0.5 * T(i, j, k);
    END_UX
  }
  /* surface: ux.c:203-216 */
  BEGIN_UX
    srca[cnt] += o->ifrcx[cnt];
  END_UX
#undef BEGIN_UX
#undef END_UX
}

#define BEGIN_C \
  { size_t cnt = 0; KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++, cnt++) {
#define END_C } }

/* compute_rhs_uy: predict/uy.c:70-221 */
static void compute_rhs_uy(orc_t * o) {
  SIZES;
  const double * dxf = o->dxf;
  const double dy = o->dy;
  const double * ux = o->ux, * uy = o->uy, * p = o->p;
  double * srca = o->srcuy[rk_a], * srcg = o->srcuy[rk_g];
  const double dif = o->m_dif;
  const double * lapx = o->lapx_c;
  const double lapy[3] = {+1. / dy / dy, -2. / dy / dy, +1. / dy / dy};
  BEGIN_C
    const double ux_l = +0.5 * UX(i, j - 1, k) + 0.5 * UX(i, j, k);
    const double ux_u = +0.5 * UX(i + 1, j - 1, k) + 0.5 * UX(i + 1, j, k);
    const double l = +0.5 / DXF(i) * ux_l;
    const double u = -0.5 / DXF(i) * ux_u;
    const double c = -l - u;
    srca[cnt] += +l * UY(i - 1, j, k) + c * UY(i, j, k) + u * UY(i + 1, j, k);
  END_C
  BEGIN_C
    const double uy_l = +0.5 * UY(i, j - 1, k) + 0.5 * UY(i, j, k);
    const double uy_u = +0.5 * UY(i, j, k) + 0.5 * UY(i, j + 1, k);
    const double l = +0.5 / dy * uy_l;
    const double u = -0.5 / dy * uy_u;
    const double c = -l - u;
    srca[cnt] += +l * UY(i, j - 1, k) + c * UY(i, j, k) + u * UY(i, j + 1, k);
  END_C
#if NDIMS == 3
  const double dz = o->dz;
  const double * uz = o->uz;
  const double lapz[3] = {+1. / dz / dz, -2. / dz / dz, +1. / dz / dz};
  BEGIN_C
    const double uz_l = +0.5 * UZ(i, j - 1, k) + 0.5 * UZ(i, j, k);
    const double uz_u = +0.5 * UZ(i, j - 1, k + 1) + 0.5 * UZ(i, j, k + 1);
    const double l = +0.5 / dz * uz_l;
    const double u = -0.5 / dz * uz_u;
    const double c = -l - u;
    srca[cnt] += +l * UY(i, j, k - 1) + c * UY(i, j, k) + u * UY(i, j, k + 1);
  END_C
#endif
  BEGIN_C
    srcg[cnt] += dif * (lapx[3 * (i - 1) + 0] * UY(i - 1, j, k) + lapx[3 * (i - 1) + 1] * UY(i, j, k)
                        + lapx[3 * (i - 1) + 2] * UY(i + 1, j, k));
  END_C
  BEGIN_C
    srca[cnt] += dif * (lapy[0] * UY(i, j - 1, k) + lapy[1] * UY(i, j, k) + lapy[2] * UY(i, j + 1, k));
  END_C
#if NDIMS == 3
  BEGIN_C
    srca[cnt] += dif * (lapz[0] * UY(i, j, k - 1) + lapz[1] * UY(i, j, k) + lapz[2] * UY(i, j, k + 1));
  END_C
#endif
  BEGIN_C
    srcg[cnt] -= 1. / dy * (-P(i, j - 1, k) + P(i, j, k));
  END_C
  BEGIN_C
    srca[cnt] += o->ifrcy[cnt];
  END_C
}

#if NDIMS == 3
/* compute_rhs_uz: 3D extrapolation of predict/uy.c with the roles of y and z
 * exchanged */
static void compute_rhs_uz(orc_t * o) {
  SIZES;
  const double * dxf = o->dxf;
  const double dy = o->dy, dz = o->dz;
  const double * ux = o->ux, * uy = o->uy, * uz = o->uz, * p = o->p;
  double * srca = o->srcuz[rk_a], * srcg = o->srcuz[rk_g];
  const double dif = o->m_dif;
  const double * lapx = o->lapx_c;
  const double lapy[3] = {+1. / dy / dy, -2. / dy / dy, +1. / dy / dy};
  const double lapz[3] = {+1. / dz / dz, -2. / dz / dz, +1. / dz / dz};
  BEGIN_C
    const double ux_l = +0.5 * UX(i, j, k - 1) + 0.5 * UX(i, j, k);
    const double ux_u = +0.5 * UX(i + 1, j, k - 1) + 0.5 * UX(i + 1, j, k);
    const double l = +0.5 / DXF(i) * ux_l;
    const double u = -0.5 / DXF(i) * ux_u;
    const double c = -l - u;
    srca[cnt] += +l * UZ(i - 1, j, k) + c * UZ(i, j, k) + u * UZ(i + 1, j, k);
  END_C
  BEGIN_C
    const double uy_l = +0.5 * UY(i, j, k - 1) + 0.5 * UY(i, j, k);
    const double uy_u = +0.5 * UY(i, j + 1, k - 1) + 0.5 * UY(i, j + 1, k);
    const double l = +0.5 / dy * uy_l;
    const double u = -0.5 / dy * uy_u;
    const double c = -l - u;
    srca[cnt] += +l * UZ(i, j - 1, k) + c * UZ(i, j, k) + u * UZ(i, j + 1, k);
  END_C
  BEGIN_C
    const double uz_l = +0.5 * UZ(i, j, k - 1) + 0.5 * UZ(i, j, k);
    const double uz_u = +0.5 * UZ(i, j, k) + 0.5 * UZ(i, j, k + 1);
    const double l = +0.5 / dz * uz_l;
    const double u = -0.5 / dz * uz_u;
    const double c = -l - u;
    srca[cnt] += +l * UZ(i, j, k - 1) + c * UZ(i, j, k) + u * UZ(i, j, k + 1);
  END_C
  BEGIN_C
    srcg[cnt] += dif * (lapx[3 * (i - 1) + 0] * UZ(i - 1, j, k) + lapx[3 * (i - 1) + 1] * UZ(i, j, k)
                        + lapx[3 * (i - 1) + 2] * UZ(i + 1, j, k));
  END_C
  BEGIN_C
    srca[cnt] += dif * (lapy[0] * UZ(i, j - 1, k) + lapy[1] * UZ(i, j, k) + lapy[2] * UZ(i, j + 1, k));
  END_C
  BEGIN_C
    srca[cnt] += dif * (lapz[0] * UZ(i, j, k - 1) + lapz[1] * UZ(i, j, k) + lapz[2] * UZ(i, j, k + 1));
  END_C
  BEGIN_C
    srcg[cnt] -= 1. / dz * (-P(i, j, k - 1) + P(i, j, k));
  END_C
  BEGIN_C
    srca[cnt] += o->ifrcz[cnt];
  END_C
}
#endif

/* compute_rhs_t: predict/t.c:68-182 */
static void compute_rhs_t(orc_t * o) {
  SIZES;
  const double * dxf = o->dxf;
  const double dy = o->dy;
  const double * ux = o->ux, * uy = o->uy, * t = o->t;
  double * srca = o->srct[rk_a], * srcg = o->srct[rk_g];
  const double dif = o->t_dif;
  const double * lapx = o->lapx_c;
  const double lapy[3] = {+1. / dy / dy, -2. / dy / dy, +1. / dy / dy};
  BEGIN_C
    const double l = +0.5 / DXF(i) * UX(i, j, k);
    const double u = -0.5 / DXF(i) * UX(i + 1, j, k);
    const double c = -l - u;
    srca[cnt] += +l * T(i - 1, j, k) + c * T(i, j, k) + u * T(i + 1, j, k);
  END_C
  BEGIN_C
    const double l = +0.5 / dy * UY(i, j, k);
    const double u = -0.5 / dy * UY(i, j + 1, k);
    const double c = -l - u;
    srca[cnt] += +l * T(i, j - 1, k) + c * T(i, j, k) + u * T(i, j + 1, k);
  END_C
#if NDIMS == 3
  const double dz = o->dz;
  const double * uz = o->uz;
  const double lapz[3] = {+1. / dz / dz, -2. / dz / dz, +1. / dz / dz};
  BEGIN_C
    const double l = +0.5 / dz * UZ(i, j, k);
    const double u = -0.5 / dz * UZ(i, j, k + 1);
    const double c = -l - u;
    srca[cnt] += +l * T(i, j, k - 1) + c * T(i, j, k) + u * T(i, j, k + 1);
  END_C
#endif
  BEGIN_C
    srcg[cnt] += dif * (lapx[3 * (i - 1) + 0] * T(i - 1, j, k) + lapx[3 * (i - 1) + 1] * T(i, j, k)
                        + lapx[3 * (i - 1) + 2] * T(i + 1, j, k));
  END_C
  BEGIN_C
    srca[cnt] += dif * (lapy[0] * T(i, j - 1, k) + lapy[1] * T(i, j, k) + lapy[2] * T(i, j + 1, k));
  END_C
#if NDIMS == 3
  BEGIN_C
    srca[cnt] += dif * (lapz[0] * T(i, j, k - 1) + lapz[1] * T(i, j, k) + lapz[2] * T(i, j, k + 1));
  END_C
#endif
}

/* increments + implicit x solve (tdm.solve via solve_in_x): ux.c:251-272,
 * 306-388 (and the uy, t counterparts) */
static void predict_one(orc_t * o, int rkstep, double dt, double * const * srcs,
                        int n, int nlines, double diffusivity, const double * lapx) {
  const double coef_a = rkcoefs[rkstep][rk_a];
  const double coef_b = rkcoefs[rkstep][rk_b];
  const double coef_g = rkcoefs[rkstep][rk_g];
  const double * sa = srcs[rk_a], * sb = srcs[rk_b], * sg = srcs[rk_g];
  double * d = o->x1pncl;
  const size_t nitems = (size_t)n * nlines;
  for (size_t m = 0; m < nitems; m++) {
    d[m] = +coef_a * dt * sa[m] + coef_b * dt * sb[m] + coef_g * dt * sg[m];
  }
  const double prefactor = 0.5 * rkcoefs[rkstep][rk_g] * dt * diffusivity;
  double * l = malloc(sizeof(double) * n), * c = malloc(sizeof(double) * n);
  double * u = malloc(sizeof(double) * n), * v = malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) {
    l[i] = -prefactor * lapx[3 * i + 0];
    c[i] = 1. - prefactor * lapx[3 * i + 1];
    u[i] = -prefactor * lapx[3 * i + 2];
  }
  tdm_solve_real(n, nlines, 0, l, c, u, v, d, NULL);
  free(l); free(c); free(u); free(v);
}

int orc_fluid_predict_field(orc_t * o, int rkstep, double dt) {
  SIZES;
  const size_t nux = orc_field_count(o, "srcux_a");
  const size_t nc = orc_field_count(o, "srct_a");
  reset_srcs3(rkstep, &o->srcux[0], &o->srcux[1], &o->srcux[2], nux);
  reset_srcs3(rkstep, &o->srcuy[0], &o->srcuy[1], &o->srcuy[2], nc);
#if NDIMS == 3
  reset_srcs3(rkstep, &o->srcuz[0], &o->srcuz[1], &o->srcuz[2], nc);
#endif
  reset_srcs3(rkstep, &o->srct[0], &o->srct[1], &o->srct[2], nc);
  compute_rhs_ux(o);
  compute_rhs_uy(o);
#if NDIMS == 3
  compute_rhs_uz(o);
#endif
  compute_rhs_t(o);
  const int nlines = jsize * ksize;
  /* ux */
  predict_one(o, rkstep, dt, o->srcux, isize - 1, nlines, o->m_dif, o->lapx_ux);
  {
    double * ux = o->ux; const double * d = o->x1pncl; size_t cnt = 0;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++, cnt++) UX(i, j, k) += d[cnt];
    bc_ux(o);
  }
  /* uy */
  predict_one(o, rkstep, dt, o->srcuy, isize, nlines, o->m_dif, o->lapx_c);
  {
    double * uy = o->uy; const double * d = o->x1pncl; size_t cnt = 0;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++, cnt++) UY(i, j, k) += d[cnt];
    bc_uy(o);
  }
#if NDIMS == 3
  predict_one(o, rkstep, dt, o->srcuz, isize, nlines, o->m_dif, o->lapx_c);
  {
    double * uz = o->uz; const double * d = o->x1pncl; size_t cnt = 0;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++, cnt++) UZ(i, j, k) += d[cnt];
    bc_uz(o);
  }
#endif
  /* t */
  predict_one(o, rkstep, dt, o->srct, isize, nlines, o->t_dif, o->lapx_c);
  {
    double * t = o->t; const double * d = o->x1pncl; size_t cnt = 0;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++, cnt++) T(i, j, k) += d[cnt];
    bc_t(o);
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Poisson: src/fluid/compute_potential/{dct,dft}.c */

/* assign_input: dct.c:285-315 / dft.c:315-345 */
static void assign_input(orc_t * o, int rkstep, double dt, double norm, double * rhs) {
  SIZES;
  const double * dxf = o->dxf;
  const double dy = o->dy;
  const double * ux = o->ux, * uy = o->uy;
#if NDIMS == 3
  const double dz = o->dz; const double * uz = o->uz;
#endif
  const double prefactor = 1. / (rkcoefs[rkstep][rk_g] * dt) / norm;
  size_t cnt = 0;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++, cnt++) {
    const double dx = DXF(i);
    const double ux_xm = UX(i, j, k), ux_xp = UX(i + 1, j, k);
    const double uy_ym = UY(i, j, k), uy_yp = UY(i, j + 1, k);
#if NDIMS == 2
    rhs[cnt] = prefactor * ((ux_xp - ux_xm) / dx + (uy_yp - uy_ym) / dy);
#else
    const double uz_zm = UZ(i, j, k), uz_zp = UZ(i, j, k + 1);
    rhs[cnt] = prefactor * ((ux_xp - ux_xm) / dx + (uy_yp - uy_ym) / dy + (uz_zp - uz_zm) / dz);
#endif
  }
}

/* extract_output: dct.c:317-332 */
static void extract_output(orc_t * o, const double * rhs) {
  SIZES;
  double * psi = o->psi;
  size_t cnt = 0;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++, cnt++) PSI(i, j, k) = rhs[cnt];
  bc_neumann(o, o->psi);
}

#if NDIMS == 2
/* fluid_compute_potential_dct: dct.c:373-421 (x uniform) */
static void potential_dct(orc_t * o, int rkstep, double dt) {
  const int nx = o->isize, ny = o->jsize;
  double * buf0 = o->pbuf0, * buf1 = o->pbuf1;
  assign_input(o, rkstep, dt, 2. * nx, buf0);
  double * ct = dct_table(nx);
  for (int j = 0; j < ny; j++) redft10_line(nx, ct, buf0 + (size_t)j * nx, buf1 + (size_t)j * nx);
  /* x1 -> y1 transpose: buf0[i][j] */
  for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) buf0[(size_t)i * ny + j] = buf1[(size_t)j * nx + i];
  /* eigenvalues: dct.c:233-261 */
  const double dx = o->lengths[0] / nx;
  double * l = malloc(sizeof(double) * ny), * u = malloc(sizeof(double) * ny);
  double * c = malloc(sizeof(double) * ny), * v = malloc(sizeof(double) * ny);
  double * q1 = malloc(sizeof(double) * ny);
  for (int j = 0; j < ny; j++) { l[j] = 1. / o->dy / o->dy; u[j] = 1. / o->dy / o->dy; }
  for (int m = 0; m < nx; m++) {
    const double s = sin(g_pi * m / (2. * nx));
    const double eval = -4. / (dx * dx) * (s * s);
    for (int n = 0; n < ny; n++) c[n] = -l[n] - u[n] + eval;
    tdm_solve_real(ny, 1, 1, l, c, u, v, buf0 + (size_t)m * ny, q1);
  }
  free(l); free(u); free(c); free(v); free(q1);
  for (int j = 0; j < ny; j++) for (int i = 0; i < nx; i++) buf1[(size_t)j * nx + i] = buf0[(size_t)i * ny + j];
  for (int j = 0; j < ny; j++) redft01_line(nx, ct, buf1 + (size_t)j * nx, buf0 + (size_t)j * nx);
  free(ct);
  extract_output(o, buf0);
}
#endif

/* fluid_compute_potential_dft: dft.c:406-468; 3D adds a complex transform in
 * z between the y transform and the x solve, eigenvalues lambda_y + lambda_z */
static void potential_dft(orc_t * o, int rkstep, double dt) {
  const int nx = o->isize, ny = o->jsize, nz = o->ksize;
  const int nh = ny / 2 + 1;
  double * buf0 = o->pbuf0, * buf1 = o->pbuf1;
  assign_input(o, rkstep, dt, (double)ny * (NDIMS == 3 ? (double)nz : 1.), buf0);
  /* y transform per (i, k): spectral layout spec[k][ky][i] (x1 pencil) */
  twiddle_t twy; twiddle_make(&twy, ny);
  const int nw = ny > nz ? ny : nz;
  double * line = malloc(sizeof(double) * ny);
  double * cl = malloc(sizeof(double) * 2 * nw);
  double * w = malloc(sizeof(double) * 6 * nw);
  double * spec = buf1;
  for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) {
    for (int j = 0; j < ny; j++) line[j] = buf0[(size_t)i + (size_t)nx * (j + (size_t)ny * k)];
    rfft_line(&twy, ny, line, cl, w, w + 2 * nw, w + 4 * nw);
    for (int m = 0; m < nh; m++) {
      const size_t idx = 2 * ((size_t)i + (size_t)nx * (m + (size_t)nh * k));
      spec[idx] = cl[2 * m]; spec[idx + 1] = cl[2 * m + 1];
    }
  }
#if NDIMS == 3
  twiddle_t twz; twiddle_make(&twz, nz);
  for (int m = 0; m < nh; m++) for (int i = 0; i < nx; i++) {
    for (int k = 0; k < nz; k++) {
      const size_t idx = 2 * ((size_t)i + (size_t)nx * (m + (size_t)nh * k));
      cl[2 * k] = spec[idx]; cl[2 * k + 1] = spec[idx + 1];
    }
    dft(&twz, nz, -1, cl, w, w + 2 * nw);
    for (int k = 0; k < nz; k++) {
      const size_t idx = 2 * ((size_t)i + (size_t)nx * (m + (size_t)nh * k));
      spec[idx] = cl[2 * k]; spec[idx + 1] = cl[2 * k + 1];
    }
  }
#endif
  /* eigenvalues: dft.c:263-291 */
  double * ey = malloc(sizeof(double) * nh);
  {
    const double gy = o->lengths[1] / ny;
    for (int m = 0; m < nh; m++) {
      const double s = sin(g_pi * m / (double)ny);
      ey[m] = -4. / (gy * gy) * (s * s);
    }
  }
  double * ez = malloc(sizeof(double) * nz);
  for (int kz = 0; kz < nz; kz++) ez[kz] = 0.;
#if NDIMS == 3
  {
    const double gz = o->lengths[2] / nz;
    for (int kz = 0; kz < nz; kz++) {
      const double s = sin(g_pi * kz / (double)nz);
      ez[kz] = -4. / (gz * gz) * (s * s);
    }
  }
#endif
  /* tri-diagonal systems in x: dft.c:160-190, 364-395 */
  const double * dxf = o->dxf, * dxc = o->dxc;
  double * l = malloc(sizeof(double) * nx), * u = malloc(sizeof(double) * nx);
  double * c = malloc(sizeof(double) * nx), * v = malloc(sizeof(double) * nx);
  for (int i = 1; i <= nx; i++) {
    l[i - 1] = 1. / DXC(i) / DXF(i);
    u[i - 1] = 1. / DXC(i + 1) / DXF(i);
  }
  for (int kz = 0; kz < nz; kz++) for (int m = 0; m < nh; m++) {
#if NDIMS == 2
    const double eval = ey[m];
#else
    const double eval = ey[m] + ez[kz];
#endif
    for (int n = 0; n < nx; n++) c[n] = -l[n] - u[n] + eval;
    c[0] += l[0];
    c[nx - 1] += u[nx - 1];
    tdm_solve_cplx(nx, 1, 0, l, c, u, v, spec + 2 * (size_t)nx * (m + (size_t)nh * kz), NULL);
  }
  free(l); free(u); free(c); free(v); free(ey); free(ez);
#if NDIMS == 3
  for (int m = 0; m < nh; m++) for (int i = 0; i < nx; i++) {
    for (int k = 0; k < nz; k++) {
      const size_t idx = 2 * ((size_t)i + (size_t)nx * (m + (size_t)nh * k));
      cl[2 * k] = spec[idx]; cl[2 * k + 1] = spec[idx + 1];
    }
    dft(&twz, nz, +1, cl, w, w + 2 * nw);
    for (int k = 0; k < nz; k++) {
      const size_t idx = 2 * ((size_t)i + (size_t)nx * (m + (size_t)nh * k));
      spec[idx] = cl[2 * k]; spec[idx + 1] = cl[2 * k + 1];
    }
  }
  twiddle_free(&twz);
#endif
  for (int k = 0; k < nz; k++) for (int i = 0; i < nx; i++) {
    for (int m = 0; m < nh; m++) {
      const size_t idx = 2 * ((size_t)i + (size_t)nx * (m + (size_t)nh * k));
      cl[2 * m] = spec[idx]; cl[2 * m + 1] = spec[idx + 1];
    }
    irfft_line(&twy, ny, cl, line, w, w + 2 * nw, w + 4 * nw);
    for (int j = 0; j < ny; j++) buf0[(size_t)i + (size_t)nx * (j + (size_t)ny * k)] = line[j];
  }
  twiddle_free(&twy);
  free(line); free(cl); free(w);
  extract_output(o, buf0);
}

int orc_fluid_compute_potential(orc_t * o, int rkstep, double dt, int use_dct) {
#if NDIMS == 2
  if (use_dct) { potential_dct(o, rkstep, dt); return 0; }
#else
  (void)use_dct;
#endif
  potential_dft(o, rkstep, dt);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* fluid_correct_velocity: correct_velocity/{main,ux,uy}.c */
int orc_fluid_correct_velocity(orc_t * o, int rkstep, double dt) {
  SIZES;
  const double prefactor = rkcoefs[rkstep][rk_g] * dt;
  const double * dxc = o->dxc;
  const double dy = o->dy;
  const double * psi = o->psi;
  {
    double * ux = o->ux;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++) {
      const double dx = DXC(i);
      UX(i, j, k) -= prefactor / dx * (+PSI(i, j, k) - PSI(i - 1, j, k));
    }
    bc_ux(o);
  }
  {
    double * uy = o->uy;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      UY(i, j, k) -= prefactor / dy * (+PSI(i, j, k) - PSI(i, j - 1, k));
    }
    bc_uy(o);
  }
#if NDIMS == 3
  {
    const double dz = o->dz;
    double * uz = o->uz;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      UZ(i, j, k) -= prefactor / dz * (+PSI(i, j, k) - PSI(i, j, k - 1));
    }
    bc_uz(o);
  }
#endif
  return 0;
}

/* fluid_update_pressure: update_pressure.c:12-109 (x implicit, y/z explicit) */
int orc_fluid_update_pressure(orc_t * o, int rkstep, double dt) {
  SIZES;
  const double * psi = o->psi;
  double * p = o->p;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) P(i, j, k) += PSI(i, j, k);
  const double prefactor = 0.5 * rkcoefs[rkstep][rk_g] * dt * o->m_dif;
  const double * dxf = o->dxf, * dxc = o->dxc;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    const double dpsidx_xm = (-PSI(i - 1, j, k) + PSI(i, j, k)) / DXC(i);
    const double dpsidx_xp = (-PSI(i, j, k) + PSI(i + 1, j, k)) / DXC(i + 1);
    P(i, j, k) -= prefactor / DXF(i) * (-dpsidx_xm + dpsidx_xp);
  }
  bc_neumann(o, o->p);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* integrate: src/integrate.c:11-68 */
int orc_integrate(orc_t * o, double * dt) {
  if (0 != orc_decide_dt(o, dt)) return 1;
  for (int rkstep = 0; rkstep < 3; rkstep++) {
    if (orc_interface_compute_curvature_tensor(o)) return 1;
    if (orc_interface_compute_force(o)) return 1;
    if (orc_interface_update_vof(o, rkstep, *dt)) return 1;
    if (orc_fluid_predict_field(o, rkstep, *dt)) return 1;
    if (orc_fluid_compute_potential(o, rkstep, *dt, o->x_uniform)) return 1;
    if (orc_fluid_correct_velocity(o, rkstep, *dt)) return 1;
    if (orc_fluid_update_pressure(o, rkstep, *dt)) return 1;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* diagnostics: logging/divergence.c:37-55 */
double orc_max_divergence(const orc_t * o) {
  SIZES;
  const double * dxf = o->dxf;
  const double dy = o->dy;
  const double * ux = o->ux, * uy = o->uy;
#if NDIMS == 3
  const double * uz = o->uz; const double dz = o->dz;
#endif
  double divmax = 0.;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    const double dx = DXF(i);
    double div = (UX(i + 1, j, k) - UX(i, j, k)) / dx + (UY(i, j + 1, k) - UY(i, j, k)) / dy;
#if NDIMS == 3
    div += (UZ(i, j, k + 1) - UZ(i, j, k)) / dz;
#endif
    divmax = fmax(divmax, fabs(div));
  }
  return divmax;
}

/* logging/interface.c:29-51 */
void orc_vof_stats(const orc_t * o, double * vmin, double * vmax, double * volume) {
  SIZES;
  const double * dxf = o->dxf;
  const double cell_yz = o->dy * (NDIMS == 3 ? o->dz : 1.);
  const double * vof = o->vof;
  double mn = 1., mx = 0., s0 = 0., s1 = 0.;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    const double cellsize = DXF(i) * cell_yz;
    const double lvof = VOF(i, j, k);
    mn = fmin(mn, lvof);
    mx = fmax(mx, lvof);
    s0 += lvof * cellsize;
    s1 += cellsize;
  }
  *vmin = mn; *vmax = mx; *volume = s0 / s1;
}

/* logging/energy.c:43-65 */
void orc_energy(const orc_t * o, double * out) {
  SIZES;
  const double * dxf = o->dxf, * dxc = o->dxc;
  const double dy = o->dy, dz = o->dz;
  (void)dz;
  const double * ux = o->ux, * uy = o->uy, * t = o->t;
  double q[4] = {0., 0., 0., 0.};
#if NDIMS == 2
#define CELLSIZE(DX) ((DX) * dy)
#else
#define CELLSIZE(DX) ((DX) * dy * dz)
#endif
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++) {
    q[0] += 0.5 * (UX(i, j, k) * UX(i, j, k)) * CELLSIZE(DXC(i));
  }
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    q[1] += 0.5 * (UY(i, j, k) * UY(i, j, k)) * CELLSIZE(DXF(i));
  }
#if NDIMS == 3
  const double * uz = o->uz;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    q[2] += 0.5 * (UZ(i, j, k) * UZ(i, j, k)) * CELLSIZE(DXF(i));
  }
#endif
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    q[NDIMS] += 0.5 * (T(i, j, k) * T(i, j, k)) * CELLSIZE(DXF(i));
  }
#undef CELLSIZE
  for (int d = 0; d <= NDIMS; d++) out[d] = q[d];
}

/* logging/momentum.c:39-53: x momentum on the x faces i = 2..nx (cell size
 * DXC), y (and z) momentum on the cell centres (cell size DXF) */
void orc_momentum(const orc_t * o, double * out) {
  SIZES;
  const double * dxf = o->dxf, * dxc = o->dxc;
  const double dy = o->dy, dz = o->dz;
  (void)dz;
  const double * ux = o->ux, * uy = o->uy;
  double q[3] = {0., 0., 0.};
#if NDIMS == 2
#define CELLSIZE(DX) ((DX) * dy)
#else
#define CELLSIZE(DX) ((DX) * dy * dz)
#endif
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++) {
    q[0] += UX(i, j, k) * CELLSIZE(DXC(i));
  }
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    q[1] += UY(i, j, k) * CELLSIZE(DXF(i));
  }
#if NDIMS == 3
  const double * uz = o->uz;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    q[2] += UZ(i, j, k) * CELLSIZE(DXF(i));
  }
#endif
#undef CELLSIZE
  for (int d = 0; d < NDIMS; d++) out[d] = q[d];
}

/* interface measure sum |grad vof| dV (length in 2D, area in 3D) with
 * central differences at the cell centres; the wall values of vof (0) are
 * the outer neighbours at i = 1 and nx.  For a planar interface normal to
 * x it returns the wall area exactly (the three non-zero differences
 * telescope to 1/(2 dx) each side). */
double orc_interface_area(const orc_t * o) {
  SIZES;
  const double * dxf = o->dxf, * dxc = o->dxc;
  const double dy = o->dy, dz = o->dz;
  (void)dz;
  const double * vof = o->vof;
  double s = 0.;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
    const double gx = (VOF(i + 1, j, k) - VOF(i - 1, j, k)) / (DXC(i) + DXC(i + 1));
    const double gy = (VOF(i, j + 1, k) - VOF(i, j - 1, k)) / (2. * dy);
#if NDIMS == 2
    const double g2 = gx * gx + gy * gy;
    s += sqrt(g2) * (DXF(i) * dy);
#else
    const double gz = (VOF(i, j, k + 1) - VOF(i, j, k - 1)) / (2. * dz);
    const double g2 = gx * gx + gy * gy + gz * gz;
    s += sqrt(g2) * (DXF(i) * dy * dz);
#endif
  }
  return s;
}

/* ---- logging/nusselt (heat_flux.c, kinetic_energy_*.c, thermal_*.c): four Nusselt numbers ---------------------------
 * 2D as the reference; 3D adds the z terms of each integral and the
 * reference flux t_dif * ly * lz.  out = {heat flux on the walls, kinetic
 * energy injection, kinetic energy dissipation, thermal dissipation}. */
void orc_nusselt(const orc_t * o, double * out) {
  SIZES;
  const double * dxf = o->dxf, * dxc = o->dxc;
  const double dy = o->dy, dz = o->dz;
  (void)dz;
  const double * ux = o->ux, * uy = o->uy, * t = o->t;
#if NDIMS == 3
  const double * uz = o->uz;
  const double dyz = dy * dz;
  const double ref = o->t_dif * o->lengths[1] * o->lengths[2];
#else
  const double dyz = dy;
  const double ref = o->t_dif * o->lengths[1];
#endif
  const double md = o->m_dif, td = o->t_dif;
  /* heat_flux.c:20-45 */
  double hf = 0.;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) {
    const double dtdx_xm = (+T(1, j, k) - T(0, j, k)) / DXC(1);
    const double dtdx_xp = (+T(isize + 1, j, k) - T(isize, j, k)) / DXC(isize + 1);
    hf -= 0.5 * td * (dtdx_xm + dtdx_xp) * dyz;
  }
  out[0] = hf / ref;
  /* kinetic_energy_injection.c:14-45 */
  double inj = 0.;
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 2; i <= isize; i++) {
    const double t_ = +0.5 * T(i - 1, j, k) + 0.5 * T(i, j, k);
    inj += UX(i, j, k) * t_ * (DXC(i) * dyz);
  }
  out[1] = inj / ref + 1.;
  /* kinetic_energy_dissipation.c:14-160, one accumulator per term */
  double kd = 0.;
  {
    double s = 0.;   /* ux, x */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize + 1; i++) {
      const double cs = DXC(i) * dyz;
      if (1 != i) {
        const double d = 1. / DXF(i - 1) * (UX(i, j, k) - UX(i - 1, j, k));
        s += md * 0.5 * (DXF(i - 1) / DXC(i)) * (d * d) * cs;
      }
      if (isize + 1 != i) {
        const double d = 1. / DXF(i) * (UX(i + 1, j, k) - UX(i, j, k));
        s += md * 0.5 * (DXF(i) / DXC(i)) * (d * d) * cs;
      }
    }
    kd += s;
  }
  {
    double s = 0.;   /* ux, y */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize + 1; i++) {
      const double cs = DXC(i) * dyz;
      const double dm = 1. / dy * (UX(i, j, k) - UX(i, j - 1, k));
      s += md * 0.5 * (dm * dm) * cs;
      const double dp = 1. / dy * (UX(i, j + 1, k) - UX(i, j, k));
      s += md * 0.5 * (dp * dp) * cs;
    }
    kd += s;
  }
  {
    double s = 0.;   /* uy, x */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / DXC(i) * (UY(i, j, k) - UY(i - 1, j, k));
      s += md * 0.5 * (DXC(i) / DXF(i)) * (1 == i ? 2. : 1.) * (dm * dm) * cs;
      const double dp = 1. / DXC(i + 1) * (UY(i + 1, j, k) - UY(i, j, k));
      s += md * 0.5 * (DXC(i + 1) / DXF(i)) * (isize == i ? 2. : 1.) * (dp * dp) * cs;
    }
    kd += s;
  }
  {
    double s = 0.;   /* uy, y */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / dy * (UY(i, j, k) - UY(i, j - 1, k));
      s += md * 0.5 * (dm * dm) * cs;
      const double dp = 1. / dy * (UY(i, j + 1, k) - UY(i, j, k));
      s += md * 0.5 * (dp * dp) * cs;
    }
    kd += s;
  }
#if NDIMS == 3
  {
    double s = 0.;   /* ux, z */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize + 1; i++) {
      const double cs = DXC(i) * dyz;
      const double dm = 1. / dz * (UX(i, j, k) - UX(i, j, k - 1));
      s += md * 0.5 * (dm * dm) * cs;
      const double dp = 1. / dz * (UX(i, j, k + 1) - UX(i, j, k));
      s += md * 0.5 * (dp * dp) * cs;
    }
    kd += s;
  }
  {
    double s = 0.;   /* uy, z */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / dz * (UY(i, j, k) - UY(i, j, k - 1));
      s += md * 0.5 * (dm * dm) * cs;
      const double dp = 1. / dz * (UY(i, j, k + 1) - UY(i, j, k));
      s += md * 0.5 * (dp * dp) * cs;
    }
    kd += s;
  }
  {
    double s = 0.;   /* uz, x */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / DXC(i) * (UZ(i, j, k) - UZ(i - 1, j, k));
      s += md * 0.5 * (DXC(i) / DXF(i)) * (1 == i ? 2. : 1.) * (dm * dm) * cs;
      const double dp = 1. / DXC(i + 1) * (UZ(i + 1, j, k) - UZ(i, j, k));
      s += md * 0.5 * (DXC(i + 1) / DXF(i)) * (isize == i ? 2. : 1.) * (dp * dp) * cs;
    }
    kd += s;
  }
  {
    double s = 0.;   /* uz, y */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / dy * (UZ(i, j, k) - UZ(i, j - 1, k));
      s += md * 0.5 * (dm * dm) * cs;
      const double dp = 1. / dy * (UZ(i, j + 1, k) - UZ(i, j, k));
      s += md * 0.5 * (dp * dp) * cs;
    }
    kd += s;
  }
  {
    double s = 0.;   /* uz, z */
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / dz * (UZ(i, j, k) - UZ(i, j, k - 1));
      s += md * 0.5 * (dm * dm) * cs;
      const double dp = 1. / dz * (UZ(i, j, k + 1) - UZ(i, j, k));
      s += md * 0.5 * (dp * dp) * cs;
    }
    kd += s;
  }
#endif
  out[2] = kd / ref + 1.;
  /* thermal_energy_dissipation.c:14-90 */
  double th = 0.;
  {
    double s = 0.;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / DXC(i) * (T(i, j, k) - T(i - 1, j, k));
      s += td * 0.5 * (DXC(i) / DXF(i)) * (1 == i ? 2. : 1.) * (dm * dm) * cs;
      const double dp = 1. / DXC(i + 1) * (T(i + 1, j, k) - T(i, j, k));
      s += td * 0.5 * (DXC(i + 1) / DXF(i)) * (isize == i ? 2. : 1.) * (dp * dp) * cs;
    }
    th += s;
  }
  {
    double s = 0.;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / dy * (T(i, j, k) - T(i, j - 1, k));
      s += td * 0.5 * (dm * dm) * cs;
      const double dp = 1. / dy * (T(i, j + 1, k) - T(i, j, k));
      s += td * 0.5 * (dp * dp) * cs;
    }
    th += s;
  }
#if NDIMS == 3
  {
    double s = 0.;
    KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) for (int i = 1; i <= isize; i++) {
      const double cs = DXF(i) * dyz;
      const double dm = 1. / dz * (T(i, j, k) - T(i, j, k - 1));
      s += td * 0.5 * (dm * dm) * cs;
      const double dp = 1. / dz * (T(i, j, k + 1) - T(i, j, k));
      s += td * 0.5 * (dp * dp) * cs;
    }
    th += s;
  }
#endif
  out[3] = th / ref;
}

/* ---- statistics.c:117-223: time-averaging accumulators -------------------
 * ux1 ux2 uxt on the x faces i = 1..nx+1, uy1 uy2 (uz1 uz2) t1 t2 on the
 * cells i = 0..nx+1, all interior rows; reduce sums over y (and z) as
 * statistics.c:232-275 */
struct orc_stats {
  size_t num;
  int wf, wc;        /* row widths: nx + 1 faces, nx + 2 cells */
  size_t rows;       /* ny (* nz) */
  double * a[9];     /* ux1 ux2 uy1 uy2 t1 t2 uxt uz1 uz2 */
};

static const char * const kStatNames[9] = {"ux1", "ux2", "uy1", "uy2", "t1", "t2", "uxt", "uz1", "uz2"};

orc_stats_t * orc_stats_create(const orc_t * o) {
  orc_stats_t * s = calloc(1, sizeof(orc_stats_t));
  s->wf = o->isize + 1;
  s->wc = o->isize + 2;
  s->rows = (size_t)o->jsize * (size_t)o->ksize;
  for (int n = 0; n < 9; n++) {
    const int w = (n == 0 || n == 1 || n == 6) ? s->wf : s->wc;
    s->a[n] = calloc((size_t)w * s->rows, sizeof(double));
  }
  return s;
}

void orc_stats_destroy(orc_stats_t * s) {
  if (!s) return;
  for (int n = 0; n < 9; n++) free(s->a[n]);
  free(s);
}

void orc_stats_collect(orc_stats_t * s, const orc_t * o) {
  SIZES;
  const double * ux = o->ux, * uy = o->uy, * t = o->t;
#if NDIMS == 3
  const double * uz = o->uz;
#endif
  KLOOP(1, ksize) for (int j = 1; j <= jsize; j++) {
    const size_t r = (size_t)(j - 1) + (size_t)jsize * (size_t)(k - 1);
    for (int i = 1; i <= isize + 1; i++) {
      const double u = UX(i, j, k);
      s->a[0][(size_t)(i - 1) + s->wf * r] += u;
      s->a[1][(size_t)(i - 1) + s->wf * r] += u * u;
      const double t_ = +0.5 * T(i - 1, j, k) + 0.5 * T(i, j, k);
      s->a[6][(size_t)(i - 1) + s->wf * r] += u * t_;
    }
    for (int i = 0; i <= isize + 1; i++) {
      const double v = UY(i, j, k);
      s->a[2][(size_t)i + s->wc * r] += v;
      s->a[3][(size_t)i + s->wc * r] += v * v;
      const double tt = T(i, j, k);
      s->a[4][(size_t)i + s->wc * r] += tt;
      s->a[5][(size_t)i + s->wc * r] += tt * tt;
#if NDIMS == 3
      const double w = UZ(i, j, k);
      s->a[7][(size_t)i + s->wc * r] += w;
      s->a[8][(size_t)i + s->wc * r] += w * w;
#endif
    }
  }
  s->num += 1;
}

size_t orc_stats_num(const orc_stats_t * s) { return s->num; }

/* sum over the rows in row order; returns the profile length or 0 */
int orc_stats_reduce(const orc_stats_t * s, const char * name, double * out) {
  for (int n = 0; n < 9; n++) {
    if (strcmp(name, kStatNames[n])) continue;
    const int w = (n == 0 || n == 1 || n == 6) ? s->wf : s->wc;
    for (int i = 0; i < w; i++) out[i] = 0.;
    for (size_t r = 0; r < s->rows; r++)
      for (int i = 0; i < w; i++) out[i] += s->a[n][(size_t)i + (size_t)w * r];
    return w;
  }
  return 0;
}
