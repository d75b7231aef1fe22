// vof_kernels.h -- gfx950 device kernels of the SimpleVOFSolver time step.
//
// Compiled once per dimensionality (NDIMS = 2 or 3) inside vof_lib.hip.
// All arithmetic is FP64 and follows the reference's evaluation order; the
// library is built with -ffp-contract=off so that no FMA contraction changes
// the rounding relative to the reference's -std=c99 build.
//
// Array index maps are the reference's (include/array_macros/ of the
// reference for 2D, tools/define_arrays.py:201-229 for 3D): x fastest, then
// y, then z; each array carries its own halo padding.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#if NDIMS == 2
#define LIN(I, J, K, i0, n0, j0, n1, k0) \
  ((size_t)((I) - (i0)) + (size_t)(n0) * (size_t)((J) - (j0)))
#else
#define LIN(I, J, K, i0, n0, j0, n1, k0) \
  ((size_t)((I) - (i0)) + (size_t)(n0) * ((size_t)((J) - (j0)) + (size_t)(n1) * (size_t)((K) - (k0))))
#endif
// these expect ints nx, ny in scope
#define I_UX(I, J, K)   LIN(I, J, K, 1, nx + 1, 0, ny + 2, 0)
#define I_C1(I, J, K)   LIN(I, J, K, 0, nx + 2, 0, ny + 2, 0)   /* uy uz p psi t curv */
#define I_VOF(I, J, K)  LIN(I, J, K, 0, nx + 2, -1, ny + 4, -1)
#define I_DVOF(I, J, K) LIN(I, J, K, 1, nx + 1, 0, ny + 3, 0)
#define I_NRM(I, J, K)  LIN(I, J, K, 1, nx, 0, ny + 2, 0)
#define I_FX(I, J, K)   LIN(I, J, K, 1, nx + 1, 1, ny, 1)
#define I_FY(I, J, K)   LIN(I, J, K, 1, nx, 1, ny + 1, 1)
#define I_FZ(I, J, K)   LIN(I, J, K, 1, nx, 1, ny, 1)
#define I_SX(I, J, K)   LIN(I, J, K, 2, nx - 1, 1, ny, 1)        /* srcux, ifrcx */
#define I_SC(I, J, K)   LIN(I, J, K, 1, nx, 1, ny, 1)            /* srcuy/z/t, ifrcy/z, src */

namespace vofk {

// interface/common.c:4-14
#define NGAUSS 2
__device__ __constant__ double kGaussP[NGAUSS] = {-0.2886751345948129, +0.2886751345948129};
__device__ __constant__ double kGaussW[NGAUSS] = {+0.5, +0.5};
constexpr double kVofBeta = 1.;
constexpr double kVofMin = 1.e-8;

// Discrete decisions of the 2D step, recorded by one forward run (FZ = 1)
// and imposed on later runs (FZ = 2): the map the adjoint differentiates
// holds them fixed, so that finite differences of that map can be compared
// with it.  FZ = 0: the step as the reference computes it.
//   corner: corner gradient below the floor (1): its normal is the recorded one
//   cell:   bit 7 interface cell (normal and intercept formed), bits 0-3 the
//           number of Newton updates of its intercept
//   fx, fy: x / y face: bit 0 upwind cell is the upper one (vel < 0), bit 1
//           the upwind cell is single-phase (flux vel * vof)
struct Freeze {
  unsigned char * corner;   // I_DVOF index
  unsigned char * cell;     // I_C1 index
  unsigned char * fx;       // I_UX index of the face
  unsigned char * fy;       // I_C1 index of the face's upper cell
  double * dvof;            // corner normals of the recording run (2 per corner)
  double floor;             // |grad vof| DXC below which a corner counts as flat
};

struct Grid {
  int nx, ny, nz;
  double dy, dz;
  const double * __restrict__ dxf;  // [nx]   DXF(i) = dxf[i-1]
  const double * __restrict__ dxc;  // [nx+1] DXC(i) = dxc[i-1]
  int bands;                        // 3D tile kernels: one band of tile rows per XCD
  // quotients the stencils multiply by, computed on the host with the same
  // IEEE divisions the expressions spell (so the products are unchanged):
  // hdy = 0.5/dy, rdy = 1./dy, lyy = 1./dy/dy, cyy = -2./dy/dy (and z)
  double hdy, rdy, lyy, cyy, hdz, rdz, lzz, czz;
  // per x index i (0..nx+1), kXT entries: 0.5/DXC(i), 1./DXC(i),
  // 0.5*DXF(i-1)/DXC(i), 0.5*DXF(i)/DXC(i), 0.5/DXF(i), 1./DXF(i)
  const double * __restrict__ xt;
};
constexpr int kXT = 8;
enum { XT_HXC = 0, XT_RXC, XT_WXM, XT_WXP, XT_HXF, XT_RXF };

// Work-groups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Tile kernels renumber their work-groups so that every XCD processes
// one contiguous run of tiles (x fastest): tiles that share cache lines or
// stencil planes then meet in the same L2.  The map is a bijection of
// [0, total), so each tile is still processed exactly once.
constexpr unsigned kXcds = 8;

// Every kernel body takes its work-group coordinates and grid shape as a
// VBlock: the hardware's for a plain launch, or a case's own tile for a
// batched launch that covers many cases (one work-group = one tile of one
// case).
struct VBlock {
  unsigned x, y, z;
  unsigned gx, gy, gz;
};

__device__ inline VBlock hw_block() {
  VBlock b;
  b.x = blockIdx.x;
  b.y = blockIdx.y;
  b.z = blockIdx.z;
  b.gx = gridDim.x;
  b.gy = gridDim.y;
  b.gz = gridDim.z;
  return b;
}

struct Tile3 {
  unsigned x, y, z;
};

__device__ inline Tile3 xcd_tile(const VBlock & vb) {
  const unsigned gx = vb.gx, gy = vb.gy;
  const unsigned total = gx * gy * vb.gz;
  const unsigned w = vb.x + gx * (vb.y + gy * vb.z);
  const unsigned q = total / kXcds, r = total % kXcds;
  const unsigned xcd = w % kXcds, slot = w / kXcds;
  const unsigned l = xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
  Tile3 t;
  t.x = l % gx;
  t.y = (l / gx) % gy;
  t.z = l / (gx * gy);
  return t;
}

__device__ inline Tile3 plain_tile(const VBlock & vb) {
  Tile3 t;
  t.x = vb.x;
  t.y = vb.y;
  t.z = vb.z;
  return t;
}

// 3D tile kernels: XCD x (work-group w goes to XCD w % 8) takes the band
// of tile rows y in [x gy/8, (x+1) gy/8) over all planes, walking x fastest,
// then the rows of its band, then z.  Neighbour tiles in x, y and z then
// share one L2 (except at band edges), and the planes in flight on an XCD
// (its work-groups span a few planes of its band) fit its L2.  A bijection of
// the work-groups when gy is a multiple of 8; otherwise the plain order.
__device__ inline Tile3 band_tile(const VBlock & vb) {
  if (vb.gy % kXcds != 0 || vb.gz == 1) return plain_tile(vb);
  const unsigned w = vb.x + vb.gx * (vb.y + vb.gy * vb.z);
  const unsigned x = w % kXcds, s = w / kXcds;
  const unsigned rows = vb.gy / kXcds;
  Tile3 t;
  t.x = s % vb.gx;
  const unsigned r = s / vb.gx;
  t.y = x * rows + r % rows;
  t.z = r / rows;
  return t;
}

#if NDIMS == 2
#define KIDX 1
#define THREAD_IJK(ilo, jlo, klo)                               \
  const int i = (int)(vb.x * blockDim.x + threadIdx.x) + (ilo); \
  const int j = (int)(vb.y * blockDim.y + threadIdx.y) + (jlo); \
  const int k = 1; (void)k;
#else
#define THREAD_IJK(ilo, jlo, klo)                               \
  const Tile3 tile_ = g.bands ? band_tile(vb) : plain_tile(vb); \
  const int i = (int)(tile_.x * blockDim.x + threadIdx.x) + (ilo); \
  const int j = (int)(tile_.y * blockDim.y + threadIdx.y) + (jlo); \
  const int k = (int)tile_.z + (klo);
#endif

// ---------------------------------------------------------------------------
// boundary conditions (src/halo.c:18-76, fluid/boundary/*.c,
// interface/boundary.c:7-39)

// Every halo element of a padded array takes its periodic image.  The halo
// region is enumerated as (A) the y-halo rows of every plane and, in 3D,
// (B) the z-halo planes restricted to interior rows; each element is written
// once from its fully wrapped interior source, which this kernel never
// writes, so the copy is race free and equals the reference's y-then-z
// exchange.
__device__ __forceinline__ void k_halo_body(const VBlock & vb, double * __restrict__ a, int n0, int ny, int nz, int h) {
  const size_t t = (size_t)vb.x * blockDim.x + threadIdx.x;
  const int n1 = ny + 2 * h;
#if NDIMS == 2
  const int np = 1;
#else
  const int np = nz + 2 * h;
#endif
  const size_t cntA = (size_t)n0 * (2 * h) * np;
#if NDIMS == 2
  const size_t cntB = 0;
#else
  const size_t cntB = (size_t)n0 * ny * (2 * h);
#endif
  int x, r, p;
  if (t < cntA) {
    x = (int)(t % n0);
    const size_t rem = t / n0;
    const int rr = (int)(rem % (2 * h));
    p = (int)(rem / (2 * h));
    r = rr < h ? rr : ny + rr;
  } else if (t < cntA + cntB) {
    const size_t u = t - cntA;
    x = (int)(u % n0);
    const size_t rem = u / n0;
    r = h + (int)(rem % ny);
    const int pp = (int)(rem / ny);
    p = pp < h ? pp : nz + pp;
  } else {
    return;
  }
  const int rs = r < h ? r + ny : (r >= ny + h ? r - ny : r);
#if NDIMS == 2
  const int ps = 0;
  (void)nz;
#else
  const int ps = p < h ? p + nz : (p >= nz + h ? p - nz : p);
#endif
  a[x + (size_t)n0 * (r + (size_t)n1 * p)] = a[x + (size_t)n0 * (rs + (size_t)n1 * ps)];
}

__global__ void k_halo(double * __restrict__ a, int n0, int ny, int nz, int h) {
  k_halo_body(hw_block(), a, n0, ny, nz, h);
}

// x walls on rows [r0, r0+nr) of planes [p0, p0+np): mode 0 sets Dirichlet
// values (vlo at column xlo, vhi at xhi), mode 1 copies the neighbour
// (Neumann, p.c:15-16)
__device__ __forceinline__ void k_walls_body(const VBlock & vb, double * __restrict__ a, int n0, int n1, int r0, int nr, int p0,
                        int np, int xlo, int xhi, int mode, double vlo, double vhi) {
  const int t = (int)(vb.x * blockDim.x + threadIdx.x);
  if (t >= nr * np) return;
  const int r = r0 + t % nr;
  const int p = p0 + t / nr;
  double * row = a + (size_t)n0 * (r + (size_t)n1 * p);
  if (mode == 0) {
    row[xlo] = vlo;
    row[xhi] = vhi;
  } else {
    row[xlo] = row[xlo + 1];
    row[xhi] = row[xhi - 1];
  }
}

__global__ void k_walls(double * __restrict__ a, int n0, int n1, int r0, int nr, int p0,
                        int np, int xlo, int xhi, int mode, double vlo, double vhi) {
  k_walls_body(hw_block(), a, n0, n1, r0, nr, p0, np, xlo, xhi, mode, vlo, vhi);
}

// ---------------------------------------------------------------------------
// interface_compute_curvature_tensor: src/interface/curvature_tensor.c

// compute_gradient (:14-48): corners i = 1..nx+1, j = 0..ny+2 (k = 0..nz+2)
__device__ __forceinline__ void k_gradient_body(const VBlock & vb, Grid g, const double * __restrict__ vof, double * __restrict__ dvof) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 0, 0)
  if (i > nx + 1 || j > ny + 2) return;
  const double dx = g.dxc[i - 1];
#if NDIMS == 2
  const double a = vof[I_VOF(i - 1, j - 1, k)], b = vof[I_VOF(i, j - 1, k)];
  const double c = vof[I_VOF(i - 1, j, k)], d = vof[I_VOF(i, j, k)];
  const double gx = 1. / dx * (-a + b - c + d);
  const double gy = 1. / g.dy * (-a - b + c + d);
  const double norm = sqrt(gx * gx + gy * gy);
  const double ninv = 1. / fmax(norm, DBL_EPSILON);
  double * o = dvof + 2 * I_DVOF(i, j, k);
  o[0] = gx * ninv;
  o[1] = gy * ninv;
#else
  if (k > g.nz + 2) return;
  const double a0 = vof[I_VOF(i - 1, j - 1, k - 1)], b0 = vof[I_VOF(i, j - 1, k - 1)];
  const double c0 = vof[I_VOF(i - 1, j, k - 1)], d0 = vof[I_VOF(i, j, k - 1)];
  const double a1 = vof[I_VOF(i - 1, j - 1, k)], b1 = vof[I_VOF(i, j - 1, k)];
  const double c1 = vof[I_VOF(i - 1, j, k)], d1 = vof[I_VOF(i, j, k)];
  // per-plane partial sums in the 2D order: a z-uniform field reproduces the
  // 2D corner normals exactly
  const double sx0 = -a0 + b0 - c0 + d0, sx1 = -a1 + b1 - c1 + d1;
  const double sy0 = -a0 - b0 + c0 + d0, sy1 = -a1 - b1 + c1 + d1;
  const double sz0 = a0 + b0 + c0 + d0, sz1 = a1 + b1 + c1 + d1;
  const double gx = 1. / dx * (sx0 + sx1);
  const double gy = 1. / g.dy * (sy0 + sy1);
  const double gz = 1. / g.dz * (-sz0 + sz1);
  const double norm = sqrt(gx * gx + gy * gy + gz * gz);
  const double ninv = 1. / fmax(norm, DBL_EPSILON);
  double * o = dvof + 3 * I_DVOF(i, j, k);
  o[0] = gx * ninv;
  o[1] = gy * ninv;
  o[2] = gz * ninv;
#endif
}

__global__ void k_gradient(Grid g, const double * __restrict__ vof, double * __restrict__ dvof) {
  k_gradient_body(hw_block(), g, vof, dvof);
}

// compute_intercept (:50-92): Newton on D = exp(-2 beta d), at most 8
// iterations, residual tested after the update
template <int FZ = 0>
__device__ inline double intercept(const double vof, const double * n, int * iters = nullptr) {
  const int cntmax = 8;
  const double resmax = 1.e-12;
#if NDIMS == 2
  double e[4];
  for (int jj = 0; jj < NGAUSS; jj++)
    for (int ii = 0; ii < NGAUSS; ii++)
      e[jj * NGAUSS + ii] = exp(-2. * kVofBeta * (n[0] * kGaussP[ii] + n[1] * kGaussP[jj]));
#else
  double e[8];
  for (int kk = 0; kk < NGAUSS; kk++)
    for (int jj = 0; jj < NGAUSS; jj++)
      for (int ii = 0; ii < NGAUSS; ii++)
        e[(kk * NGAUSS + jj) * NGAUSS + ii] =
            exp(-2. * kVofBeta * (n[0] * kGaussP[ii] + n[1] * kGaussP[jj] + n[2] * kGaussP[kk]));
#endif
  double val = 1. / vof - 1.;
  // FZ = 2: exactly the recorded number of updates (no stopping test)
  const int nit = FZ == 2 ? *iters : cntmax;
  int done = cntmax;
  for (int cnt = 0; cnt < nit; cnt++) {
    double f0 = -1. * vof;
    double f1 = 0.;
#if NDIMS == 2
    for (int jj = 0; jj < NGAUSS; jj++)
      for (int ii = 0; ii < NGAUSS; ii++) {
        const double w = kGaussW[ii] * kGaussW[jj];
        const double p = e[jj * NGAUSS + ii];
        const double den = 1. / (1. + p * val);
        f0 += w * den;
        f1 -= w * p * den * den;
      }
#else
    // z weights innermost: with a zero z normal the sums reduce exactly to 2D
    for (int jj = 0; jj < NGAUSS; jj++)
      for (int ii = 0; ii < NGAUSS; ii++) {
        const double w = kGaussW[ii] * kGaussW[jj];
        double s0 = 0., s1 = 0.;
        for (int kk = 0; kk < NGAUSS; kk++) {
          const double p = e[(kk * NGAUSS + jj) * NGAUSS + ii];
          const double den = 1. / (1. + p * val);
          s0 += kGaussW[kk] * den;
          s1 += kGaussW[kk] * (w * p * den * den);
        }
        f0 += w * s0;
        f1 -= s1;
      }
#endif
    val -= f0 / f1;
    if (FZ != 2 && fabs(f0) < resmax) {
      done = cnt + 1;
      break;
    }
  }
  if (FZ == 1) *iters = done;
  return -0.5 / kVofBeta * log(val);
}

// compute_normal (:94-145) + compute_curvature (:147-176), cells
// i = 1..nx, j = 0..ny+1 (k = 0..nz+1); both read the same corner normals
__device__ __forceinline__ void k_normal_curv_body(const VBlock & vb, Grid g, const double * __restrict__ vof,
                              const double * __restrict__ dvof, double * __restrict__ normal,
                              double * __restrict__ curv) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 0, 0)
  if (i > nx || j > ny + 1) return;
#if NDIMS == 3
  if (k > g.nz + 1) return;
#endif
  const double dx = g.dxf[i - 1];
#if NDIMS == 2
  const double * d00 = dvof + 2 * I_DVOF(i, j, k);
  const double * d10 = dvof + 2 * I_DVOF(i + 1, j, k);
  const double * d01 = dvof + 2 * I_DVOF(i, j + 1, k);
  const double * d11 = dvof + 2 * I_DVOF(i + 1, j + 1, k);
  const double dnxdx = 1. / dx * (-d00[0] + d10[0] - d01[0] + d11[0]);
  const double dnydy = 1. / g.dy * (-d00[1] - d10[1] + d01[1] + d11[1]);
  curv[I_C1(i, j, k)] = 0.5 * (-dnxdx - dnydy);
  const double lvof = vof[I_VOF(i, j, k)];
  if (lvof < kVofMin || 1. - kVofMin < lvof) return;
  double nxv = d00[0] + d10[0] + d01[0] + d11[0];
  double nyv = d00[1] + d10[1] + d01[1] + d11[1];
  nxv /= dx;
  nyv /= g.dy;
  const double norm = sqrt(nxv * nxv + nyv * nyv);
  const double ninv = 1. / fmax(norm, DBL_EPSILON);
  nxv *= ninv;
  nyv *= ninv;
  const double nn[2] = {nxv, nyv};
  const double seg = intercept(lvof, nn);
  double * o = normal + 3 * I_NRM(i, j, k);
  o[0] = nxv;
  o[1] = nyv;
  o[2] = seg;
#else
  const double * d000 = dvof + 3 * I_DVOF(i, j, k);
  const double * d100 = dvof + 3 * I_DVOF(i + 1, j, k);
  const double * d010 = dvof + 3 * I_DVOF(i, j + 1, k);
  const double * d110 = dvof + 3 * I_DVOF(i + 1, j + 1, k);
  const double * d001 = dvof + 3 * I_DVOF(i, j, k + 1);
  const double * d101 = dvof + 3 * I_DVOF(i + 1, j, k + 1);
  const double * d011 = dvof + 3 * I_DVOF(i, j + 1, k + 1);
  const double * d111 = dvof + 3 * I_DVOF(i + 1, j + 1, k + 1);
  const double dnxdx = 1. / dx * ((-d000[0] + d100[0] - d010[0] + d110[0])
                                  + (-d001[0] + d101[0] - d011[0] + d111[0]));
  const double dnydy = 1. / g.dy * ((-d000[1] - d100[1] + d010[1] + d110[1])
                                    + (-d001[1] - d101[1] + d011[1] + d111[1]));
  const double dnzdz = 1. / g.dz * (-(d000[2] + d100[2] + d010[2] + d110[2])
                                    + (d001[2] + d101[2] + d011[2] + d111[2]));
  curv[I_C1(i, j, k)] = 0.25 * (-dnxdx - dnydy - dnzdz);
  const double lvof = vof[I_VOF(i, j, k)];
  if (lvof < kVofMin || 1. - kVofMin < lvof) return;
  double nv[3];
  for (int d = 0; d < 3; d++) {
    nv[d] = (d000[d] + d100[d] + d010[d] + d110[d]) + (d001[d] + d101[d] + d011[d] + d111[d]);
  }
  nv[0] /= dx;
  nv[1] /= g.dy;
  nv[2] /= g.dz;
  const double norm = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  const double ninv = 1. / fmax(norm, DBL_EPSILON);
  nv[0] *= ninv;
  nv[1] *= ninv;
  nv[2] *= ninv;
  const double seg = intercept(lvof, nv);
  double * o = normal + 4 * I_NRM(i, j, k);
  o[0] = nv[0];
  o[1] = nv[1];
  o[2] = nv[2];
  o[3] = seg;
#endif
}

__global__ void k_normal_curv(Grid g, const double * __restrict__ vof,
                              const double * __restrict__ dvof, double * __restrict__ normal,
                              double * __restrict__ curv) {
  k_normal_curv_body(hw_block(), g, vof, dvof, normal, curv);
}

// ---------------------------------------------------------------------------
// interface_compute_force: src/interface/force.c:11-76 (cells 1..nx x 1..ny)
__device__ __forceinline__ void k_force_body(const VBlock & vb, Grid g, double tension, const double * __restrict__ vof,
                        const double * __restrict__ curv, double * __restrict__ ifrcx,
                        double * __restrict__ ifrcy, double * __restrict__ ifrcz) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
  const double v = vof[I_VOF(i, j, k)];
  const double cv = curv[I_C1(i, j, k)];
  if (i >= 2) {
    const double dx = g.dxc[i - 1];
    const double grad = 1. / dx * (-vof[I_VOF(i - 1, j, k)] + v);
    const double kappa = 0.5 * (curv[I_C1(i - 1, j, k)] + cv);
    ifrcx[I_SX(i, j, k)] = tension * grad * kappa;
  }
  {
    const double grad = 1. / g.dy * (-vof[I_VOF(i, j - 1, k)] + v);
    const double kappa = 0.5 * (curv[I_C1(i, j - 1, k)] + cv);
    ifrcy[I_SC(i, j, k)] = tension * grad * kappa;
  }
#if NDIMS == 3
  {
    const double grad = 1. / g.dz * (-vof[I_VOF(i, j, k - 1)] + v);
    const double kappa = 0.5 * (curv[I_C1(i, j, k - 1)] + cv);
    ifrcz[I_SC(i, j, k)] = tension * grad * kappa;
  }
#else
  (void)ifrcz;
#endif
}

__global__ void k_force(Grid g, double tension, const double * __restrict__ vof,
                        const double * __restrict__ curv, double * __restrict__ ifrcx,
                        double * __restrict__ ifrcy, double * __restrict__ ifrcz) {
  k_force_body(hw_block(), g, tension, vof, curv, ifrcx, ifrcy, ifrcz);
}

// ---------------------------------------------------------------------------
// interface_update_vof: src/interface/update/{main,flxx,flxy}.c

// indicator (update/main.c:16-26)
__device__ inline double indicator(const double * n, const double * x) {
#if NDIMS == 2
  return 1. / (1. + exp(-2. * kVofBeta * (n[0] * x[0] + n[1] * x[1] + n[2])));
#else
  return 1. / (1. + exp(-2. * kVofBeta * (n[0] * x[0] + n[1] * x[1] + n[2] * x[2] + n[3])));
#endif
}

// THINC flux through a face normal to direction D (0 x, 1 y, 2 z) with face
// velocity vel between the cells m (lower index) and p: the upwind cell's
// vof, or where it holds an interface its indicator integrated over the face
// with the 2-point Gauss rule (compute_flux_x flxx.c:11-44, compute_flux_y
// flxy.c:11-44 and the z counterpart; the same quadrature order).  Both
// cells' vof are loaded by the caller, so the upwind choice does not wait on
// the velocity; only the normal is read through the choice.
template <int D, int FZ = 0>
__device__ inline double face_flux(double vel, double vm, double vp, const double * nm, const double * np,
                                   unsigned char * dec = nullptr) {
  bool up, sp;   // upwind cell is p; the upwind cell is single-phase
  if (FZ == 2) {
    const unsigned b = *dec;
    up = (b & 1u) != 0;
    sp = (b & 2u) != 0;
  } else {
    up = vel < 0.;
    const double lv = up ? vp : vm;
    sp = lv < kVofMin || 1. - kVofMin < lv;
    if (FZ == 1) *dec = (unsigned char)((up ? 1u : 0u) | (sp ? 2u : 0u));
  }
  const double lvof = up ? vp : vm;
  if (sp) return vel * lvof;
  const double * n = up ? np : nm;
  const double c = up ? -0.5 : +0.5;
  double flux = 0.;
#if NDIMS == 2
  for (int q = 0; q < NGAUSS; q++) {
    const double pt[2] = {D == 0 ? c : kGaussP[q], D == 0 ? kGaussP[q] : c};
    flux += kGaussW[q] * indicator(n, pt);
  }
#else
  for (int q = 0; q < NGAUSS; q++) {
    double sr = 0.;
    for (int r = 0; r < NGAUSS; r++) {
      // x faces: (c, P[q], P[r]); y faces: (P[q], c, P[r]); z faces: (P[q], P[r], c)
      const double pt[3] = {D == 0 ? c : kGaussP[q], D == 0 ? kGaussP[q] : D == 1 ? c : kGaussP[r],
                            D == 2 ? c : kGaussP[r]};
      sr += kGaussW[r] * indicator(n, pt);
    }
    flux += kGaussW[q] * sr;
  }
#endif
  return vel * flux;
}

// x face i (2..nx)
__device__ inline double flux_x_at(const double * __restrict__ ux, const double * __restrict__ vof,
                                   const double * __restrict__ normal, int nx, int ny, int i, int j,
                                   int k) {
  constexpr int NC = NDIMS + 1;
  return face_flux<0>(ux[I_UX(i, j, k)], vof[I_VOF(i - 1, j, k)], vof[I_VOF(i, j, k)],
                      normal + NC * I_NRM(i - 1, j, k), normal + NC * I_NRM(i, j, k));
}

// y face j (1..ny+1)
__device__ inline double flux_y_at(const double * __restrict__ uy, const double * __restrict__ vof,
                                   const double * __restrict__ normal, int nx, int ny, int i, int j,
                                   int k) {
  constexpr int NC = NDIMS + 1;
  return face_flux<1>(uy[I_C1(i, j, k)], vof[I_VOF(i, j - 1, k)], vof[I_VOF(i, j, k)],
                      normal + NC * I_NRM(i, j - 1, k), normal + NC * I_NRM(i, j, k));
}

#if NDIMS == 3
// z face k (1..nz+1)
__device__ inline double flux_z_at(const double * __restrict__ uz, const double * __restrict__ vof,
                                   const double * __restrict__ normal, int nx, int ny, int i, int j,
                                   int k) {
  constexpr int NC = NDIMS + 1;
  return face_flux<2>(uz[I_C1(i, j, k)], vof[I_VOF(i, j, k - 1)], vof[I_VOF(i, j, k)],
                      normal + NC * I_NRM(i, j, k - 1), normal + NC * I_NRM(i, j, k));
}
#endif

// thread (i, j, k) over 1..nx x 1..ny+1 x 1..nz+1 owns FLXX(i), FLXY(j),
// FLXZ(k) where those exist
__device__ __forceinline__ void k_flux_body(const VBlock & vb, Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                       const double * __restrict__ uz, const double * __restrict__ vof,
                       const double * __restrict__ normal, double * __restrict__ flxx,
                       double * __restrict__ flxy, double * __restrict__ flxz) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny + 1) return;
#if NDIMS == 3
  const int nz = g.nz;
  if (k > nz + 1) return;
  const bool kin = k <= nz;
#else
  const bool kin = true;
  (void)uz; (void)flxz;
#endif
  if (i >= 2 && j <= ny && kin) flxx[I_FX(i, j, k)] = flux_x_at(ux, vof, normal, nx, ny, i, j, k);
  if (kin) flxy[I_FY(i, j, k)] = flux_y_at(uy, vof, normal, nx, ny, i, j, k);
#if NDIMS == 3
  if (j <= ny) flxz[I_FZ(i, j, k)] = flux_z_at(uz, vof, normal, nx, ny, i, j, k);
#endif
}

__global__ void k_flux(Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                       const double * __restrict__ uz, const double * __restrict__ vof,
                       const double * __restrict__ normal, double * __restrict__ flxx,
                       double * __restrict__ flxy, double * __restrict__ flxz) {
  k_flux_body(hw_block(), g, ux, uy, uz, vof, normal, flxx, flxy, flxz);
}

// interface_compute_rhs (update/main.c:47-78) + interface_advect_vof
// (:80-110): src_a starts from zero (reset_srcs :28-45)
__device__ __forceinline__ void k_vof_update_body(const VBlock & vb, Grid g, const double * __restrict__ flxx,
                             const double * __restrict__ flxy, const double * __restrict__ flxz,
                             double * __restrict__ srca, const double * __restrict__ srcb,
                             double * __restrict__ vof, double dta, double dtb, int use_b) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
  double s = 0.;
  s += 1. / g.dxf[i - 1] * (flxx[I_FX(i, j, k)] - flxx[I_FX(i + 1, j, k)]);
  s += 1. / g.dy * (flxy[I_FY(i, j, k)] - flxy[I_FY(i, j + 1, k)]);
#if NDIMS == 3
  s += 1. / g.dz * (flxz[I_FZ(i, j, k)] - flxz[I_FZ(i, j, k + 1)]);
#else
  (void)flxz;
#endif
  const size_t is = I_SC(i, j, k);
  srca[is] = s;
  double v = vof[I_VOF(i, j, k)];
  v += dta * s;
  if (use_b) v += dtb * srcb[is];
  vof[I_VOF(i, j, k)] = v;
}

__global__ void k_vof_update(Grid g, const double * __restrict__ flxx,
                             const double * __restrict__ flxy, const double * __restrict__ flxz,
                             double * __restrict__ srca, const double * __restrict__ srcb,
                             double * __restrict__ vof, double dta, double dtb, int use_b) {
  k_vof_update_body(hw_block(), g, flxx, flxy, flxz, srca, srcb, vof, dta, dtb, use_b);
}

// ---------------------------------------------------------------------------
// fluid_predict_field: src/fluid/predict/{ux,uy,t}.c.  Each thread forms the
// explicit (srca) and implicit (srcg) source of one DOF term by term in the
// reference's accumulation order.

struct RhsArgs {
  const double * __restrict__ ux;
  const double * __restrict__ uy;
  const double * __restrict__ uz;
  const double * __restrict__ p;
  const double * __restrict__ t;
  const double * __restrict__ ifrc;
  const double * __restrict__ lapx;   // 3 per x point
  double * __restrict__ srca;
  double * __restrict__ srcg;          // implicit source (reference layout) or
  double * __restrict__ delta;         // ... the increment, when non-null
  const double * __restrict__ srcb;    // previous explicit source (delta only)
  double ca, cb, cg;                   // coef * dt (delta only)
  double dif;
  int add_buoyancy;
};

// store the explicit source and either the implicit source (the
// per-function entry point keeps the reference's fluid_t state) or directly
// the increment a dt srca + b dt srcb + g dt srcg (ux.c:335-344)
__device__ inline void rhs_store(const RhsArgs & a, size_t is, double sa, double sg) {
  a.srca[is] = sa;
  if (a.delta) {
    a.delta[is] = +a.ca * sa + a.cb * a.srcb[is] + a.cg * sg;
  } else {
    a.srcg[is] = sg;
  }
}

// Stencil values of one ux DOF (i, j, k): ux and its six neighbours, uy at
// (i-1, j), (i, j), (i-1, j+1), (i, j+1), uz at (i-1, j, k), (i, j, k),
// (i-1, j, k+1), (i, j, k+1), p and t at (i-1, j, k), (i, j, k).
struct UxSt {
  double c, xm, xp, ym, yp, zm, zp;
  double vlm, vlc, vum, vuc;
  double wlm, wlc, wum, wuc;
  double pm, pc, tm, tc;
};

// compute_rhs_ux (ux.c:71-249), DOFs i = 2..nx, from the stencil values
__device__ inline void rhs_ux_v(const Grid & g, const RhsArgs & a, int i, const UxSt & s,
                                double ifrc, double & sa_out, double & sg_out) {
  const double * xt = g.xt + kXT * i;
  double sa = 0., sg = 0.;
  {  // advection_x: +-0.5 / DXC(i) * ...
    const double ux_l = +0.5 * s.xm + 0.5 * s.c;
    const double ux_u = +0.5 * s.c + 0.5 * s.xp;
    const double l = xt[XT_HXC] * ux_l;
    const double u = -xt[XT_HXC] * ux_u;
    const double c = -l - u;
    sa += +l * s.xm + c * s.c + u * s.xp;
  }
  const double w_xm = xt[XT_WXM];   // 0.5 * DXF(i-1) / DXC(i)
  const double w_xp = xt[XT_WXP];   // 0.5 * DXF(i) / DXC(i)
  {  // advection_y
    const double uy_l = w_xm * s.vlm + w_xp * s.vlc;
    const double uy_u = w_xm * s.vum + w_xp * s.vuc;
    const double l = g.hdy * uy_l;
    const double u = -g.hdy * uy_u;
    const double c = -l - u;
    sa += +l * s.ym + c * s.c + u * s.yp;
  }
#if NDIMS == 3
  {  // advection_z
    const double uz_l = w_xm * s.wlm + w_xp * s.wlc;
    const double uz_u = w_xm * s.wum + w_xp * s.wuc;
    const double l = g.hdz * uz_l;
    const double u = -g.hdz * uz_u;
    const double c = -l - u;
    sa += +l * s.zm + c * s.c + u * s.zp;
  }
#endif
  {  // diffusion_x -> srcg
    const double * L = a.lapx + 3 * (i - 2);
    sg += a.dif * (L[0] * s.xm + L[1] * s.c + L[2] * s.xp);
  }
  {  // diffusion_y -> srca
    sa += a.dif * (g.lyy * s.ym + g.cyy * s.c + g.lyy * s.yp);
  }
#if NDIMS == 3
  {
    sa += a.dif * (g.lzz * s.zm + g.czz * s.c + g.lzz * s.zp);
  }
#endif
  sg -= xt[XT_RXC] * (-s.pm + s.pc);
  if (a.add_buoyancy) sa += +0.5 * s.tm + 0.5 * s.tc;
  sa += ifrc;
  sa_out = sa;
  sg_out = sg;
}

__device__ inline void rhs_ux_at(const Grid & g, const RhsArgs & a, int i, int j, int k,
                                 double ifrc, double & sa_out, double & sg_out) {
  const int nx = g.nx, ny = g.ny;
  (void)nx; (void)ny;
  UxSt s;
  s.c = a.ux[I_UX(i, j, k)];
  s.xm = a.ux[I_UX(i - 1, j, k)];
  s.xp = a.ux[I_UX(i + 1, j, k)];
  s.ym = a.ux[I_UX(i, j - 1, k)];
  s.yp = a.ux[I_UX(i, j + 1, k)];
  s.vlm = a.uy[I_C1(i - 1, j, k)];
  s.vlc = a.uy[I_C1(i, j, k)];
  s.vum = a.uy[I_C1(i - 1, j + 1, k)];
  s.vuc = a.uy[I_C1(i, j + 1, k)];
#if NDIMS == 3
  s.zm = a.ux[I_UX(i, j, k - 1)];
  s.zp = a.ux[I_UX(i, j, k + 1)];
  s.wlm = a.uz[I_C1(i - 1, j, k)];
  s.wlc = a.uz[I_C1(i, j, k)];
  s.wum = a.uz[I_C1(i - 1, j, k + 1)];
  s.wuc = a.uz[I_C1(i, j, k + 1)];
#else
  s.zm = s.zp = s.wlm = s.wlc = s.wum = s.wuc = 0.;
#endif
  s.pm = a.p[I_C1(i - 1, j, k)];
  s.pc = a.p[I_C1(i, j, k)];
  s.tm = a.add_buoyancy ? a.t[I_C1(i - 1, j, k)] : 0.;
  s.tc = a.add_buoyancy ? a.t[I_C1(i, j, k)] : 0.;
  rhs_ux_v(g, a, i, s, ifrc, sa_out, sg_out);
}

__device__ __forceinline__ void k_rhs_ux_body(const VBlock & vb, Grid g, RhsArgs a) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(2, 1, 1)
  if (i > nx || j > ny) return;
  double sa, sg;
  rhs_ux_at(g, a, i, j, k, a.ifrc[I_SX(i, j, k)], sa, sg);
  rhs_store(a, I_SX(i, j, k), sa, sg);
}

__global__ void k_rhs_ux(Grid g, RhsArgs a) {
  k_rhs_ux_body(hw_block(), g, a);
}

// Stencil values of one uy DOF (i, j, k): uy and its six neighbours, ux at
// (i, j-1), (i, j), (i+1, j-1), (i+1, j), uz at (i, j-1, k), (i, j, k),
// (i, j-1, k+1), (i, j, k+1), p at (i, j-1, k), (i, j, k).
struct UySt {
  double c, xm, xp, ym, yp, zm, zp;
  double ulm, ulc, uum, uuc;
  double wlm, wlc, wum, wuc;
  double pm, pc;
};

// compute_rhs_uy (uy.c:70-221), DOFs i = 1..nx, from the stencil values
__device__ inline void rhs_uy_v(const Grid & g, const RhsArgs & a, int i, const UySt & s,
                                double ifrc, double & sa_out, double & sg_out) {
  const double hxf = g.xt[kXT * i + XT_HXF];   // 0.5 / DXF(i)
  double sa = 0., sg = 0.;
  {  // advection_x
    const double ux_l = +0.5 * s.ulm + 0.5 * s.ulc;
    const double ux_u = +0.5 * s.uum + 0.5 * s.uuc;
    const double l = hxf * ux_l;
    const double u = -hxf * ux_u;
    const double c = -l - u;
    sa += +l * s.xm + c * s.c + u * s.xp;
  }
  {  // advection_y
    const double uy_l = +0.5 * s.ym + 0.5 * s.c;
    const double uy_u = +0.5 * s.c + 0.5 * s.yp;
    const double l = g.hdy * uy_l;
    const double u = -g.hdy * uy_u;
    const double c = -l - u;
    sa += +l * s.ym + c * s.c + u * s.yp;
  }
#if NDIMS == 3
  {  // advection_z
    const double uz_l = +0.5 * s.wlm + 0.5 * s.wlc;
    const double uz_u = +0.5 * s.wum + 0.5 * s.wuc;
    const double l = g.hdz * uz_l;
    const double u = -g.hdz * uz_u;
    const double c = -l - u;
    sa += +l * s.zm + c * s.c + u * s.zp;
  }
#endif
  {
    const double * L = a.lapx + 3 * (i - 1);
    sg += a.dif * (L[0] * s.xm + L[1] * s.c + L[2] * s.xp);
  }
  {
    sa += a.dif * (g.lyy * s.ym + g.cyy * s.c + g.lyy * s.yp);
  }
#if NDIMS == 3
  {
    sa += a.dif * (g.lzz * s.zm + g.czz * s.c + g.lzz * s.zp);
  }
#endif
  sg -= g.rdy * (-s.pm + s.pc);
  sa += ifrc;
  sa_out = sa;
  sg_out = sg;
}

__device__ inline void rhs_uy_at(const Grid & g, const RhsArgs & a, int i, int j, int k,
                                 double ifrc, double & sa_out, double & sg_out) {
  const int nx = g.nx, ny = g.ny;
  (void)nx; (void)ny;
  UySt s;
  s.c = a.uy[I_C1(i, j, k)];
  s.xm = a.uy[I_C1(i - 1, j, k)];
  s.xp = a.uy[I_C1(i + 1, j, k)];
  s.ym = a.uy[I_C1(i, j - 1, k)];
  s.yp = a.uy[I_C1(i, j + 1, k)];
  s.ulm = a.ux[I_UX(i, j - 1, k)];
  s.ulc = a.ux[I_UX(i, j, k)];
  s.uum = a.ux[I_UX(i + 1, j - 1, k)];
  s.uuc = a.ux[I_UX(i + 1, j, k)];
#if NDIMS == 3
  s.zm = a.uy[I_C1(i, j, k - 1)];
  s.zp = a.uy[I_C1(i, j, k + 1)];
  s.wlm = a.uz[I_C1(i, j - 1, k)];
  s.wlc = a.uz[I_C1(i, j, k)];
  s.wum = a.uz[I_C1(i, j - 1, k + 1)];
  s.wuc = a.uz[I_C1(i, j, k + 1)];
#else
  s.zm = s.zp = s.wlm = s.wlc = s.wum = s.wuc = 0.;
#endif
  s.pm = a.p[I_C1(i, j - 1, k)];
  s.pc = a.p[I_C1(i, j, k)];
  rhs_uy_v(g, a, i, s, ifrc, sa_out, sg_out);
}

__device__ __forceinline__ void k_rhs_uy_body(const VBlock & vb, Grid g, RhsArgs a) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
  double sa, sg;
  rhs_uy_at(g, a, i, j, k, a.ifrc[I_SC(i, j, k)], sa, sg);
  rhs_store(a, I_SC(i, j, k), sa, sg);
}

__global__ void k_rhs_uy(Grid g, RhsArgs a) {
  k_rhs_uy_body(hw_block(), g, a);
}

#if NDIMS == 3
// Stencil values of one uz DOF (i, j, k): uz and its six neighbours, ux at
// (i, j, k-1), (i, j, k), (i+1, j, k-1), (i+1, j, k), uy at (i, j, k-1),
// (i, j, k), (i, j+1, k-1), (i, j+1, k), p at (i, j, k-1), (i, j, k).
struct UzSt {
  double c, xm, xp, ym, yp, zm, zp;
  double ulm, ulc, uum, uuc;
  double vlm, vlc, vum, vuc;
  double pm, pc;
};

// compute_rhs_uz: the z counterpart of compute_rhs_uy, DOFs i = 1..nx
__device__ inline void rhs_uz_v(const Grid & g, const RhsArgs & a, int i, const UzSt & s,
                                double ifrc, double & sa_out, double & sg_out) {
  const double hxf = g.xt[kXT * i + XT_HXF];   // 0.5 / DXF(i)
  double sa = 0., sg = 0.;
  {
    const double ux_l = +0.5 * s.ulm + 0.5 * s.ulc;
    const double ux_u = +0.5 * s.uum + 0.5 * s.uuc;
    const double l = hxf * ux_l;
    const double u = -hxf * ux_u;
    const double c = -l - u;
    sa += +l * s.xm + c * s.c + u * s.xp;
  }
  {
    const double uy_l = +0.5 * s.vlm + 0.5 * s.vlc;
    const double uy_u = +0.5 * s.vum + 0.5 * s.vuc;
    const double l = g.hdy * uy_l;
    const double u = -g.hdy * uy_u;
    const double c = -l - u;
    sa += +l * s.ym + c * s.c + u * s.yp;
  }
  {
    const double uz_l = +0.5 * s.zm + 0.5 * s.c;
    const double uz_u = +0.5 * s.c + 0.5 * s.zp;
    const double l = g.hdz * uz_l;
    const double u = -g.hdz * uz_u;
    const double c = -l - u;
    sa += +l * s.zm + c * s.c + u * s.zp;
  }
  {
    const double * L = a.lapx + 3 * (i - 1);
    sg += a.dif * (L[0] * s.xm + L[1] * s.c + L[2] * s.xp);
  }
  {
    sa += a.dif * (g.lyy * s.ym + g.cyy * s.c + g.lyy * s.yp);
  }
  {
    sa += a.dif * (g.lzz * s.zm + g.czz * s.c + g.lzz * s.zp);
  }
  sg -= g.rdz * (-s.pm + s.pc);
  sa += ifrc;
  sa_out = sa;
  sg_out = sg;
}

__device__ inline void rhs_uz_at(const Grid & g, const RhsArgs & a, int i, int j, int k,
                                 double ifrc, double & sa_out, double & sg_out) {
  const int nx = g.nx, ny = g.ny;
  (void)nx; (void)ny;
  UzSt s;
  s.c = a.uz[I_C1(i, j, k)];
  s.xm = a.uz[I_C1(i - 1, j, k)];
  s.xp = a.uz[I_C1(i + 1, j, k)];
  s.ym = a.uz[I_C1(i, j - 1, k)];
  s.yp = a.uz[I_C1(i, j + 1, k)];
  s.zm = a.uz[I_C1(i, j, k - 1)];
  s.zp = a.uz[I_C1(i, j, k + 1)];
  s.ulm = a.ux[I_UX(i, j, k - 1)];
  s.ulc = a.ux[I_UX(i, j, k)];
  s.uum = a.ux[I_UX(i + 1, j, k - 1)];
  s.uuc = a.ux[I_UX(i + 1, j, k)];
  s.vlm = a.uy[I_C1(i, j, k - 1)];
  s.vlc = a.uy[I_C1(i, j, k)];
  s.vum = a.uy[I_C1(i, j + 1, k - 1)];
  s.vuc = a.uy[I_C1(i, j + 1, k)];
  s.pm = a.p[I_C1(i, j, k - 1)];
  s.pc = a.p[I_C1(i, j, k)];
  rhs_uz_v(g, a, i, s, ifrc, sa_out, sg_out);
}

__device__ __forceinline__ void k_rhs_uz_body(const VBlock & vb, Grid g, RhsArgs a) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
  double sa, sg;
  rhs_uz_at(g, a, i, j, k, a.ifrc[I_SC(i, j, k)], sa, sg);
  rhs_store(a, I_SC(i, j, k), sa, sg);
}

__global__ void k_rhs_uz(Grid g, RhsArgs a) {
  k_rhs_uz_body(hw_block(), g, a);
}
#endif

// Stencil values of one temperature DOF (i, j, k): t and its six
// neighbours, ux at (i, j, k), (i+1, j, k), uy at (i, j, k), (i, j+1, k), uz
// at (i, j, k), (i, j, k+1).
struct TSt {
  double c, xm, xp, ym, yp, zm, zp;
  double ul, uu, vl, vu, wl, wu;
};

// compute_rhs_t (t.c:68-182), DOFs i = 1..nx, from the stencil values
__device__ inline void rhs_t_v(const Grid & g, const RhsArgs & a, int i, const TSt & s,
                               double & sa_out, double & sg_out) {
  const double hxf = g.xt[kXT * i + XT_HXF];   // 0.5 / DXF(i)
  double sa = 0., sg = 0.;
  {
    const double l = hxf * s.ul;
    const double u = -hxf * s.uu;
    const double c = -l - u;
    sa += +l * s.xm + c * s.c + u * s.xp;
  }
  {
    const double l = g.hdy * s.vl;
    const double u = -g.hdy * s.vu;
    const double c = -l - u;
    sa += +l * s.ym + c * s.c + u * s.yp;
  }
#if NDIMS == 3
  {
    const double l = g.hdz * s.wl;
    const double u = -g.hdz * s.wu;
    const double c = -l - u;
    sa += +l * s.zm + c * s.c + u * s.zp;
  }
#endif
  {
    const double * L = a.lapx + 3 * (i - 1);
    sg += a.dif * (L[0] * s.xm + L[1] * s.c + L[2] * s.xp);
  }
  {
    sa += a.dif * (g.lyy * s.ym + g.cyy * s.c + g.lyy * s.yp);
  }
#if NDIMS == 3
  {
    sa += a.dif * (g.lzz * s.zm + g.czz * s.c + g.lzz * s.zp);
  }
#endif
  sa_out = sa;
  sg_out = sg;
}

__device__ inline void rhs_t_at(const Grid & g, const RhsArgs & a, int i, int j, int k,
                                double ifrc, double & sa_out, double & sg_out) {
  const int nx = g.nx, ny = g.ny;
  (void)nx; (void)ny;
  (void)ifrc;
  TSt s;
  s.c = a.t[I_C1(i, j, k)];
  s.xm = a.t[I_C1(i - 1, j, k)];
  s.xp = a.t[I_C1(i + 1, j, k)];
  s.ym = a.t[I_C1(i, j - 1, k)];
  s.yp = a.t[I_C1(i, j + 1, k)];
  s.ul = a.ux[I_UX(i, j, k)];
  s.uu = a.ux[I_UX(i + 1, j, k)];
  s.vl = a.uy[I_C1(i, j, k)];
  s.vu = a.uy[I_C1(i, j + 1, k)];
#if NDIMS == 3
  s.zm = a.t[I_C1(i, j, k - 1)];
  s.zp = a.t[I_C1(i, j, k + 1)];
  s.wl = a.uz[I_C1(i, j, k)];
  s.wu = a.uz[I_C1(i, j, k + 1)];
#else
  s.zm = s.zp = s.wl = s.wu = 0.;
#endif
  rhs_t_v(g, a, i, s, sa_out, sg_out);
}

__device__ __forceinline__ void k_rhs_t_body(const VBlock & vb, Grid g, RhsArgs a) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
  double sa, sg;
  rhs_t_at(g, a, i, j, k, 0., sa, sg);
  rhs_store(a, I_SC(i, j, k), sa, sg);
}

__global__ void k_rhs_t(Grid g, RhsArgs a) {
  k_rhs_t_body(hw_block(), g, a);
}

// ---------------------------------------------------------------------------
// increments (ux.c:335-344), implicit x solve of the predictor (tdm.c:17-50
// via solve_in_x, ux.c:251-272) and the field update (ux.c:376-386).
//
// The tri-diagonal matrix (I - prefactor Lx) is the same for every line, so
// the elimination factors the reference recomputes per line are prepared
// once on the host with the same arithmetic.
// Streaming store of a stage output: no other work-group of the writing
// kernel reads it, and a 256^3 array (134 MB) outlives the 4 MB L2 of an XCD
// before the next kernel reads it, so the line is not kept in L2, where it
// would evict the neighbour planes the stencil kernels re-read (k_rhs_march
// 704 -> 672 us at 256^3).  The value stored is the same.  Only for stores
// that fill whole 128-byte lines across a wavefront: lines written in parts
// by different wavefronts are merged in L2 by plain stores (the inverse y
// transform's psi columns: 97 us with plain stores, 121 us streaming).
__device__ __forceinline__ void st_stream(double * p, double v) { __builtin_nontemporal_store(v, p); }

template <bool NT>
__device__ __forceinline__ void st_maybe(double * p, double v) {
  if (NT)
    st_stream(p, v);
  else
    *p = v;
}

// Store v at padded position (x, row r, plane p) and at every periodic image
// of it inside the halo (h rows / planes on each side; r, p are storage
// indices of an interior entry).  In 2D the plane index is ignored.  NT:
// streaming stores (st_stream).
template <bool NT = true>
__device__ inline void store_periodic(double * __restrict__ a, int x, int n0, int n1, int r, int p,
                                      int ny, int nz, int h, double v) {
  int rs[3];
  int nr = 0;
  rs[nr++] = r;
  if (r < 2 * h) rs[nr++] = r + ny;
  if (r >= ny) rs[nr++] = r - ny;
#if NDIMS == 3
  int ps[3];
  int np = 0;
  ps[np++] = p;
  if (p < 2 * h) ps[np++] = p + nz;
  if (p >= nz) ps[np++] = p - nz;
  for (int b = 0; b < np; b++)
    for (int c = 0; c < nr; c++) st_maybe<NT>(&a[(size_t)x + (size_t)n0 * ((size_t)rs[c] + (size_t)n1 * (size_t)ps[b])], v);
#else
  (void)p; (void)nz; (void)n1;
  for (int c = 0; c < nr; c++) st_maybe<NT>(&a[(size_t)x + (size_t)n0 * (size_t)rs[c]], v);
#endif
}

// the images of (x, r, p) only (not the entry itself)
__device__ inline void store_images(double * __restrict__ a, int x, int n0, int n1, int r, int p,
                                    int ny, int nz, int h, double v) {
  int rs[3];
  int nr = 0;
  rs[nr++] = r;
  if (r < 2 * h) rs[nr++] = r + ny;
  if (r >= ny) rs[nr++] = r - ny;
#if NDIMS == 3
  int ps[3];
  int np = 0;
  ps[np++] = p;
  if (p < 2 * h) ps[np++] = p + nz;
  if (p >= nz) ps[np++] = p - nz;
  for (int b = 0; b < np; b++)
    for (int c = 0; c < nr; c++)
      if (b != 0 || c != 0) st_stream(&a[(size_t)x + (size_t)n0 * ((size_t)rs[c] + (size_t)n1 * (size_t)ps[b])], v);
#else
  (void)p; (void)nz; (void)n1;
  for (int c = 1; c < nr; c++) st_stream(&a[(size_t)x + (size_t)n0 * (size_t)rs[c]], v);
#endif
}

// Neumann wall update of a cell array at interior row (r, p): the halo
// images of the wall entries receive the wall values from before this update
// (the reference copies halos before it sets the walls), then the walls take
// the neighbouring interior values.
__device__ inline void neumann_walls(double * __restrict__ a, int i, int nx, int n1, int r, int p,
                                     int ny, int nz, double vnew) {
  const int n0 = nx + 2;
  if (i == 1) {
    const size_t w = (size_t)n0 * ((size_t)r + (size_t)n1 * (size_t)p);
    store_images(a, 0, n0, n1, r, p, ny, nz, 1, a[w]);
    a[w] = vnew;
  }
  if (i == nx) {
    const size_t w = (size_t)(nx + 1) + (size_t)n0 * ((size_t)r + (size_t)n1 * (size_t)p);
    store_images(a, nx + 1, n0, n1, r, p, ny, nz, 1, a[w]);
    a[w] = vnew;
  }
}

// Multi-field line solver: one launch solves the x lines of up to four
// fields (blockIdx.y selects the field).  256 lines per work-group, one line
// per lane; x is streamed through LDS in chunks of CH values stored x-major
// (tile[m][line]) so that the cooperative global accesses are contiguous
// rows and each lane's walk along its line is bank-conflict free.
struct LineField {
  const double * tab;        // {c0, rlast, flag, -, l[n], val[n], v[n]}
  const double * srca;       // increments from the three sources when srcg
  const double * srcb;       //   is non-null (per-function entry point) ...
  const double * srcg;
  double * q;                // ... else q holds them; forward result goes to q
  double * field;
  int n, n0, i0;             // line length, padded x width, first DOF column
};
struct LineJob {
  LineField f[4];
  int nlines, ny, n1;        // lines per field, rows per plane, padded rows (0 in 2D)
  double ca, cb, cg;
  int images, nz;            // also write periodic halo images (fused schedule)
};

constexpr int kLLines = 256;
constexpr int kLStride = 264;

template <int CH>
__device__ __forceinline__ void k_lines_body(const VBlock & vb, LineJob J) {
  __shared__ double T[CH][kLStride];
  __shared__ double U[CH][kLStride];
  const LineField & F = J.f[vb.y];
  const int tid = threadIdx.x;
  const int line0 = vb.x * kLLines;
  const int myline = line0 + tid;
  const int n = F.n;
  const int nlines = J.nlines;
  const double c0 = F.tab[0], rlast = F.tab[1];
  const bool flag = F.tab[2] != 0.;
  const double * __restrict__ L = F.tab + 4;
  const double * __restrict__ VAL = F.tab + 4 + n;
  const double * __restrict__ V = F.tab + 4 + 2 * n;
  double qprev = 0.;
  for (int c = 0; c < n; c += CH) {
    const int w = min(CH, n - c);
    for (int e = tid; e < kLLines * CH; e += kLLines) {
      const int ll = e / CH, m = e % CH;
      const int line = line0 + ll;
      if (m < w && line < nlines) {
        const size_t g = (size_t)line * n + c + m;
        T[m][ll] = F.srcg ? +J.ca * F.srca[g] + J.cb * F.srcb[g] + J.cg * F.srcg[g] : F.q[g];
      }
    }
    __syncthreads();
    if (myline < nlines) {
#pragma unroll
      for (int m = 0; m < CH; m++) {
        if (m < w) {
          const int ii = c + m;
          double qi = T[m][tid];
          if (ii == 0) {
            qi = qi / c0;
          } else if (ii < n - 1) {
            qi = VAL[ii] * (qi - L[ii] * qprev);
          } else {
            qi = flag ? rlast * (qi - L[ii] * qprev) : 0.;
          }
          T[m][tid] = qi;
          qprev = qi;
        }
      }
    }
    __syncthreads();
    for (int e = tid; e < kLLines * CH; e += kLLines) {
      const int ll = e / CH, m = e % CH;
      const int line = line0 + ll;
      if (m < w && line < nlines) F.q[(size_t)line * n + c + m] = T[m][ll];
    }
    __syncthreads();
  }
  double qnext = 0.;
  const int last_c = ((n - 1) / CH) * CH;
  for (int c = last_c; c >= 0; c -= CH) {
    const int w = min(CH, n - c);
    for (int e = tid; e < kLLines * CH; e += kLLines) {
      const int ll = e / CH, m = e % CH;
      const int line = line0 + ll;
      if (m < w && line < nlines) {
        T[m][ll] = F.q[(size_t)line * n + c + m];
        const int jj = line % J.ny, kk = line / J.ny;
        U[m][ll] = F.field[(size_t)(F.i0 + c + m) + (size_t)F.n0 * ((size_t)(jj + 1) + (size_t)J.n1 * (size_t)(kk + 1))];
      }
    }
    __syncthreads();
    if (myline < nlines) {
#pragma unroll
      for (int m = CH - 1; m >= 0; m--) {
        if (m < w) {
          const int ii = c + m;
          double qi = T[m][tid];
          if (ii < n - 1) qi -= V[ii] * qnext;
          qnext = qi;
          U[m][tid] += qi;
        }
      }
    }
    __syncthreads();
    for (int e = tid; e < kLLines * CH; e += kLLines) {
      const int ll = e / CH, m = e % CH;
      const int line = line0 + ll;
      if (m < w && line < nlines) {
        const int jj = line % J.ny, kk = line / J.ny;
        F.field[(size_t)(F.i0 + c + m) + (size_t)F.n0 * ((size_t)(jj + 1) + (size_t)J.n1 * (size_t)(kk + 1))] = U[m][ll];
      }
    }
    __syncthreads();
  }
}

template <int CH>
__global__ __launch_bounds__(kLLines) void k_lines(LineJob J) {
  k_lines_body<CH>(hw_block(), J);
}

// Whole-line variant: a work-group keeps `lpb` complete lines in LDS (odd
// row stride, so the lanes walking different lines hit different banks).
// Phase 1 copies the increments (one contiguous block: the lines of a group
// are adjacent) into LDS with every lane and several loads in flight,
// phase 2 runs forward elimination and back substitution on-chip with one
// lane per line, phase 3 adds the solution to the field rows.  Traffic per
// DOF: increments once, field read and write once.
constexpr int kWThreads = 256;
constexpr int kWUnroll = 4;

// Phases of the x-line solve of `nl` lines starting at line0, their values
// staged in S (row stride ld).  Phase 1: increments into S, by threads
// t = 0..nt-1; element e = t + nt r, (ll, m) tracked incrementally.
__device__ __forceinline__ void lines_fill(const LineField & F, const LineJob & J, double * S, int ld, int line0,
                                           int nl, int t, int nt) {
  const int n = F.n;
  const int total = nl * n;
  const size_t g0 = (size_t)line0 * n;
  {
    int e = t, ll = t / n, m = t - (t / n) * n;
    while (e < total) {
      double v[kWUnroll];
      int lls[kWUnroll], ms[kWUnroll], cnt = 0;
#pragma unroll
      for (int u = 0; u < kWUnroll; u++) {
        if (e < total) {
          const size_t g = g0 + e;
          v[u] = F.srcg ? +J.ca * F.srca[g] + J.cb * F.srcb[g] + J.cg * F.srcg[g] : F.q[g];
          lls[u] = ll;
          ms[u] = m;
          cnt = u + 1;
          e += nt;
          m += nt;
          while (m >= n) { m -= n; ll++; }
        }
      }
#pragma unroll
      for (int u = 0; u < kWUnroll; u++)
        if (u < cnt) S[lls[u] * ld + ms[u]] = v[u];
    }
  }
}

// phase 2: forward elimination and back substitution of one line q[0..n-1]
// (the reference's gtsv with the shared factors of the table)
__device__ __forceinline__ void lines_sweep(const LineField & F, double * q) {
  const int n = F.n;
  {
    const double c0 = F.tab[0], rlast = F.tab[1];
    const bool flag = F.tab[2] != 0.;
    const double * __restrict__ L = F.tab + 4;
    const double * __restrict__ VAL = F.tab + 4 + n;
    const double * __restrict__ V = F.tab + 4 + 2 * n;
    double qprev = q[0] / c0;
    q[0] = qprev;
    int i = 1;
    for (; i + 4 <= n - 1; i += 4) {
      double a0 = q[i], a1 = q[i + 1], a2 = q[i + 2], a3 = q[i + 3];
      a0 = VAL[i] * (a0 - L[i] * qprev);
      a1 = VAL[i + 1] * (a1 - L[i + 1] * a0);
      a2 = VAL[i + 2] * (a2 - L[i + 2] * a1);
      a3 = VAL[i + 3] * (a3 - L[i + 3] * a2);
      q[i] = a0; q[i + 1] = a1; q[i + 2] = a2; q[i + 3] = a3;
      qprev = a3;
    }
    for (; i < n - 1; i++) {
      const double a = VAL[i] * (q[i] - L[i] * qprev);
      q[i] = a;
      qprev = a;
    }
    {
      const double a = flag ? rlast * (q[n - 1] - L[n - 1] * qprev) : 0.;
      q[n - 1] = a;
      qprev = a;
    }
    double qnext = qprev;
    i = n - 2;
    for (; i - 3 >= 0; i -= 4) {
      double a0 = q[i], a1 = q[i - 1], a2 = q[i - 2], a3 = q[i - 3];
      a0 -= V[i] * qnext;
      a1 -= V[i - 1] * a0;
      a2 -= V[i - 2] * a1;
      a3 -= V[i - 3] * a2;
      q[i] = a0; q[i - 1] = a1; q[i - 2] = a2; q[i - 3] = a3;
      qnext = a3;
    }
    for (; i >= 0; i--) {
      const double a = q[i] - V[i] * qnext;
      q[i] = a;
      qnext = a;
    }
  }
}

// The same by a whole wavefront (the one-kernel predictor, whose waves have
// nothing else to do meanwhile): lane l < nl eliminates line l (rows at
// q0 + l ld); the other lanes follow line 0 without storing.  The factors
// L, VAL, V of 64 consecutive rows are loaded one row per lane (a coalesced
// vector load, issued one chunk ahead) and each row's pair is broadcast with
// v_readlane, so that neither scalar loads (whose out-of-order return makes
// each use wait for every outstanding LDS read) nor their latency sit on the
// recurrence; the right-hand sides of the next 8 rows are read from LDS
// before the current 8 rows' steps.  Per line the operations are gtsv's:
// q[0] / c0, then q[i] = VAL[i] (q[i] - L[i] q[i-1]), the last row
// rlast (q - L q) (0 when the last pivot is singular), back substitution
// q[i] -= V[i] q[i+1].
__device__ __forceinline__ double bcast(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ void lines_sweep_wave(const double * tab, int n, double * q0, int ld, int nl,
                                                 int lane) {
  constexpr int C = 8;
  const bool act = lane < nl;
  double * q = q0 + (size_t)(act ? lane : 0) * ld;
  const double c0 = tab[0], rlast = tab[1];
  const bool flag = tab[2] != 0.;
  const double * L = tab + 4;
  const double * VAL = tab + 4 + n;
  const double * V = tab + 4 + 2 * n;
  const int last = n - 1;
  double prev = q[0] / c0;
  if (act) q[0] = prev;
  // forward elimination, rows 1 .. last-1 in chunks of 64 (one row per lane)
  {
    double tl = L[min(1 + lane, last)], tv = VAL[min(1 + lane, last)];
    for (int c = 1; c < last; c += 64) {
      const int nxt = min(c + 64 + lane, last);
      const double nl_ = L[nxt], nv_ = VAL[nxt];
      const int m = min(64, last - c);
      int u = 0;
      for (; u + C <= m; u += C) {
        double x[C];
#pragma unroll
        for (int e = 0; e < C; e++) x[e] = q[c + u + e];
#pragma unroll
        for (int e = 0; e < C; e++) {
          prev = bcast(tv, u + e) * (x[e] - bcast(tl, u + e) * prev);
          if (act) q[c + u + e] = prev;
        }
      }
      for (; u < m; u++) {
        prev = bcast(tv, u) * (q[c + u] - bcast(tl, u) * prev);
        if (act) q[c + u] = prev;
      }
      tl = nl_;
      tv = nv_;
    }
  }
  prev = flag ? rlast * (q[last] - L[last] * prev) : 0.;
  if (act) q[last] = prev;
  // back substitution, rows last-1 .. 0 in descending chunks of 64
  {
    double tw = V[max(last - 1 - lane, 0)];
    for (int c = last - 1; c >= 0; c -= 64) {
      const int nxt = max(c - 64 - lane, 0);
      const double nw_ = V[nxt];
      const int m = min(64, c + 1);
      int u = 0;
      for (; u + C <= m; u += C) {
        double x[C];
#pragma unroll
        for (int e = 0; e < C; e++) x[e] = q[c - u - e];
#pragma unroll
        for (int e = 0; e < C; e++) {
          prev = x[e] - bcast(tw, u + e) * prev;
          if (act) q[c - u - e] = prev;
        }
      }
      for (; u < m; u++) {
        prev = q[c - u] - bcast(tw, u) * prev;
        if (act) q[c - u] = prev;
      }
      tw = nw_;
    }
  }
}

// phase 3: field rows += solution by wavefronts w = 0..nw-1 (lane = lane
// index).  Wavefront w owns lines w, w + nw, ... (two per pass), so a
// line's row / plane, its address and whether it has periodic halo images
// are wave-uniform; lanes walk x contiguously.
__device__ __forceinline__ void lines_update(const LineField & F, const LineJob & J, const double * S, int ld,
                                             int line0, int nl, int w, int nw, int lane) {
  const int n = F.n;
  {
    for (int la = w; la < nl; la += 2 * nw) {
      const int lb = la + nw;
      const bool hb = lb < nl;
      const int linea = line0 + la, lineb = line0 + (hb ? lb : la);
      const int ja = linea % J.ny, ka = linea / J.ny;
      const int jb = lineb % J.ny, kb = lineb / J.ny;
      double * __restrict__ ra = F.field + F.i0 + (size_t)F.n0 * ((size_t)(ja + 1) + (size_t)J.n1 * (size_t)(ka + 1));
      double * __restrict__ rb = F.field + F.i0 + (size_t)F.n0 * ((size_t)(jb + 1) + (size_t)J.n1 * (size_t)(kb + 1));
      const double * sa = S + la * ld;
      const double * sb = S + (hb ? lb : la) * ld;
      for (int m0 = 0; m0 < n; m0 += 64 * kWUnroll) {
        double va[kWUnroll], vb[kWUnroll];
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
          const int m = m0 + 64 * u + lane;
          if (m < n) {
            va[u] = ra[m] + sa[m];
            if (hb) vb[u] = rb[m] + sb[m];
          }
        }
#pragma unroll
        for (int u = 0; u < kWUnroll; u++) {
          const int m = m0 + 64 * u + lane;
          if (m < n) {
            st_stream(&ra[m], va[u]);
            if (hb) st_stream(&rb[m], vb[u]);
          }
        }
      }
      if (J.images) {
        // rows / planes next to the periodic halo: the entries just stored
        // go to their images too, and the Dirichlet wall entries of the row
        // (constants the loop never changes) take their images as the
        // reference's halo copy would (overwriting values left from the
        // initial load)
#pragma unroll
        for (int h = 0; h < 2; h++) {
          if (h == 1 && !hb) break;
          const int jj = h ? jb : ja, kk = h ? kb : ka;
          const bool edge = jj + 1 < 2 || jj + 1 >= J.ny
#if NDIMS == 3
                            || kk + 1 < 2 || kk + 1 >= J.nz
#endif
              ;
          if (!edge) continue;
          const double * row = h ? rb : ra;
          for (int m = lane; m < n; m += 64)
            store_images(F.field, F.i0 + m, F.n0, J.n1, jj + 1, kk + 1, J.ny, J.nz, 1, row[m]);
          if (lane < 2) {
            const int xw = lane == 0 ? F.i0 - 1 : F.i0 + n;
            const double wv = F.field[(size_t)xw + (size_t)F.n0 * ((size_t)(jj + 1) + (size_t)J.n1 * (size_t)(kk + 1))];
            store_images(F.field, xw, F.n0, J.n1, jj + 1, kk + 1, J.ny, J.nz, 1, wv);
          }
        }
      }
    }
  }
}

__device__ __forceinline__ void k_lines_lds_body(const VBlock & vb, const LineJob & J, int lpb) {
  extern __shared__ double S[];
  const LineField & F = J.f[vb.y];
  const int ld = F.n | 1;
  const int tid = threadIdx.x;
  const int line0 = vb.x * lpb;
  const int nl = min(lpb, J.nlines - line0);
  if (nl <= 0) return;
  lines_fill(F, J, S, ld, line0, nl, tid, kWThreads);
  __syncthreads();
  if (tid < nl) lines_sweep(F, S + tid * ld);
  __syncthreads();
  lines_update(F, J, S, ld, line0, nl, tid >> 6, kWThreads / 64, tid & 63);
}

__global__ __launch_bounds__(kWThreads) void k_lines_lds(LineJob J, int lpb) {
  k_lines_lds_body(hw_block(), J, lpb);
}

// ---------------------------------------------------------------------------
// Poisson solver: FFT in y (and z), Thomas in x (compute_potential/dft.c).
//
// Spectral layout spec[x][kz][ky] (complex, ky fastest, nh = ny/2+1): the
// y transform writes contiguous ky rows per x column, the z transform reads
// contiguous ky runs per kz, and the x solve walks x with one lane per
// (ky, kz) line so that a wavefront touches consecutive addresses.  The
// reference's pencil transposes (SimpleDecomp) are folded into the LDS
// staging of these kernels.

struct FftPlan {
  int n;
  int nst;
  int rad[16];
};

// complex element types of the transforms: FP64 (double2) for the plain
// solver, FP32 (float2) for the mixed-precision solver
template <class T2> struct Cx;
template <> struct Cx<double2> {
  using S = double;
  __device__ static double2 mk(double a, double b) { return make_double2(a, b); }
};
template <> struct Cx<float2> {
  using S = float;
  __device__ static float2 mk(float a, float b) { return make_float2(a, b); }
};

// radix-R DFT of v in place: y_k = sum_j v_j exp(sign 2 pi i j k / R)
template <int R, class T2>
__device__ inline void butterfly(T2 * v, typename Cx<T2>::S sign) {
  using C = Cx<T2>;
  using S = typename C::S;
  if constexpr (R == 2) {
    const T2 a = v[0], b = v[1];
    v[0] = C::mk(a.x + b.x, a.y + b.y);
    v[1] = C::mk(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const T2 s02 = C::mk(v[0].x + v[2].x, v[0].y + v[2].y);
    const T2 d02 = C::mk(v[0].x - v[2].x, v[0].y - v[2].y);
    const T2 s13 = C::mk(v[1].x + v[3].x, v[1].y + v[3].y);
    const T2 d13 = C::mk(v[1].x - v[3].x, v[1].y - v[3].y);
    const T2 jd = C::mk(-sign * d13.y, sign * d13.x);   // (sign i) d13
    v[0] = C::mk(s02.x + s13.x, s02.y + s13.y);
    v[2] = C::mk(s02.x - s13.x, s02.y - s13.y);
    v[1] = C::mk(d02.x + jd.x, d02.y + jd.y);
    v[3] = C::mk(d02.x - jd.x, d02.y - jd.y);
  } else if constexpr (R == 3) {
    const S c3 = S(-0.5), s3 = sign * S(0.86602540378443864676);
    const T2 t1 = C::mk(v[1].x + v[2].x, v[1].y + v[2].y);
    const T2 t2 = C::mk(v[1].x - v[2].x, v[1].y - v[2].y);
    const T2 m = C::mk(v[0].x + c3 * t1.x, v[0].y + c3 * t1.y);
    const T2 r2 = C::mk(-s3 * t2.y, s3 * t2.x);         // s3 i t2
    v[0] = C::mk(v[0].x + t1.x, v[0].y + t1.y);
    v[1] = C::mk(m.x + r2.x, m.y + r2.y);
    v[2] = C::mk(m.x - r2.x, m.y - r2.y);
  } else {
    static_assert(R == 5, "radix");
    const S c1 = S(0.30901699437494742410), c2 = S(-0.80901699437494742410);
    const S s1 = sign * S(0.95105651629515357212), s2 = sign * S(0.58778525229247312917);
    const T2 a1 = C::mk(v[1].x + v[4].x, v[1].y + v[4].y);
    const T2 b1 = C::mk(v[1].x - v[4].x, v[1].y - v[4].y);
    const T2 a2 = C::mk(v[2].x + v[3].x, v[2].y + v[3].y);
    const T2 b2 = C::mk(v[2].x - v[3].x, v[2].y - v[3].y);
    const T2 x0 = v[0];
    v[0] = C::mk(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    const T2 m1 = C::mk(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const T2 m2 = C::mk(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    const T2 n1 = C::mk(-(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x);
    const T2 n2 = C::mk(-(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x);
    v[1] = C::mk(m1.x + n1.x, m1.y + n1.y);
    v[4] = C::mk(m1.x - n1.x, m1.y - n1.y);
    v[2] = C::mk(m2.x + n2.x, m2.y + n2.y);
    v[3] = C::mk(m2.x - n2.x, m2.y - n2.y);
  }
}

// one Stockham stage: combine R sub-transforms of length Ns read at stride
// N/R (Govindaraju et al., SC'08 formulation)
template <int R, class T2>
__device__ inline void fft_stage(const T2 * a, T2 * b, int ncols, int N, int Ns,
                                 const T2 * __restrict__ tw, typename Cx<T2>::S sign) {
  using C = Cx<T2>;
  using S = typename C::S;
  const int nb = N / R;
  const int tstride = N / (Ns * R);
  for (int t = threadIdx.x; t < ncols * nb; t += blockDim.x) {
    const int col = t / nb, jj = t - col * nb;
    const T2 * in = a + (size_t)col * N;
    T2 * out = b + (size_t)col * N;
    const int kk = jj % Ns;
    T2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const T2 x = in[jj + r * nb];
      if (r == 0 || kk == 0) {
        v[r] = x;
      } else {
        const T2 w = tw[kk * r * tstride];
        const S ws = sign * w.y;
        v[r] = C::mk(x.x * w.x - x.y * ws, x.x * ws + x.y * w.x);
      }
    }
    butterfly<R, T2>(v, sign);
    const int idxD = (jj / Ns) * Ns * R + kk;
#pragma unroll
    for (int r = 0; r < R; r++) out[idxD + r * Ns] = v[r];
  }
}

// Stockham autosort FFT of `ncols` LDS columns of length P.n between buffers
// a and b; returns the buffer holding the result.
// out[k] = sum_j in[j] exp(sign 2 pi i j k / n); tw[m] = (cos, sin)(2 pi m / n)
template <class T2>
__device__ T2 * lds_fft(T2 * a, T2 * b, int ncols, const FftPlan & P, const T2 * __restrict__ tw,
                        typename Cx<T2>::S sign) {
  const int N = P.n;
  int Ns = 1;
  for (int s = 0; s < P.nst; s++) {
    const int R = P.rad[s];
    switch (R) {
      case 4: fft_stage<4, T2>(a, b, ncols, N, Ns, tw, sign); break;
      case 2: fft_stage<2, T2>(a, b, ncols, N, Ns, tw, sign); break;
      case 3: fft_stage<3, T2>(a, b, ncols, N, Ns, tw, sign); break;
      default: fft_stage<5, T2>(a, b, ncols, N, Ns, tw, sign); break;
    }
    __syncthreads();
    T2 * tmp = a; a = b; b = tmp;
    Ns *= R;
  }
  return a;
}

constexpr int kFftThreads = 256;
constexpr int kFftLoadU = 4;      // elements per thread whose global loads are batched

// 16-point DFT in registers, radix 4 x 4, in place: X[k] = sum_n v[n] w^(n k),
// w = exp(sign 2 pi i / 16), n = a + 4 b, k = kb + 4 ka; X[k] is left in
// v[dft16_slot(k)] = v[4 kb + ka]
__device__ inline int dft16_slot(int k) { return 4 * (k & 3) + (k >> 2); }
template <class T2>
__device__ __forceinline__ void dft16(T2 * v, typename Cx<T2>::S sign) {
  using S = typename Cx<T2>::S;
  // cos, sin of 2 pi m / 16, m = 0..15
  constexpr double kC[16] = {1., 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                             0., -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                             -1., -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                             0., 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613};
  constexpr double kS[16] = {0., 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                             1., 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                             0., -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                             -1., -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173};
  // DFT4 over b of v[a + 4 b], twiddled by w^(a kb), back into v[a + 4 kb]
#pragma unroll
  for (int a = 0; a < 4; a++) {
    T2 u[4] = {v[a], v[a + 4], v[a + 8], v[a + 12]};
    butterfly<4, T2>(u, sign);
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const int m = a * kb;
      if (m == 0) {
        v[a + 4 * kb] = u[kb];
      } else {
        const S wr = (S)kC[m], wi = sign * (S)kS[m];
        v[a + 4 * kb] = Cx<T2>::mk(u[kb].x * wr - u[kb].y * wi, u[kb].x * wi + u[kb].y * wr);
      }
    }
  }
  // DFT4 over a of v[a + 4 kb] -> X[kb + 4 ka], left in v[4 kb + ka]
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    T2 u[4] = {v[4 * kb], v[4 * kb + 1], v[4 * kb + 2], v[4 * kb + 3]};
    butterfly<4, T2>(u, sign);
#pragma unroll
    for (int ka = 0; ka < 4; ka++) v[4 * kb + ka] = u[ka];
  }
}

// 256-point transforms of `ncols` LDS columns in place, as two passes of
// 16-point DFTs in registers (n = n2 + 16 n1, k = k1 + 16 k2):
//   pass 1, thread (c, n2): Y[n2][k1] = W^(n2 k1) DFT16_n1(x[n2 + 16 n1])[k1]
//   pass 2, thread (c, k1): X[k1 + 16 k2] = DFT16_n2(Y[n2][k1])[k2]
// with W = exp(sign 2 pi i / 256), tw[m] = (cos, sin)(2 pi m / 256).  A
// column is a 16 x 17 matrix (row stride 17: both passes are bank-conflict
// free); x[n] sits at row n / 16, column n % 16 (f256_in), Y[n2][k1] at
// (k1, n2) and X[k] at (k % 16, k / 16) (f256_out).  Each thread owns one
// matrix column in pass 1 and one row in pass 2, so both run in place.
constexpr int kF256Cols = 8;      // transforms per work-group of the y kernels (16 x columns = 128-byte rows)
constexpr int kF256ColsZ = 16;    // transforms per work-group of the z kernel (every thread of the FFT passes busy)
constexpr int kF256Stride = 273;  // complex values per column: 16 x 17 plus one, so that the
                                  // columns of a work-group start in different LDS banks
__device__ inline int f256_in(int n) { return (n >> 4) * 17 + (n & 15); }
__device__ inline int f256_out(int k) { return (k & 15) * 17 + (k >> 4); }

template <class T2>
__device__ __forceinline__ void fft256(T2 * buf, int ncols, const T2 * __restrict__ tw, typename Cx<T2>::S sign) {
  using S = typename Cx<T2>::S;
  for (int t = threadIdx.x; t < ncols * 16; t += blockDim.x) {
    T2 * col = buf + (t >> 4) * kF256Stride;
    const int n2 = t & 15;
    T2 v[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; n1++) v[n1] = col[n1 * 17 + n2];
    dft16<T2>(v, sign);
#pragma unroll
    for (int k1 = 0; k1 < 16; k1++) {
      T2 x = v[dft16_slot(k1)];
      if (k1 != 0 && n2 != 0) {
        const T2 w = tw[n2 * k1];
        const S ws = sign * w.y;
        x = Cx<T2>::mk(x.x * w.x - x.y * ws, x.x * ws + x.y * w.x);
      }
      col[k1 * 17 + n2] = x;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < ncols * 16; t += blockDim.x) {
    T2 * row = buf + (t >> 4) * kF256Stride + (t & 15) * 17;
    T2 v[16];
#pragma unroll
    for (int n2 = 0; n2 < 16; n2++) v[n2] = row[n2];
    dft16<T2>(v, sign);
#pragma unroll
    for (int k2 = 0; k2 < 16; k2++) row[k2] = v[dft16_slot(k2)];
  }
  __syncthreads();
}

// column layout of the two transform engines: F256 = the 256-point register
// path (one padded buffer), else the mixed-radix Stockham path (two buffers)
template <bool F256> __device__ inline int fft_in(int n) { return F256 ? f256_in(n) : n; }
template <bool F256> __device__ inline int fft_out(int k) { return F256 ? f256_out(k) : k; }
template <bool F256, class T2>
__device__ __forceinline__ T2 * fft_cols(T2 * a, T2 * b, int ncols, const FftPlan & P,
                                     const T2 * __restrict__ tw, typename Cx<T2>::S sign) {
  if constexpr (F256) {
    (void)b; (void)P;
    fft256<T2>(a, ncols, tw, sign);
    return a;
  } else {
    return lds_fft<T2>(a, b, ncols, P, tw, sign);
  }
}

constexpr int kFftElems = 2048;   // complex elements per LDS buffer

// assign_input (dft.c:315-345) fused with the real transform in y.  Two x
// columns are packed into one complex column; the r2c halves are separated
// after the transform.  Output spec[x][k][ky].
template <int ELEMS, bool F256 = false>
__device__ __forceinline__ void k_div_rfft_y_body(const VBlock & vb, 
    Grid g, const FftPlan & P, const double2 * __restrict__ tw, const double * __restrict__ ux,
    const double * __restrict__ uy, const double * __restrict__ uz, double prefactor,
    double2 * __restrict__ spec) {
  __shared__ double2 sa[F256 ? kF256Cols * kF256Stride : ELEMS];
  __shared__ double2 sb[F256 ? 1 : ELEMS];
  const int nx = g.nx, ny = g.ny;
  const int N = ny;
  const int NC = F256 ? kF256Cols : ELEMS / N;   // complex columns per block
  const int CS = F256 ? kF256Stride : N;         // LDS column stride
  const Tile3 tile = xcd_tile(vb);
  const int x0 = tile.x * 2 * NC;          // first interior x (0-based)
#if NDIMS == 2
  const int k = 1;
  const int nz = 1;
  (void)uz;
#else
  const int k = (int)tile.y + 1;
  const int nz = g.nz;
#endif
  // kFftLoadU elements per thread have their loads in flight together
  const int total = 2 * NC * N;
  for (int e0 = threadIdx.x; e0 < total; e0 += kFftLoadU * blockDim.x) {
    // Every load of an element is issued in this loop, unconditionally
    // (indices past the grid are clamped to it and their values unused), so
    // all of them are in flight together: a load in the second loop, or
    // loads under a divergent branch, made each element wait for the others.
    double xm[kFftLoadU], xp[kFftLoadU], ym[kFftLoadU], yp[kFftLoadU], zm[kFftLoadU], zp[kFftLoadU];
    double dxv[kFftLoadU];
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      const int xl = e % (2 * NC);
      const int j = min(e / (2 * NC) + 1, ny);
      const int i = min(x0 + xl + 1, nx);
      dxv[u] = g.dxf[i - 1];
      xm[u] = ux[I_UX(i, j, k)];
      xp[u] = ux[I_UX(i + 1, j, k)];
      ym[u] = uy[I_C1(i, j, k)];
      yp[u] = uy[I_C1(i, j + 1, k)];
#if NDIMS == 3
      zm[u] = uz[I_C1(i, j, k)];
      zp[u] = uz[I_C1(i, j, k + 1)];
#else
      zm[u] = zp[u] = 0.;
#endif
    }
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      if (e >= total) break;
      const int xl = e % (2 * NC), j = e / (2 * NC) + 1;
      const int i = x0 + xl + 1;
      double r = 0.;
      if (i <= nx) {
        const double dx = dxv[u];
#if NDIMS == 2
        r = prefactor * ((xp[u] - xm[u]) / dx + (yp[u] - ym[u]) / g.dy);
#else
        r = prefactor * ((xp[u] - xm[u]) / dx + (yp[u] - ym[u]) / g.dy + (zp[u] - zm[u]) / g.dz);
#endif
      }
      double * slot = reinterpret_cast<double *>(&sa[(xl >> 1) * CS + fft_in<F256>(j - 1)]);
      slot[xl & 1] = r;
    }
  }
  __syncthreads();
  double2 * res = fft_cols<F256, double2>(sa, sb, NC, P, tw, -1.);
  const int nh = N / 2 + 1;
  for (int e = threadIdx.x; e < NC * nh; e += blockDim.x) {
    const int c = e / nh, ky = e % nh;
    const double2 z = res[c * CS + fft_out<F256>(ky)];
    const double2 zc = res[c * CS + fft_out<F256>((N - ky) % N)];
    // A = (Z[k] + conj Z[N-k]) / 2, B = (Z[k] - conj Z[N-k]) / (2 i)
    const double2 A = make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y - zc.y));
    const double2 B = make_double2(0.5 * (z.y + zc.y), -0.5 * (z.x - zc.x));
    const int xa = x0 + 2 * c, xb = xa + 1;
    if (xa < nx) spec[((size_t)xa * nz + (k - 1)) * nh + ky] = A;
    if (xb < nx) spec[((size_t)xb * nz + (k - 1)) * nh + ky] = B;
  }
}

template <int ELEMS, bool F256 = false>
__global__ __launch_bounds__(kFftThreads) void k_div_rfft_y(
    Grid g, FftPlan P, const double2 * __restrict__ tw, const double * __restrict__ ux,
    const double * __restrict__ uy, const double * __restrict__ uz, double prefactor,
    double2 * __restrict__ spec) {
  k_div_rfft_y_body<ELEMS, F256>(hw_block(), g, P, tw, ux, uy, uz, prefactor, spec);
}

#if NDIMS == 3
// complex transform in z of spec[x][kz][ky] for a run of ky at one x
template <int ELEMS, class T2, bool F256 = false>
__device__ __forceinline__ void k_fft_z_body(const VBlock & vb, Grid g, const FftPlan & P, const T2 * __restrict__ tw,
                                                      T2 * __restrict__ spec, int nh, typename Cx<T2>::S sign) {
  __shared__ T2 sa[F256 ? kF256ColsZ * kF256Stride : ELEMS];
  __shared__ T2 sb[F256 ? 1 : ELEMS];
  const int N = g.nz;
  const int NC = F256 ? kF256ColsZ : ELEMS / N;
  const int CS = F256 ? kF256Stride : N;
  const Tile3 tile = xcd_tile(vb);
  const int ky0 = tile.x * NC;
  const int x = tile.y;
  T2 * base = spec + (size_t)x * N * nh;
  const int total = NC * N;
  for (int e0 = threadIdx.x; e0 < total; e0 += kFftLoadU * blockDim.x) {
    T2 v[kFftLoadU];
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      const int c = e % NC, kz = e / NC;
      const int ky = ky0 + c;
      v[u] = e < total && ky < nh ? base[(size_t)kz * nh + ky] : Cx<T2>::mk(0, 0);
    }
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      if (e < total) sa[(e % NC) * CS + fft_in<F256>(e / NC)] = v[u];
    }
  }
  __syncthreads();
  T2 * res;
  if constexpr (F256) {
    using S = typename Cx<T2>::S;
    // the direction as a constant, so the twiddle signs fold
    if (sign < 0)
      res = fft_cols<true, T2>(sa, sb, NC, P, tw, (S)-1);
    else
      res = fft_cols<true, T2>(sa, sb, NC, P, tw, (S)+1);
  } else {
    res = lds_fft<T2>(sa, sb, NC, P, tw, sign);
  }
  for (int e = threadIdx.x; e < NC * N; e += blockDim.x) {
    const int c = e % NC, kz = e / NC;
    const int ky = ky0 + c;
    if (ky < nh) base[(size_t)kz * nh + ky] = res[c * CS + fft_out<F256>(kz)];
  }
}

template <int ELEMS, class T2, bool F256 = false>
__global__ __launch_bounds__(kFftThreads) void k_fft_z(Grid g, FftPlan P, const T2 * __restrict__ tw,
                                                      T2 * __restrict__ spec, int nh, typename Cx<T2>::S sign) {
  k_fft_z_body<ELEMS, T2, F256>(hw_block(), g, P, tw, spec, nh, sign);
}
#endif

// solve_linear_systems (dft.c:364-395): per (ky, kz) line the complex
// tri-diagonal system in x with c = -l - u + lambda and the Neumann fold,
// eliminated as gtsv (tdm.c:17-50).  v goes to vbuf[x][line] when store_v
// (the factors depend only on the grid and the wavenumbers, so they are
// written by the first solve and read back by every later one).  One
// lane per line; the rows of the next chunk of U rows are loaded before the
// current chunk is eliminated, so U loads per lane stay in flight.
constexpr int kTsU = 8;   // default rows per chunk (the U of the templates below)

__device__ inline double2 to_d2(double2 a) { return a; }
__device__ inline double2 to_d2(float2 a) { return make_double2((double)a.x, (double)a.y); }
template <class T2> __device__ inline T2 from_d2(double2 a);
template <> __device__ inline double2 from_d2<double2>(double2 a) { return a; }
template <> __device__ inline float2 from_d2<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }

// T2 is the storage type of the spectrum (double2, or float2 in the
// mixed-precision solver); the elimination itself is FP64 either way
// Two lanes per spectral line: the real and imaginary parts are two
// systems with the same (real) coefficients (the reference's real-by-complex
// arithmetic), so lane 2 line + c sweeps component c of the line.  Both
// lanes form the same factors and both store them (store_v: the same value
// to the same address, so each lane's back substitution reads its own
// store).  At 256^3
// that is 66048 lanes, one wavefront for every SIMD of the chip, where one
// lane per line left half of them idle.  Per component the operations are
// those of the one-lane-per-line elimination, in the same order.
template <class T2, int U = kTsU>
__device__ __forceinline__ void k_thomas_spec_body(const VBlock & vb, int nx, int nlines, int nh, const double * __restrict__ pl,
                              const double * __restrict__ pu, const double * __restrict__ ey,
                              const double * __restrict__ ez, T2 * __restrict__ spec,
                              double * __restrict__ vbuf, int store_v) {
  using S = typename Cx<T2>::S;   // component type of the spectrum
  const int lane = (int)(vb.x * blockDim.x + threadIdx.x);
  const int line = lane >> 1, comp = lane & 1;
  if (line >= nlines) return;
  const int ky = line % nh;
#if NDIMS == 2
  const double eval = ey[ky];
  (void)ez;
#else
  const int kz = line / nh;
  const double eval = ey[ky] + ez[kz];
#endif
  const size_t st = 2 * (size_t)nlines;   // components between rows
  S * q = reinterpret_cast<S *>(spec) + 2 * (size_t)line + comp;
  double * v = vbuf + line;
  const size_t sv = (size_t)nlines;
  const bool sv0 = store_v != 0;
  const int n = nx;
  // row 0 (with the Neumann fold c[0] += l[0])
  double cc = -pl[0] - pu[0] + eval;
  cc += pl[0];
  double vprev = pu[0] / cc;
  double qprev = (double)q[0];
  qprev = qprev / cc;
  q[0] = (S)qprev;
  if (sv0) v[0] = vprev;
  {
    // rows are kept in the storage type until used: a conversion right
    // after a conditional load would wait for that load and serialise them
    S cur[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (1 + u < n - 1) cur[u] = q[(size_t)(1 + u) * st];
    }
    for (int i0 = 1; i0 < n - 1; i0 += U) {
      S nxt[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int i = i0 + U + u;
        if (i < n - 1) nxt[u] = q[(size_t)i * st];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int i = i0 + u;
        if (i < n - 1) {
          const double c = -pl[i] - pu[i] + eval;
          const double val = 1. / (c - pl[i] * vprev);
          const double vi = val * (pu[i]);
          double qi = (double)cur[u];
          qi = val * (qi - pl[i] * qprev);
          q[(size_t)i * st] = (S)qi;
          if (sv0) v[(size_t)i * sv] = vi;
          qprev = qi;
          vprev = vi;
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) cur[u] = nxt[u];
    }
  }
  {
    const int i = n - 1;
    double c = -pl[i] - pu[i] + eval;
    c += pu[i];
    const double val = c - pl[i] * vprev;
    double qi = (double)q[(size_t)i * st];
    // the (ky, kz) = (0, 0) line is singular (Neumann in x).  Its right-hand
    // side is compatible to FP64 rounding, and then the guard below picks the
    // gauge; an FP32 spectrum is compatible only to FP32 rounding, which
    // 1 / val would blow up, so that line always takes the gauge branch
    const bool gauge = sizeof(T2) == sizeof(float2) && eval == 0.;
    if (!gauge && fabs(val) > DBL_EPSILON) {
      const double r = 1. / val;
      qi = r * (qi - pl[i] * qprev);
    } else {
      qi = 0.;
    }
    q[(size_t)i * st] = (S)qi;
    qprev = qi;
  }
  {
    S cq[U];
    double cv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = n - 2 - u;
      if (i >= 0) {
        cq[u] = q[(size_t)i * st];
        cv[u] = v[(size_t)i * sv];
      }
    }
    for (int i0 = n - 2; i0 >= 0; i0 -= U) {
      S nq[U];
      double nv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int i = i0 - U - u;
        if (i >= 0) {
          nq[u] = q[(size_t)i * st];
          nv[u] = v[(size_t)i * sv];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int i = i0 - u;
        if (i >= 0) {
          double qi = (double)cq[u];
          qi -= cv[u] * qprev;
          q[(size_t)i * st] = (S)qi;
          qprev = qi;
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        cq[u] = nq[u];
        cv[u] = nv[u];
      }
    }
  }
}

template <class T2, int U = kTsU>
__global__ __launch_bounds__(64) void k_thomas_spec(int nx, int nlines, int nh, const double * __restrict__ pl,
                              const double * __restrict__ pu, const double * __restrict__ ey,
                              const double * __restrict__ ez, T2 * __restrict__ spec,
                              double * __restrict__ vbuf, int store_v) {
  k_thomas_spec_body<T2, U>(hw_block(), nx, nlines, nh, pl, pu, ey, ez, spec, vbuf, store_v);
}

// The elimination of k_thomas_spec for one whole line, both components by
// one lane (per component the operations and their order are those of
// k_thomas_spec with store_v; the gauge choice of its last row included).
template <class T2>
__device__ void thomas_line_seq(int n, int nlines, int line, double eval, const double * __restrict__ pl,
                                const double * __restrict__ pu, T2 * __restrict__ spec, double * __restrict__ vbuf) {
  T2 * q = spec + line;
  double * v = vbuf + line;
  const size_t st = (size_t)nlines;
  double cc = -pl[0] - pu[0] + eval;
  cc += pl[0];
  double vprev = pu[0] / cc;
  double2 qprev = to_d2(q[0]);
  qprev.x = qprev.x / cc;
  qprev.y = qprev.y / cc;
  q[0] = from_d2<T2>(qprev);
  v[0] = vprev;
  for (int i = 1; i < n - 1; i++) {
    const double c = -pl[i] - pu[i] + eval;
    const double val = 1. / (c - pl[i] * vprev);
    const double vi = val * (pu[i]);
    double2 qi = to_d2(q[(size_t)i * st]);
    qi.x = val * (qi.x - pl[i] * qprev.x);
    qi.y = val * (qi.y - pl[i] * qprev.y);
    q[(size_t)i * st] = from_d2<T2>(qi);
    v[(size_t)i * st] = vi;
    qprev = qi;
    vprev = vi;
  }
  {
    const int i = n - 1;
    double c = -pl[i] - pu[i] + eval;
    c += pu[i];
    const double val = c - pl[i] * vprev;
    double2 qi = to_d2(q[(size_t)i * st]);
    const bool gauge = sizeof(T2) == sizeof(float2) && eval == 0.;
    if (!gauge && fabs(val) > DBL_EPSILON) {
      const double r = 1. / val;
      qi.x = r * (qi.x - pl[i] * qprev.x);
      qi.y = r * (qi.y - pl[i] * qprev.y);
    } else {
      qi = make_double2(0., 0.);
    }
    q[(size_t)i * st] = from_d2<T2>(qi);
    qprev = qi;
  }
  for (int i = n - 2; i >= 0; i--) {
    double2 qi = to_d2(q[(size_t)i * st]);
    const double vi = v[(size_t)i * st];
    qi.x -= vi * qprev.x;
    qi.y -= vi * qprev.y;
    q[(size_t)i * st] = from_d2<T2>(qi);
    qprev = qi;
  }
}

// The elimination of k_thomas_spec for one line staged in LDS (ql), both
// components by one lane, with the line's factors given (tab: c0, rlast,
// flag; val and v of rows 0 .. n-1 in vl, fl; the rows' lower coefficients
// in pl): per component the operations of k_thomas_spec with the divisions
// done in advance, its gauge choice for the last row included
template <class T2>
__device__ void thomas_line_tab(int n, const double * tab, const double * pl, const double * vl,
                                const double * fl, double2 * ql) {
  const double c0 = tab[0], rlast = tab[1];
  const bool flag = tab[2] != 0.;
  double2 qprev = ql[0];
  qprev.x = qprev.x / c0;
  qprev.y = qprev.y / c0;
  ql[0] = to_d2(from_d2<T2>(qprev));
#pragma unroll 4
  for (int i = 1; i < n - 1; i++) {
    double2 qi = ql[i];
    qi.x = vl[i] * (qi.x - pl[i] * qprev.x);
    qi.y = vl[i] * (qi.y - pl[i] * qprev.y);
    ql[i] = to_d2(from_d2<T2>(qi));
    qprev = qi;
  }
  {
    const int i = n - 1;
    double2 qi = ql[i];
    const bool gauge = sizeof(T2) == sizeof(float2);
    if (!gauge && flag) {
      qi.x = rlast * (qi.x - pl[i] * qprev.x);
      qi.y = rlast * (qi.y - pl[i] * qprev.y);
    } else {
      qi = make_double2(0., 0.);
    }
    ql[i] = to_d2(from_d2<T2>(qi));
    qprev = qi;
  }
#pragma unroll 4
  for (int i = n - 2; i >= 0; i--) {
    double2 qi = ql[i];
    qi.x -= fl[i] * qprev.x;
    qi.y -= fl[i] * qprev.y;
    ql[i] = to_d2(from_d2<T2>(qi));
    qprev = qi;
  }
}

// Partitioned x elimination of the spectrum: the tri-diagonal systems of
// k_thomas_spec (solve_linear_systems, dft.c:364-395: c = -l - u + lambda,
// Neumann folds in the first and last rows), solved in another order.  Each
// line is cut into P = min(32, nx / 2) segments of 2 .. 8 rows (nx <= 256),
// one lane per (line, segment), 8 lines per work-group of 256 lanes:
//  1. a lane loads its segment's rows (both components) into registers and
//     eliminates inside the segment, forward and then backward, so that each
//     interior row reads x_i = d_i - a_i x_s - c_i x_e in the segment's first
//     and last unknowns x_s, x_e;
//  2. the first and last rows of the P segments of a line form a
//     tri-diagonal system in those 2P unknowns; the lanes write it to LDS,
//     where one wavefront per two lines condenses it to the P unknowns x_e
//     and solves that by cyclic reduction (log2 P steps, a row per lane);
//  3. every lane forms its interior rows from x_s, x_e and stores its rows.
// The spectrum crosses HBM once each way, where the sequential elimination
// also stores and re-reads the modified right-hand side and reads its
// factors.  All systems but one are strictly diagonally dominant (lambda < 0),
// so no pivoting is needed in either elimination; the results differ from
// the sequential order's by rounding (a few 1e-14 of the line's largest
// entry).  The (ky, kz) = (0, 0) line (lambda = 0) is singular, Neumann in
// x; it keeps the sequential elimination and its choice of the constant: its
// segment lanes stage the line in LDS, where one lane eliminates it while
// the other lines solve their reduced systems, with the factors of that
// line's sweep (stab, built once on the host by the same operations) staged
// beside it (thomas_line_tab: no division left on the recurrence).  Only
// line 0 has lambda = 0 (every other eigenvalue is negative), and it is line
// t = 0 of the first work-group; a singular line elsewhere would take
// thomas_line_seq on global memory.
constexpr int kTpLines = 8;    // lines per work-group
constexpr int kTpSeg = 32;     // segment lanes per line
constexpr int kTpRows = 8;     // rows of the longest segment (nx <= 256)

template <class T2>
__device__ __forceinline__ void k_thomas_part_body(const VBlock & vb, int nx, int nlines, int nh,
                                                   const double * __restrict__ pl, const double * __restrict__ pu,
                                                   const double * __restrict__ ey, const double * __restrict__ ez,
                                                   const double * __restrict__ stab, T2 * __restrict__ spec,
                                                   double * __restrict__ vbuf) {
  // the end rows of every segment: r1 (first row: lower, diagonal, upper,
  // right-hand side re, im), r2 (last row: lower, upper, right-hand side re,
  // im; its diagonal is 1).  Segments of a line are contiguous, lines 33
  // apart (the segment lanes of 8 lines then fall in distinct banks).  The
  // reduction then publishes its rows in r1 and leaves each segment's x_s
  // and x_e (re, im) in r2.
  constexpr int RS = kTpSeg + 1;
  __shared__ double r1s[5][kTpLines][RS];
  __shared__ double r2s[4][kTpLines][RS];
  __shared__ double2 sline[kTpSeg * kTpRows];   // the staged singular line
  __shared__ double sval[kTpSeg * kTpRows];      // its factors and lower coefficients
  __shared__ double sfac[kTpSeg * kTpRows];
  __shared__ double spl[kTpSeg * kTpRows];
  const int tid = (int)threadIdx.x;
  const int t = tid % kTpLines, seg = tid / kTpLines;   // 8 consecutive lines per segment: 128-byte rows
  const int line = (int)vb.x * kTpLines + t;
  const int n = nx;
  const int P = min(kTpSeg, n / 2);
  const bool has_line = line < nlines;
  double eval = 0.;
  if (has_line) {
#if NDIMS == 2
    eval = ey[line % nh];
    (void)ez;
#else
    eval = ey[line % nh] + ez[line / nh];
#endif
  }
  const bool sing = eval == 0.;
  const bool stage = has_line && sing && t == 0 && seg < P;   // rows of the staged singular line
  // the work-group of the singular line (its line t = 0) stages its factors
  {
    const int l0 = (int)vb.x * kTpLines;
    bool wsing = false;
    if (l0 < nlines) {
#if NDIMS == 2
      wsing = ey[l0 % nh] == 0.;
#else
      wsing = ey[l0 % nh] + ez[l0 / nh] == 0.;
#endif
    }
    if (wsing)
      for (int e = tid; e < n; e += kTpLines * kTpSeg) {
        sval[e] = stab[3 + e];
        sfac[e] = stab[3 + n + e];
        spl[e] = pl[e];
      }
  }
  const bool act = has_line && !sing && seg < P;
  const int s = seg * n / P, e = (seg + 1) * n / P - 1;
  const int m = e - s + 1;
  const size_t st = (size_t)nlines;
  T2 * __restrict__ q = spec + line;
  // row i of the line: lower, diagonal and upper coefficients (the folds
  // as in k_thomas_spec; no lower entry in row 0, no upper in row n - 1)
  auto coef = [&](int i, double & lo, double & dg, double & up) {
    lo = pl[i];
    up = pu[i];
    dg = -lo - up + eval;
    if (i == 0) {
      dg += lo;
      lo = 0.;
    }
    if (i == n - 1) {
      dg += up;
      up = 0.;
    }
  };
  double A[kTpRows], C[kTpRows];
  double2 D[kTpRows];
  double2 xs = make_double2(0., 0.), xe = make_double2(0., 0.);
  if (stage) {
#pragma unroll
    for (int r = 0; r < kTpRows; r++)
      if (r < m) sline[s + r] = to_d2(q[(size_t)(s + r) * st]);
  }
  if (act) {
#pragma unroll
    for (int r = 0; r < kTpRows; r++)
      if (r < m) D[r] = to_d2(q[(size_t)(s + r) * st]);
    // forward: row r as a_r x_s + x_r + c_r x_{r+1} = d_r
    {
      double lo, dg, up;
      coef(s + 1, lo, dg, up);
      const double rc = 1. / dg;
      A[1] = lo * rc;
      C[1] = up * rc;
      D[1].x = D[1].x * rc;
      D[1].y = D[1].y * rc;
    }
#pragma unroll
    for (int r = 2; r < kTpRows; r++) {
      if (r < m) {
        double lo, dg, up;
        coef(s + r, lo, dg, up);
        const double rc = 1. / (dg - lo * C[r - 1]);
        A[r] = -lo * A[r - 1] * rc;
        C[r] = up * rc;
        D[r].x = (D[r].x - lo * D[r - 1].x) * rc;
        D[r].y = (D[r].y - lo * D[r - 1].y) * rc;
      }
    }
    // the last row, as it stands, is a reduced row
    double ae = A[1], ce = C[1];
    double2 de = D[1];
#pragma unroll
    for (int r = 2; r < kTpRows; r++) {
      if (r == m - 1) {
        ae = A[r];
        ce = C[r];
        de = D[r];
      }
    }
    // backward: interior rows as x_r = d_r - a_r x_s - c_r x_e
#pragma unroll
    for (int r = kTpRows - 3; r >= 1; r--) {
      if (r <= m - 3) {
        A[r] = A[r] - C[r] * A[r + 1];
        D[r].x = D[r].x - C[r] * D[r + 1].x;
        D[r].y = D[r].y - C[r] * D[r + 1].y;
        C[r] = -C[r] * C[r + 1];
      }
    }
    // the first row with x_{s+1} eliminated
    double lo, dg, up;
    coef(s, lo, dg, up);
    double d1 = dg, u1 = up;
    double2 r1 = D[0];
    if (m >= 3) {
      d1 = dg - up * A[1];
      u1 = -up * C[1];
      r1.x = D[0].x - up * D[1].x;
      r1.y = D[0].y - up * D[1].y;
    }
    r1s[0][t][seg] = lo;
    r1s[1][t][seg] = d1;
    r1s[2][t][seg] = u1;
    r1s[3][t][seg] = r1.x;
    r1s[4][t][seg] = r1.y;
    r2s[0][t][seg] = ae;
    r2s[1][t][seg] = ce;
    r2s[2][t][seg] = de.x;
    r2s[3][t][seg] = de.y;
  }
  __syncthreads();
  if (tid == 0 && has_line && sing) thomas_line_tab<T2>(n, stab, spl, sval, sfac, sline);
  // the reduced system of a line, one wavefront per two lines (lanes 0-31
  // line 2w, 32-63 line 2w + 1; lane k for segment k).  Segment k's rows
  //   r1:  l1 x_e(k-1) + d1 x_s(k) + u1 x_e(k) = r1
  //   r2:  a2 x_s(k) + x_e(k) + c2 x_s(k+1) = r2
  // with x_s(k) and x_s(k+1) taken from the r1 rows of segments k and k + 1
  // leave a tri-diagonal system in the P unknowns x_e(k), solved by cyclic
  // reduction: step h combines row k with rows k - h and k + h so that it
  // couples k - 2h, k, k + 2h; after ceil(log2 P) steps every row stands
  // alone.  Then x_s(k) follows from r1.  The rows are diagonally dominant as
  // the segments' rows are.  A lane keeps its row in registers and publishes
  // it, with the reciprocal of its diagonal, for its neighbours each step.  A
  // wavefront's LDS accesses complete in order, so its steps need no
  // barrier: the wavefront-scope fences only keep the compiler from moving
  // one step's reads across the previous step's writes.
  {
    const int lw = tid & 63;
    const int tl = 2 * (tid >> 6) + (lw >> 5);
    const int k = lw & 31;
    const int lr = (int)vb.x * kTpLines + tl;
    bool on = false;
    if (lr < nlines && k < P) {
#if NDIMS == 2
      on = ey[lr % nh] != 0.;
#else
      on = ey[lr % nh] + ez[lr / nh] != 0.;
#endif
    }
    double l1 = 0., d1 = 1., u1 = 0., q1x = 0., q1y = 0.;
    double own[5] = {0., 1., 0., 0., 0.};
    double rd = 1.;
    if (on) {
      l1 = r1s[0][tl][k];
      d1 = r1s[1][tl][k];
      u1 = r1s[2][tl][k];
      q1x = r1s[3][tl][k];
      q1y = r1s[4][tl][k];
      const double a2 = r2s[0][tl][k], c2 = r2s[1][tl][k];
      const double w = a2 / d1;
      own[0] = -w * l1;
      own[1] = 1. - w * u1;
      own[3] = r2s[2][tl][k] - w * q1x;
      own[4] = r2s[3][tl][k] - w * q1y;
      if (k + 1 < P) {
        const double g = c2 / r1s[1][tl][k + 1];
        own[1] -= g * r1s[0][tl][k + 1];
        own[2] = -g * r1s[2][tl][k + 1];
        own[3] -= g * r1s[3][tl][k + 1];
        own[4] -= g * r1s[4][tl][k + 1];
      }
      rd = 1. / own[1];
    }
    for (int h = 1; h < P; h *= 2) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
      if (on) {
        r1s[0][tl][k] = own[0];
        r1s[1][tl][k] = rd;
        r1s[2][tl][k] = own[2];
        r1s[3][tl][k] = own[3];
        r1s[4][tl][k] = own[4];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
      if (on) {
        double nlo = 0., nup = 0., ndg = own[1], nre = own[3], nim = own[4];
        if (k - h >= 0) {
          const double al = -own[0] * r1s[1][tl][k - h];
          nlo = al * r1s[0][tl][k - h];
          ndg += al * r1s[2][tl][k - h];
          nre += al * r1s[3][tl][k - h];
          nim += al * r1s[4][tl][k - h];
        }
        if (k + h < P) {
          const double ga = -own[2] * r1s[1][tl][k + h];
          nup = ga * r1s[2][tl][k + h];
          ndg += ga * r1s[0][tl][k + h];
          nre += ga * r1s[3][tl][k + h];
          nim += ga * r1s[4][tl][k + h];
        }
        own[0] = nlo;
        own[1] = ndg;
        own[2] = nup;
        own[3] = nre;
        own[4] = nim;
        rd = 1. / ndg;
      }
    }
    // x_e(k), then x_s(k) = (r1 - l1 x_e(k-1) - u1 x_e(k)) / d1
    const double yx = own[3] * rd, yy = own[4] * rd;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    if (on) {
      r2s[2][tl][k] = yx;
      r2s[3][tl][k] = yy;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
    if (on) {
      double px = 0., py = 0.;
      if (k > 0) {
        px = r2s[2][tl][k - 1];
        py = r2s[3][tl][k - 1];
      }
      r2s[0][tl][k] = (q1x - l1 * px - u1 * yx) / d1;
      r2s[1][tl][k] = (q1y - l1 * py - u1 * yy) / d1;
    }
  }
  __syncthreads();
  if (act) {
    xs = make_double2(r2s[0][t][seg], r2s[1][t][seg]);
    xe = make_double2(r2s[2][t][seg], r2s[3][t][seg]);
    q[(size_t)s * st] = from_d2<T2>(xs);
#pragma unroll
    for (int r = 1; r < kTpRows - 1; r++) {
      if (r < m - 1) {
        double2 x;
        x.x = D[r].x - A[r] * xs.x - C[r] * xe.x;
        x.y = D[r].y - A[r] * xs.y - C[r] * xe.y;
        q[(size_t)(s + r) * st] = from_d2<T2>(x);
      }
    }
    q[(size_t)e * st] = from_d2<T2>(xe);
  }
  if (stage) {
#pragma unroll
    for (int r = 0; r < kTpRows; r++)
      if (r < m) q[(size_t)(s + r) * st] = from_d2<T2>(sline[s + r]);
  }
  if (has_line && sing && t != 0 && seg == 0) thomas_line_seq<T2>(n, nlines, line, eval, pl, pu, spec, vbuf);
}

template <class T2>
__global__ __launch_bounds__(kTpLines * kTpSeg) void k_thomas_part(int nx, int nlines, int nh,
                                                                    const double * __restrict__ pl,
                                                                    const double * __restrict__ pu,
                                                                    const double * __restrict__ ey,
                                                                    const double * __restrict__ ez,
                                                                    const double * __restrict__ stab,
                                                                    T2 * __restrict__ spec, double * __restrict__ vbuf) {
  k_thomas_part_body<T2>(hw_block(), nx, nlines, nh, pl, pu, ey, ez, stab, spec, vbuf);
}

// inverse real transform in y (c2r: the imaginary parts of the k = 0 and
// k = n/2 bins do not contribute) and extract_output (dft.c:347-362)
template <int ELEMS, bool F256 = false>
__device__ __forceinline__ void k_irfft_y_psi_body(const VBlock & vb, 
    Grid g, const FftPlan & P, const double2 * __restrict__ tw, const double2 * __restrict__ spec,
    double * __restrict__ psi, int images) {
  __shared__ double2 sa[F256 ? kF256Cols * kF256Stride : ELEMS];
  __shared__ double2 sb[F256 ? 1 : ELEMS];
  const int nx = g.nx, ny = g.ny;
  const int N = ny;
  const int NC = F256 ? kF256Cols : ELEMS / N;
  const int CS = F256 ? kF256Stride : N;
  const Tile3 tile = xcd_tile(vb);
  const int x0 = tile.x * 2 * NC;
  const int nh = N / 2 + 1;
#if NDIMS == 2
  const int k = 1;
  const int nz = 1;
#else
  const int k = (int)tile.y + 1;
  const int nz = g.nz;
#endif
  // each stored bin kk = 0 .. N/2 is read once and placed at ky = kk and,
  // conjugated, at ky = N - kk
  const int total = NC * nh;
  for (int e0 = threadIdx.x; e0 < total; e0 += kFftLoadU * blockDim.x) {
    double2 As[kFftLoadU], Bs[kFftLoadU];
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      const int c = e / nh, kk = e % nh;
      const int xa = x0 + 2 * c, xb = xa + 1;
      As[u] = make_double2(0., 0.);
      Bs[u] = make_double2(0., 0.);
      if (e < total && xa < nx) As[u] = spec[((size_t)xa * nz + (k - 1)) * nh + kk];
      if (e < total && xb < nx) Bs[u] = spec[((size_t)xb * nz + (k - 1)) * nh + kk];
    }
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      if (e >= total) break;
      const int c = e / nh, kk = e % nh;
      double2 A = As[u], B = Bs[u];
      if (kk == 0 || 2 * kk == N) { A.y = 0.; B.y = 0.; }
      // Z = A + i B at ky = kk; the conjugates (imaginary parts times -1) at N - kk
      sa[c * CS + fft_in<F256>(kk)] = make_double2(A.x - B.y, A.y + B.x);
      if (kk != 0 && 2 * kk != N) {
        const double ay = A.y * -1., by = B.y * -1.;
        sa[c * CS + fft_in<F256>(N - kk)] = make_double2(A.x - by, ay + B.x);
      }
    }
  }
  __syncthreads();
  double2 * res = fft_cols<F256, double2>(sa, sb, NC, P, tw, +1.);
  for (int e = threadIdx.x; e < 2 * NC * N; e += blockDim.x) {
    const int xl = e % (2 * NC), j = e / (2 * NC);
    const int i = x0 + xl + 1;
    if (i <= nx) {
      const double2 z = res[(xl >> 1) * CS + fft_out<F256>(j)];
      const double v = (xl & 1) ? z.y : z.x;
      if (images) {
        // fluid_update_boundaries_psi folded in: halo images, Neumann walls
        const int pl = NDIMS == 3 ? k : 0;
        store_periodic<false>(psi, i, nx + 2, ny + 2, j + 1, pl, ny, nz, 1, v);
        neumann_walls(psi, i, nx, ny + 2, j + 1, pl, ny, nz, v);
      } else {
        psi[I_C1(i, j + 1, k)] = v;
      }
    }
  }
}

template <int ELEMS, bool F256 = false>
__global__ __launch_bounds__(kFftThreads) void k_irfft_y_psi(
    Grid g, FftPlan P, const double2 * __restrict__ tw, const double2 * __restrict__ spec,
    double * __restrict__ psi, int images) {
  k_irfft_y_psi_body<ELEMS, F256>(hw_block(), g, P, tw, spec, psi, images);
}

// ---------------------------------------------------------------------------
// Mixed-precision Poisson solve: the transforms run in FP32 on an FP32
// spectrum, the x eliminations in FP64 (k_thomas_spec<float2>), and FP64
// residual corrections refine psi:
//   psi = M(r0);  repeat: psi += M(r0 - A psi)
// with r0 = div(u*) / (gamma dt) and A the discrete Laplacian the FP64
// solver inverts (second differences in y and z, the l/u coefficients and
// Neumann walls in x).  M includes the 1/(ny nz) of the unnormalised
// transforms.

// mode 0: r = prefactor * div u, stored in rhs when store_rhs (a refinement
// pass follows); mode 1: r = rhs - A psi.  Then (float)(r * scale) is
// transformed in y as k_div_rfft_y does.
template <int ELEMS, bool F256 = false>
__device__ __forceinline__ void k_res_rfft32_y_body(const VBlock & vb, 
    Grid g, const FftPlan & P, const float2 * __restrict__ tw, const double * __restrict__ ux,
    const double * __restrict__ uy, const double * __restrict__ uz, double prefactor, double scale,
    double * __restrict__ rhs, const double * __restrict__ psi, const double * __restrict__ pl,
    const double * __restrict__ pu, int mode, int store_rhs, float2 * __restrict__ spec) {
  __shared__ float2 sa[F256 ? kF256Cols * kF256Stride : ELEMS];
  __shared__ float2 sb[F256 ? 1 : ELEMS];
  const int nx = g.nx, ny = g.ny;
  const int N = ny;
  const int NC = F256 ? kF256Cols : ELEMS / N;
  const int CS = F256 ? kF256Stride : N;
  const Tile3 tile = xcd_tile(vb);
  const int x0 = tile.x * 2 * NC;
#if NDIMS == 2
  const int k = 1;
  const int nz = 1;
  (void)uz;
#else
  const int k = (int)tile.y + 1;
  const int nz = g.nz;
#endif
  const double rdy2 = 1. / (g.dy * g.dy);
#if NDIMS == 3
  const double rdz2 = 1. / (g.dz * g.dz);
#endif
  // kFftLoadU elements per thread have their loads in flight together
  const int total = 2 * NC * N;
  for (int e0 = threadIdx.x; e0 < total; e0 += kFftLoadU * blockDim.x) {
    // mode 0: ux(i+1) ux(i) uy(j+1) uy(j) uz(k+1) uz(k), DXF(i); mode 1: psi
    // at the centre, x-, x+, y-, y+, z-, z+ and rhs (every load of an element
    // in the first loop, so all of them are in flight together)
    double a[8][kFftLoadU];
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      // indices past the grid are clamped to it (values unused), so the
      // loads need no divergent branch
      const int e = e0 + u * (int)blockDim.x;
      const int xl = e % (2 * NC);
      const int j = min(e / (2 * NC) + 1, ny);
      const int i = min(x0 + xl + 1, nx);
      a[4][u] = a[5][u] = a[6][u] = 0.;
      if (mode == 0) {
        a[7][u] = g.dxf[i - 1];
        a[0][u] = ux[I_UX(i + 1, j, k)];
        a[1][u] = ux[I_UX(i, j, k)];
        a[2][u] = uy[I_C1(i, j + 1, k)];
        a[3][u] = uy[I_C1(i, j, k)];
#if NDIMS == 3
        a[4][u] = uz[I_C1(i, j, k + 1)];
        a[5][u] = uz[I_C1(i, j, k)];
#endif
      } else {
        const size_t ir = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
        a[0][u] = psi[I_C1(i, j, k)];
        a[1][u] = psi[I_C1(i - 1, j, k)];
        a[2][u] = psi[I_C1(i + 1, j, k)];
        a[3][u] = psi[I_C1(i, j - 1, k)];
        a[4][u] = psi[I_C1(i, j + 1, k)];
#if NDIMS == 3
        a[5][u] = psi[I_C1(i, j, k - 1)];
        a[6][u] = psi[I_C1(i, j, k + 1)];
#endif
        a[7][u] = rhs[ir];
      }
    }
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      if (e >= total) break;
      const int xl = e % (2 * NC), j = e / (2 * NC) + 1;
      const int i = x0 + xl + 1;
      float r32 = 0.f;
      if (i <= nx) {
        const size_t ir = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
        double r;
        if (mode == 0) {
          const double dx = a[7][u];
          double div = (a[0][u] - a[1][u]) / dx + (a[2][u] - a[3][u]) / g.dy;
#if NDIMS == 3
          div += (a[4][u] - a[5][u]) / g.dz;
#endif
          r = prefactor * div;
          if (store_rhs) rhs[ir] = r;
        } else {
          const double pc = a[0][u];
          // psi(0) = psi(1) and psi(nx+1) = psi(nx) (Neumann images) fold the
          // wall rows of the x operator
          double ap = pl[i - 1] * (a[1][u] - pc) + pu[i - 1] * (a[2][u] - pc);
          ap += rdy2 * (a[3][u] - 2. * pc + a[4][u]);
#if NDIMS == 3
          ap += rdz2 * (a[5][u] - 2. * pc + a[6][u]);
#endif
          r = a[7][u] - ap;
        }
        r32 = (float)(r * scale);
      }
      float * slot = reinterpret_cast<float *>(&sa[(xl >> 1) * CS + fft_in<F256>(j - 1)]);
      slot[xl & 1] = r32;
    }
  }
  __syncthreads();
  float2 * res = fft_cols<F256, float2>(sa, sb, NC, P, tw, -1.f);
  const int nh = N / 2 + 1;
  for (int e = threadIdx.x; e < NC * nh; e += blockDim.x) {
    const int c = e / nh, ky = e % nh;
    const float2 z = res[c * CS + fft_out<F256>(ky)];
    const float2 zc = res[c * CS + fft_out<F256>((N - ky) % N)];
    const float2 A = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
    const float2 B = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
    const int xa = x0 + 2 * c, xb = xa + 1;
    if (xa < nx) spec[((size_t)xa * nz + (k - 1)) * nh + ky] = A;
    if (xb < nx) spec[((size_t)xb * nz + (k - 1)) * nh + ky] = B;
  }
}

template <int ELEMS, bool F256 = false>
__global__ __launch_bounds__(kFftThreads) void k_res_rfft32_y(
    Grid g, FftPlan P, const float2 * __restrict__ tw, const double * __restrict__ ux,
    const double * __restrict__ uy, const double * __restrict__ uz, double prefactor, double scale,
    double * __restrict__ rhs, const double * __restrict__ psi, const double * __restrict__ pl,
    const double * __restrict__ pu, int mode, int store_rhs, float2 * __restrict__ spec) {
  k_res_rfft32_y_body<ELEMS, F256>(hw_block(), g, P, tw, ux, uy, uz, prefactor, scale, rhs, psi, pl, pu, mode,
                                   store_rhs, spec);
}

// inverse real transform in y in FP32; psi = delta (accumulate 0) or
// psi += delta (accumulate 1) in FP64, with halo images and Neumann walls
template <int ELEMS, bool F256 = false>
__device__ __forceinline__ void k_irfft32_y_psi_body(const VBlock & vb, 
    Grid g, const FftPlan & P, const float2 * __restrict__ tw, const float2 * __restrict__ spec,
    double * __restrict__ psi, int accumulate) {
  __shared__ float2 sa[F256 ? kF256Cols * kF256Stride : ELEMS];
  __shared__ float2 sb[F256 ? 1 : ELEMS];
  const int nx = g.nx, ny = g.ny;
  const int N = ny;
  const int NC = F256 ? kF256Cols : ELEMS / N;
  const int CS = F256 ? kF256Stride : N;
  const Tile3 tile = xcd_tile(vb);
  const int x0 = tile.x * 2 * NC;
  const int nh = N / 2 + 1;
#if NDIMS == 2
  const int k = 1;
  const int nz = 1;
#else
  const int k = (int)tile.y + 1;
  const int nz = g.nz;
#endif
  // each stored bin kk = 0 .. N/2 is read once and placed at ky = kk and,
  // conjugated, at ky = N - kk
  const int total = NC * nh;
  for (int e0 = threadIdx.x; e0 < total; e0 += kFftLoadU * blockDim.x) {
    float2 As[kFftLoadU], Bs[kFftLoadU];
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      const int c = e / nh, kk = e % nh;
      const int xa = x0 + 2 * c, xb = xa + 1;
      As[u] = make_float2(0.f, 0.f);
      Bs[u] = make_float2(0.f, 0.f);
      if (e < total && xa < nx) As[u] = spec[((size_t)xa * nz + (k - 1)) * nh + kk];
      if (e < total && xb < nx) Bs[u] = spec[((size_t)xb * nz + (k - 1)) * nh + kk];
    }
#pragma unroll
    for (int u = 0; u < kFftLoadU; u++) {
      const int e = e0 + u * (int)blockDim.x;
      if (e >= total) break;
      const int c = e / nh, kk = e % nh;
      float2 A = As[u], B = Bs[u];
      if (kk == 0 || 2 * kk == N) { A.y = 0.f; B.y = 0.f; }
      sa[c * CS + fft_in<F256>(kk)] = make_float2(A.x - B.y, A.y + B.x);
      if (kk != 0 && 2 * kk != N) {
        const float ay = A.y * -1.f, by = B.y * -1.f;
        sa[c * CS + fft_in<F256>(N - kk)] = make_float2(A.x - by, ay + B.x);
      }
    }
  }
  __syncthreads();
  float2 * res = fft_cols<F256, float2>(sa, sb, NC, P, tw, +1.f);
  for (int e = threadIdx.x; e < 2 * NC * N; e += blockDim.x) {
    const int xl = e % (2 * NC), j = e / (2 * NC);
    const int i = x0 + xl + 1;
    if (i <= nx) {
      const float2 z = res[(xl >> 1) * CS + fft_out<F256>(j)];
      double v = (double)((xl & 1) ? z.y : z.x);
      if (accumulate) v = psi[I_C1(i, j + 1, k)] + v;
      const int pln = NDIMS == 3 ? k : 0;
      store_periodic<false>(psi, i, nx + 2, ny + 2, j + 1, pln, ny, nz, 1, v);
      neumann_walls(psi, i, nx, ny + 2, j + 1, pln, ny, nz, v);
    }
  }
}

template <int ELEMS, bool F256 = false>
__global__ __launch_bounds__(kFftThreads) void k_irfft32_y_psi(
    Grid g, FftPlan P, const float2 * __restrict__ tw, const float2 * __restrict__ spec,
    double * __restrict__ psi, int accumulate) {
  k_irfft32_y_psi_body<ELEMS, F256>(hw_block(), g, P, tw, spec, psi, accumulate);
}

// ---------------------------------------------------------------------------
// fluid_correct_velocity (correct_velocity/{ux,uy}.c) for one component
__device__ __forceinline__ void k_correct_body(const VBlock & vb, Grid g, double prefactor, const double * __restrict__ psi,
                          double * __restrict__ ux, double * __restrict__ uy,
                          double * __restrict__ uz) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
  const double pc = psi[I_C1(i, j, k)];
  if (i >= 2) {
    const double dx = g.dxc[i - 1];
    ux[I_UX(i, j, k)] -= prefactor / dx * (+pc - psi[I_C1(i - 1, j, k)]);
  }
  uy[I_C1(i, j, k)] -= prefactor / g.dy * (+pc - psi[I_C1(i, j - 1, k)]);
#if NDIMS == 3
  uz[I_C1(i, j, k)] -= prefactor / g.dz * (+pc - psi[I_C1(i, j, k - 1)]);
#else
  (void)uz;
#endif
}

__global__ void k_correct(Grid g, double prefactor, const double * __restrict__ psi,
                          double * __restrict__ ux, double * __restrict__ uy,
                          double * __restrict__ uz) {
  k_correct_body(hw_block(), g, prefactor, psi, ux, uy, uz);
}

// fluid_update_pressure (update_pressure.c:12-109): p += psi, then the
// implicit-x correction
__device__ __forceinline__ void k_pressure_body(const VBlock & vb, Grid g, double prefactor, const double * __restrict__ psi,
                           double * __restrict__ p) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
  const double pc = psi[I_C1(i, j, k)];
  const double pm = psi[I_C1(i - 1, j, k)], pp = psi[I_C1(i + 1, j, k)];
  double v = p[I_C1(i, j, k)];
  v += pc;
  const double dpsidx_xm = (-pm + pc) / g.dxc[i - 1];
  const double dpsidx_xp = (-pc + pp) / g.dxc[i];
  v -= prefactor / g.dxf[i - 1] * (-dpsidx_xm + dpsidx_xp);
  p[I_C1(i, j, k)] = v;
}

__global__ void k_pressure(Grid g, double prefactor, const double * __restrict__ psi,
                           double * __restrict__ p) {
  k_pressure_body(hw_block(), g, prefactor, psi, p);
}

// ---------------------------------------------------------------------------
// fluid_decide_dt (decide_dt.c:28-64): min over DOFs of spacing / (|u| + 1e-8)
constexpr int kRedThreads = 256;

__device__ inline double block_min(double v) {
  __shared__ double red[kRedThreads];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kRedThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  return red[0];
}

constexpr int kDtU = 16;  // planes per pass of k_dt_partial (a whole 256-plane column at 16 z groups)

__device__ __forceinline__ void k_dt_partial_body(const VBlock & vb, Grid g, const double * __restrict__ ux,
                                                            const double * __restrict__ uy,
                                                            const double * __restrict__ uz,
                                                            double * __restrict__ part) {
  const int nx = g.nx, ny = g.ny;
#if NDIMS == 2
  const size_t ncell = (size_t)nx * ny;
  (void)uz;
#else
  const size_t ncell = (size_t)nx * ny * g.nz;
#endif
  const double small = 1.e-8;
  double m = 1.;
  // 64 x 4 tiles of (i, j): thread (i, j0) of work-group (x, y, z) walks
  // the rows j0, j0 + 4 gy, ... and planes z + 1, z + 1 + gz, ... of its
  // column (no index division)
  const int i = (int)(vb.x * 64 + threadIdx.x % 64) + 1;
  (void)ncell;
  if (i <= nx) {
    const double dxc = i >= 2 ? g.dxc[i - 1] : 1.;
    for (int j = (int)(vb.y * 4 + threadIdx.x / 64) + 1; j <= ny; j += 4 * (int)vb.gy) {
#if NDIMS == 2
      if (i >= 2) m = fmin(m, dxc / (fabs(ux[I_UX(i, j, 1)]) + small));
      m = fmin(m, g.dy / (fabs(uy[I_C1(i, j, 1)]) + small));
#else
      // kDtU planes per pass with their loads issued together (planes past
      // nz are clamped to it and skipped; the minimum does not depend on
      // the order its terms are taken in)
      const int gz = (int)vb.gz;
      for (int k0 = (int)vb.z + 1; k0 <= g.nz; k0 += kDtU * gz) {
        double a[kDtU], b[kDtU], c[kDtU];
#pragma unroll
        for (int u = 0; u < kDtU; u++) {
          const int k = min(k0 + u * gz, g.nz);
          a[u] = ux[I_UX(i, j, k)];   // i = 1: the wall entry, loaded and skipped
          b[u] = uy[I_C1(i, j, k)];
          c[u] = uz[I_C1(i, j, k)];
        }
#pragma unroll
        for (int u = 0; u < kDtU; u++) {
          if (k0 + u * gz > g.nz) break;
          if (i >= 2) m = fmin(m, dxc / (fabs(a[u]) + small));
          m = fmin(m, g.dy / (fabs(b[u]) + small));
          m = fmin(m, g.dz / (fabs(c[u]) + small));
        }
      }
#endif
    }
  }
  m = block_min(m);
  if (threadIdx.x == 0) part[vb.x + vb.gx * (vb.y + vb.gy * vb.z)] = m;
}

__global__ __launch_bounds__(kRedThreads) void k_dt_partial(Grid g, const double * __restrict__ ux,
                                                            const double * __restrict__ uy,
                                                            const double * __restrict__ uz,
                                                            double * __restrict__ part) {
  k_dt_partial_body(hw_block(), g, ux, uy, uz, part);
}

__device__ __forceinline__ void k_min_final_body(const VBlock & vb, const double * __restrict__ part, int n,
                                                           double * __restrict__ out) {
  double m = 1.;
  for (int t = threadIdx.x; t < n; t += blockDim.x) m = fmin(m, part[t]);
  m = block_min(m);
  if (threadIdx.x == 0) out[0] = m;
}

__global__ __launch_bounds__(kRedThreads) void k_min_final(const double * __restrict__ part, int n,
                                                           double * __restrict__ out) {
  k_min_final_body(hw_block(), part, n, out);
}

// ---------------------------------------------------------------------------
// diagnostics (logging/divergence.c, interface.c, energy.c): per-block
// partials {max|div|, vof min, vof max, sum vof dV, sum dV, KE x, y, z, TE}
// partial sums: 0 max|div|, 1 vof min, 2 vof max, 3 vof volume, 4 volume,
// 5-7 kinetic energy x y z, 8 thermal energy, 9-11 momentum x y z,
// 12 interface measure sum |grad vof| dV
constexpr int kDiagN = 13;

__device__ __forceinline__ void k_diag_partial_body(const VBlock & vb, 
    Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
    const double * __restrict__ uz, const double * __restrict__ t,
    const double * __restrict__ vof, double * __restrict__ part) {
  const int nx = g.nx, ny = g.ny;
#if NDIMS == 2
  const size_t ncell = (size_t)nx * ny;
  const double dyz = g.dy;
  (void)uz;
#else
  const size_t ncell = (size_t)nx * ny * g.nz;
  const double dyz = g.dy * g.dz;
#endif
  double q[kDiagN];
  for (int m = 0; m < kDiagN; m++) q[m] = m == 1 ? 1. : 0.;
  for (size_t s = (size_t)vb.x * blockDim.x + threadIdx.x; s < ncell;
       s += (size_t)vb.gx * blockDim.x) {
    const int i = (int)(s % nx) + 1;
    const size_t r = s / nx;
    const int j = (int)(r % ny) + 1;
#if NDIMS == 2
    const int k = 1;
#else
    const int k = (int)(r / ny) + 1;
#endif
    const double dx = g.dxf[i - 1];
    double div = (ux[I_UX(i + 1, j, k)] - ux[I_UX(i, j, k)]) / dx +
                 (uy[I_C1(i, j + 1, k)] - uy[I_C1(i, j, k)]) / g.dy;
#if NDIMS == 3
    div += (uz[I_C1(i, j, k + 1)] - uz[I_C1(i, j, k)]) / g.dz;
    const double w = uz[I_C1(i, j, k)];
    q[7] += 0.5 * (w * w) * (dx * dyz);
    q[11] += w * (dx * dyz);
#endif
    q[0] = fmax(q[0], fabs(div));
    const double lv = vof[I_VOF(i, j, k)];
    q[1] = fmin(q[1], lv);
    q[2] = fmax(q[2], lv);
    q[3] += lv * (dx * dyz);
    q[4] += dx * dyz;
    if (i >= 2) {
      const double u = ux[I_UX(i, j, k)];
      q[5] += 0.5 * (u * u) * (g.dxc[i - 1] * dyz);
      q[9] += u * (g.dxc[i - 1] * dyz);
    }
    const double v = uy[I_C1(i, j, k)];
    q[6] += 0.5 * (v * v) * (dx * dyz);
    q[10] += v * (dx * dyz);
    const double tt = t[I_C1(i, j, k)];
    q[8] += 0.5 * (tt * tt) * (dx * dyz);
    {
      const double gx = (vof[I_VOF(i + 1, j, k)] - vof[I_VOF(i - 1, j, k)]) / (g.dxc[i - 1] + g.dxc[i]);
      const double gy = (vof[I_VOF(i, j + 1, k)] - vof[I_VOF(i, j - 1, k)]) / (2. * g.dy);
#if NDIMS == 2
      const double g2 = gx * gx + gy * gy;
#else
      const double gz = (vof[I_VOF(i, j, k + 1)] - vof[I_VOF(i, j, k - 1)]) / (2. * g.dz);
      const double g2 = gx * gx + gy * gy + gz * gz;
#endif
      q[12] += sqrt(g2) * (dx * dyz);
    }
  }
  __shared__ double red[kDiagN][kRedThreads];
  for (int m = 0; m < kDiagN; m++) red[m][threadIdx.x] = q[m];
  __syncthreads();
  for (int st = kRedThreads / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      const int o = threadIdx.x + st;
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][o]);
      red[1][threadIdx.x] = fmin(red[1][threadIdx.x], red[1][o]);
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][o]);
      for (int m = 3; m < kDiagN; m++) red[m][threadIdx.x] += red[m][o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < kDiagN; m++) part[(size_t)vb.x * kDiagN + m] = red[m][0];
}

__global__ __launch_bounds__(kRedThreads) void k_diag_partial(
    Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
    const double * __restrict__ uz, const double * __restrict__ t,
    const double * __restrict__ vof, double * __restrict__ part) {
  k_diag_partial_body(hw_block(), g, ux, uy, uz, t, vof, part);
}

// logging/nusselt: partial sums of {wall heat flux, kinetic energy
// injection, kinetic energy dissipation, thermal dissipation} over the x
// faces i = 1..nx+1 of every interior row (cell terms use i <= nx); 3D adds
// the z terms of each integral
constexpr int kNuN = 4;

__device__ inline double sq(double a) { return a * a; }

__device__ __forceinline__ void k_nusselt_partial_body(const VBlock & vb, 
    Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
    const double * __restrict__ uz, const double * __restrict__ t, double md, double td,
    double * __restrict__ part) {
  const int nx = g.nx, ny = g.ny;
#if NDIMS == 2
  const size_t nrow = (size_t)ny;
  const double dyz = g.dy;
  (void)uz;
#else
  const size_t nrow = (size_t)ny * g.nz;
  const double dyz = g.dy * g.dz;
  const double dz = g.dz;
#endif
  const double dy = g.dy;
  const double * dxf = g.dxf;
  const double * dxc = g.dxc;
  double q[kNuN] = {0., 0., 0., 0.};
  const size_t n = (size_t)(nx + 1) * nrow;
  for (size_t s = (size_t)vb.x * blockDim.x + threadIdx.x; s < n; s += (size_t)vb.gx * blockDim.x) {
    const int i = (int)(s % (size_t)(nx + 1)) + 1;
    const size_t r = s / (size_t)(nx + 1);
    const int j = (int)(r % ny) + 1;
#if NDIMS == 2
    const int k = 1;
#else
    const int k = (int)(r / ny) + 1;
#endif
    // DXC(i) = dxc[i - 1], DXF(i) = dxf[i - 1]
    if (i == 1) {
      const double dtdx_xm = (+t[I_C1(1, j, k)] - t[I_C1(0, j, k)]) / dxc[0];
      const double dtdx_xp = (+t[I_C1(nx + 1, j, k)] - t[I_C1(nx, j, k)]) / dxc[nx];
      q[0] -= 0.5 * td * (dtdx_xm + dtdx_xp) * dyz;
    }
    if (i >= 2 && i <= nx) {
      const double t_ = +0.5 * t[I_C1(i - 1, j, k)] + 0.5 * t[I_C1(i, j, k)];
      q[1] += ux[I_UX(i, j, k)] * t_ * (dxc[i - 1] * dyz);
    }
    {
      // ux terms on the faces i = 1..nx+1
      const double cs = dxc[i - 1] * dyz;
      const double uc = ux[I_UX(i, j, k)];
      if (i != 1) q[2] += md * 0.5 * (dxf[i - 2] / dxc[i - 1]) * sq(1. / dxf[i - 2] * (uc - ux[I_UX(i - 1, j, k)])) * cs;
      if (i != nx + 1) q[2] += md * 0.5 * (dxf[i - 1] / dxc[i - 1]) * sq(1. / dxf[i - 1] * (ux[I_UX(i + 1, j, k)] - uc)) * cs;
      q[2] += md * 0.5 * sq(1. / dy * (uc - ux[I_UX(i, j - 1, k)])) * cs;
      q[2] += md * 0.5 * sq(1. / dy * (ux[I_UX(i, j + 1, k)] - uc)) * cs;
#if NDIMS == 3
      q[2] += md * 0.5 * sq(1. / dz * (uc - ux[I_UX(i, j, k - 1)])) * cs;
      q[2] += md * 0.5 * sq(1. / dz * (ux[I_UX(i, j, k + 1)] - uc)) * cs;
#endif
    }
    if (i <= nx) {
      const double cs = dxf[i - 1] * dyz;
      const double w1m = 1 == i ? 2. : 1.;
      const double w1p = nx == i ? 2. : 1.;
      const double wm = dxc[i - 1] / dxf[i - 1];
      const double wp = dxc[i] / dxf[i - 1];
      // uy (and uz) terms on the cell centres
      const double vc = uy[I_C1(i, j, k)];
      q[2] += md * 0.5 * wm * w1m * sq(1. / dxc[i - 1] * (vc - uy[I_C1(i - 1, j, k)])) * cs;
      q[2] += md * 0.5 * wp * w1p * sq(1. / dxc[i] * (uy[I_C1(i + 1, j, k)] - vc)) * cs;
      q[2] += md * 0.5 * sq(1. / dy * (vc - uy[I_C1(i, j - 1, k)])) * cs;
      q[2] += md * 0.5 * sq(1. / dy * (uy[I_C1(i, j + 1, k)] - vc)) * cs;
#if NDIMS == 3
      q[2] += md * 0.5 * sq(1. / dz * (vc - uy[I_C1(i, j, k - 1)])) * cs;
      q[2] += md * 0.5 * sq(1. / dz * (uy[I_C1(i, j, k + 1)] - vc)) * cs;
      const double wc = uz[I_C1(i, j, k)];
      q[2] += md * 0.5 * wm * w1m * sq(1. / dxc[i - 1] * (wc - uz[I_C1(i - 1, j, k)])) * cs;
      q[2] += md * 0.5 * wp * w1p * sq(1. / dxc[i] * (uz[I_C1(i + 1, j, k)] - wc)) * cs;
      q[2] += md * 0.5 * sq(1. / dy * (wc - uz[I_C1(i, j - 1, k)])) * cs;
      q[2] += md * 0.5 * sq(1. / dy * (uz[I_C1(i, j + 1, k)] - wc)) * cs;
      q[2] += md * 0.5 * sq(1. / dz * (wc - uz[I_C1(i, j, k - 1)])) * cs;
      q[2] += md * 0.5 * sq(1. / dz * (uz[I_C1(i, j, k + 1)] - wc)) * cs;
#endif
      const double tc = t[I_C1(i, j, k)];
      q[3] += td * 0.5 * wm * w1m * sq(1. / dxc[i - 1] * (tc - t[I_C1(i - 1, j, k)])) * cs;
      q[3] += td * 0.5 * wp * w1p * sq(1. / dxc[i] * (t[I_C1(i + 1, j, k)] - tc)) * cs;
      q[3] += td * 0.5 * sq(1. / dy * (tc - t[I_C1(i, j - 1, k)])) * cs;
      q[3] += td * 0.5 * sq(1. / dy * (t[I_C1(i, j + 1, k)] - tc)) * cs;
#if NDIMS == 3
      q[3] += td * 0.5 * sq(1. / dz * (tc - t[I_C1(i, j, k - 1)])) * cs;
      q[3] += td * 0.5 * sq(1. / dz * (t[I_C1(i, j, k + 1)] - tc)) * cs;
#endif
    }
  }
  __shared__ double red[kNuN][kRedThreads];
  for (int m = 0; m < kNuN; m++) red[m][threadIdx.x] = q[m];
  __syncthreads();
  for (int st = kRedThreads / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int m = 0; m < kNuN; m++) red[m][threadIdx.x] += red[m][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int m = 0; m < kNuN; m++) part[(size_t)vb.x * kNuN + m] = red[m][0];
}

__global__ __launch_bounds__(kRedThreads) void k_nusselt_partial(
    Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
    const double * __restrict__ uz, const double * __restrict__ t, double md, double td,
    double * __restrict__ part) {
  k_nusselt_partial_body(hw_block(), g, ux, uy, uz, t, md, td, part);
}

// statistics.c:117-197: add one sample to the accumulators; faces
// i = 1..nx+1 (ux1 ux2 uxt, row width nx+1), cells i = 0..nx+1 (uy1 uy2 t1
// t2 [uz1 uz2], row width nx+2); rows are (j, k) in y-fastest order
struct StatAcc {
  double * a[9];   // ux1 ux2 uy1 uy2 t1 t2 uxt uz1 uz2
};

__device__ __forceinline__ void k_stats_collect_body(const VBlock & vb, Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                                const double * __restrict__ uz, const double * __restrict__ t, StatAcc acc) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(vb.x * blockDim.x + threadIdx.x);   // 0..nx+1
  const int j = (int)(vb.y * blockDim.y + threadIdx.y) + 1;
  const int k = (int)vb.z + 1;
  if (i > nx + 1 || j > ny) return;
  const size_t r = (size_t)(j - 1) + (size_t)ny * (size_t)(k - 1);
  if (i >= 1) {
    const size_t e = (size_t)(i - 1) + (size_t)(nx + 1) * r;
    const double u = ux[I_UX(i, j, k)];
    acc.a[0][e] += u;
    acc.a[1][e] += u * u;
    const double t_ = +0.5 * t[I_C1(i - 1, j, k)] + 0.5 * t[I_C1(i, j, k)];
    acc.a[6][e] += u * t_;
  }
  const size_t e = (size_t)i + (size_t)(nx + 2) * r;
  const double v = uy[I_C1(i, j, k)];
  acc.a[2][e] += v;
  acc.a[3][e] += v * v;
  const double tt = t[I_C1(i, j, k)];
  acc.a[4][e] += tt;
  acc.a[5][e] += tt * tt;
#if NDIMS == 3
  const double w = uz[I_C1(i, j, k)];
  acc.a[7][e] += w;
  acc.a[8][e] += w * w;
#else
  (void)uz;
#endif
}

__global__ void k_stats_collect(Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                                const double * __restrict__ uz, const double * __restrict__ t, StatAcc acc) {
  k_stats_collect_body(hw_block(), g, ux, uy, uz, t, acc);
}

// statistics.c:232-275: sum over the rows in row order, one lane per x entry
__device__ __forceinline__ void k_stats_reduce_body(const VBlock & vb, const double * __restrict__ a, int w, size_t rows, double * __restrict__ out) {
  const int i = (int)(vb.x * blockDim.x + threadIdx.x);
  if (i >= w) return;
  double s = 0.;
  for (size_t r = 0; r < rows; r++) s += a[(size_t)i + (size_t)w * r];
  out[i] = s;
}

__global__ void k_stats_reduce(const double * __restrict__ a, int w, size_t rows, double * __restrict__ out) {
  k_stats_reduce_body(hw_block(), a, w, rows, out);
}

// plain copy, the calibration kernel for hardware byte counters
__device__ __forceinline__ void k_copy_body(const VBlock & vb, const double * __restrict__ a, double * __restrict__ b, size_t n) {
  for (size_t t = (size_t)vb.x * blockDim.x + threadIdx.x; t < n; t += (size_t)vb.gx * blockDim.x)
    b[t] = a[t];
}

__global__ void k_copy(const double * __restrict__ a, double * __restrict__ b, size_t n) {
  k_copy_body(hw_block(), a, b, n);
}


// ===========================================================================
// Fused schedule.  Each kernel below produces exactly the values of the
// per-function kernels above (same expressions in the same order) and also
// writes the periodic halo images and x walls the reference's boundary
// update (src/halo.c:18-76 then fluid/boundary/*.c) would produce, so no
// separate boundary launch is needed.

#if NDIMS == 2
#define PLANE1(k) 0
#define PLANE2(k) 0
#else
#define PLANE1(k) (k)
#define PLANE2(k) ((k) + 1)
#endif

// interface_compute_curvature_tensor fused: corner normals of a tile are
// computed once into LDS (compute_gradient), then each cell forms its
// curvature and, at interface cells, its normal and THINC intercept
// (compute_normal / compute_curvature).  Interior cells only; the halo
// cells get their periodic images.
#if NDIMS == 2
constexpr int kIX = 64, kIY = 8, kIZ = 1;
#else
constexpr int kIX = 32, kIY = 4, kIZ = 4;
#endif
constexpr int kIThreads = kIX * kIY * kIZ;

template <int FZ>
__device__ __forceinline__ void k_iface_fused_fz_body(const VBlock & vb, Grid g, const double * __restrict__ vof,
                                                      double * __restrict__ normal, double * __restrict__ curv,
                                                      const Freeze & fz) {
  static_assert(FZ == 0 || NDIMS == 2, "decision replay is 2D");
  constexpr int VX = kIX + 2, VY = kIY + 2;
  constexpr int CX = kIX + 1, CY = kIY + 1;
#if NDIMS == 2
  constexpr int VZ = 1, CZ = 1;
#else
  constexpr int VZ = kIZ + 2, CZ = kIZ + 1;
#endif
  __shared__ double sv[VX * VY * VZ];
  __shared__ double sc[NDIMS][CX * CY * CZ];
  // interface cells of the tile (vof, unit normal), solved for their
  // intercepts one per lane after every cell is classified: in place, a
  // wavefront would run the Newton solve with only its few interface lanes
  __shared__ double lv[NDIMS + 1][kIThreads];
  __shared__ int lc[kIThreads];
  __shared__ int ln;
  const int nx = g.nx, ny = g.ny;
  const Tile3 tile = g.bands ? band_tile(vb) : plain_tile(vb);
  const int i0 = 1 + tile.x * kIX, j0 = 1 + tile.y * kIY;
#if NDIMS == 3
  const int nz = g.nz;
  const int k0 = 1 + tile.z * kIZ;
#else
  const int nz = 1;
  const int k0 = 1;
#endif
  const int tid = threadIdx.x;
  if (tid == 0) ln = 0;
  // vof tile: cells (i0-1 .. i0+kIX, j0-1 .., k0-1 ..)
  for (int e = tid; e < VX * VY * VZ; e += kIThreads) {
    const int a = e % VX, r = e / VX;
    const int b = r % VY;
    const int c = r / VY;
    const int i = i0 - 1 + a, j = j0 - 1 + b;
#if NDIMS == 3
    const int k = k0 - 1 + c;
    const bool ok = i <= nx + 1 && j <= ny + 2 && k <= nz + 2;
#else
    const int k = 1;
    (void)c;
    const bool ok = i <= nx + 1 && j <= ny + 2;
#endif
    sv[e] = ok ? vof[I_VOF(i, j, k)] : 0.;
  }
  __syncthreads();
  // corner normals (i0 .. i0+kIX, j0 .., k0 ..): compute_gradient
  for (int e = tid; e < CX * CY * CZ; e += kIThreads) {
    const int a = e % CX, r = e / CX;
    const int b = r % CY;
    const int c = r / CY;
    const int i = i0 + a;
    const double rdx = g.xt[kXT * min(i, nx + 1) + XT_RXC];   // 1 / DXC
#define SV(A, B, C) sv[(A) + VX * ((B) + VY * (C))]
#if NDIMS == 2
    (void)c;
    const double va = SV(a, b, 0), vb = SV(a + 1, b, 0);
    const double vc = SV(a, b + 1, 0), vd = SV(a + 1, b + 1, 0);
    const double gx = rdx * (-va + vb - vc + vd);
    const double gy = g.rdy * (-va - vb + vc + vd);
    const double norm = sqrt(gx * gx + gy * gy);
    const double ninv = 1. / fmax(norm, DBL_EPSILON);
    sc[0][e] = gx * ninv;
    sc[1][e] = gy * ninv;
    if (FZ != 0) {
      const int j = j0 + b;
      if (i <= nx + 1 && j <= ny + 1) {
        const size_t ci = I_DVOF(i, j, 1);
        if (FZ == 1) {
          fz.corner[ci] = norm < fz.floor * rdx ? 1 : 0;
          fz.dvof[2 * ci] = sc[0][e];
          fz.dvof[2 * ci + 1] = sc[1][e];
        } else if (fz.corner[ci]) {
          sc[0][e] = fz.dvof[2 * ci];
          sc[1][e] = fz.dvof[2 * ci + 1];
        }
      }
    }
#else
    const double a0 = SV(a, b, c), b0 = SV(a + 1, b, c), c0 = SV(a, b + 1, c), d0 = SV(a + 1, b + 1, c);
    const double a1 = SV(a, b, c + 1), b1 = SV(a + 1, b, c + 1), c1 = SV(a, b + 1, c + 1), d1 = SV(a + 1, b + 1, c + 1);
    const double sx0 = -a0 + b0 - c0 + d0, sx1 = -a1 + b1 - c1 + d1;
    const double sy0 = -a0 - b0 + c0 + d0, sy1 = -a1 - b1 + c1 + d1;
    const double sz0 = a0 + b0 + c0 + d0, sz1 = a1 + b1 + c1 + d1;
    const double gx = rdx * (sx0 + sx1);
    const double gy = g.rdy * (sy0 + sy1);
    const double gz = g.rdz * (-sz0 + sz1);
    const double norm = sqrt(gx * gx + gy * gy + gz * gz);
    const double ninv = 1. / fmax(norm, DBL_EPSILON);
    sc[0][e] = gx * ninv;
    sc[1][e] = gy * ninv;
    sc[2][e] = gz * ninv;
#endif
#undef SV
  }
  __syncthreads();
  const int a = tid % kIX, b = (tid / kIX) % kIY;
  const int i = i0 + a, j = j0 + b;
#if NDIMS == 3
  const int c = tid / (kIX * kIY);
  const int k = k0 + c;
  const bool own = i <= nx && j <= ny && k <= nz;
#else
  const int c = 0;
  const int k = 1;
  const bool own = i <= nx && j <= ny;
#endif
#define SC(D, A, B, C) sc[D][(A) + CX * ((B) + CY * (C))]
  if (own) {
    const double dx = g.dxf[i - 1];
    const double rdx = g.xt[kXT * i + XT_RXF];   // 1 / DXF
    double cv;
#if NDIMS == 2
    {
      const double dnxdx = rdx * (-SC(0, a, b, 0) + SC(0, a + 1, b, 0) - SC(0, a, b + 1, 0) + SC(0, a + 1, b + 1, 0));
      const double dnydy = g.rdy * (-SC(1, a, b, 0) - SC(1, a + 1, b, 0) + SC(1, a, b + 1, 0) + SC(1, a + 1, b + 1, 0));
      cv = 0.5 * (-dnxdx - dnydy);
    }
#else
    {
      const double dnxdx = rdx * ((-SC(0, a, b, c) + SC(0, a + 1, b, c) - SC(0, a, b + 1, c) + SC(0, a + 1, b + 1, c))
                                      + (-SC(0, a, b, c + 1) + SC(0, a + 1, b, c + 1) - SC(0, a, b + 1, c + 1) + SC(0, a + 1, b + 1, c + 1)));
      const double dnydy = g.rdy * ((-SC(1, a, b, c) - SC(1, a + 1, b, c) + SC(1, a, b + 1, c) + SC(1, a + 1, b + 1, c))
                                        + (-SC(1, a, b, c + 1) - SC(1, a + 1, b, c + 1) + SC(1, a, b + 1, c + 1) + SC(1, a + 1, b + 1, c + 1)));
      const double dnzdz = g.rdz * (-(SC(2, a, b, c) + SC(2, a + 1, b, c) + SC(2, a, b + 1, c) + SC(2, a + 1, b + 1, c))
                                        + (SC(2, a, b, c + 1) + SC(2, a + 1, b, c + 1) + SC(2, a, b + 1, c + 1) + SC(2, a + 1, b + 1, c + 1)));
      cv = 0.25 * (-dnxdx - dnydy - dnzdz);
    }
#endif
    store_periodic(curv, i, nx + 2, ny + 2, j, PLANE1(k), ny, nz, 1, cv);
    const double lvof = sv[(a + 1) + VX * ((b + 1) + VY * (NDIMS == 3 ? c + 1 : 0))];
    bool iface;
    if (FZ == 2) {
      iface = (fz.cell[I_C1(i, j, k)] & 0x80u) != 0;
    } else {
      iface = !(lvof < kVofMin || 1. - kVofMin < lvof);
      if (FZ == 1) fz.cell[I_C1(i, j, k)] = iface ? 0x80u : 0u;
    }
    if (iface) {
      double nv[NDIMS];
#if NDIMS == 2
      {
        double nxv = SC(0, a, b, 0) + SC(0, a + 1, b, 0) + SC(0, a, b + 1, 0) + SC(0, a + 1, b + 1, 0);
        double nyv = SC(1, a, b, 0) + SC(1, a + 1, b, 0) + SC(1, a, b + 1, 0) + SC(1, a + 1, b + 1, 0);
        nxv /= dx;
        nyv /= g.dy;
        const double norm = sqrt(nxv * nxv + nyv * nyv);
        const double ninv = 1. / fmax(norm, DBL_EPSILON);
        nv[0] = nxv * ninv;
        nv[1] = nyv * ninv;
      }
#else
      for (int d = 0; d < 3; d++) {
        nv[d] = (SC(d, a, b, c) + SC(d, a + 1, b, c) + SC(d, a, b + 1, c) + SC(d, a + 1, b + 1, c))
              + (SC(d, a, b, c + 1) + SC(d, a + 1, b, c + 1) + SC(d, a, b + 1, c + 1) + SC(d, a + 1, b + 1, c + 1));
      }
      nv[0] /= dx;
      nv[1] /= g.dy;
      nv[2] /= g.dz;
      {
        const double norm = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        const double ninv = 1. / fmax(norm, DBL_EPSILON);
        nv[0] *= ninv;
        nv[1] *= ninv;
        nv[2] *= ninv;
      }
#endif
      const int slot = atomicAdd(&ln, 1);   // at most one entry per lane
      lv[0][slot] = lvof;
#pragma unroll
      for (int d = 0; d < NDIMS; d++) lv[1 + d][slot] = nv[d];
      lc[slot] = tid;
    }
  }
#undef SC
  __syncthreads();
  const int nl = ln;
  for (int e = tid; e < nl; e += kIThreads) {
    const int t = lc[e];
    const int ia = i0 + t % kIX, jb = j0 + (t / kIX) % kIY;
#if NDIMS == 3
    const int kc = k0 + t / (kIX * kIY);
#else
    const int kc = 1;
#endif
    double nv[NDIMS + 1];
#pragma unroll
    for (int d = 0; d < NDIMS; d++) nv[d] = lv[1 + d][e];
    {
      int it = FZ == 2 ? (int)(fz.cell[I_C1(ia, jb, kc)] & 15u) : 0;
      nv[NDIMS] = intercept<FZ>(lv[0][e], nv, &it);
      if (FZ == 1) fz.cell[I_C1(ia, jb, kc)] = (unsigned char)(0x80u | (unsigned)it);
    }
    // NORMAL rows are 0..ny+1 like cell arrays but x runs 1..nx without padding
    for (int d = 0; d <= NDIMS; d++) {
      // component d of each image: the component array is interleaved, so
      // store through a strided view
      double * base = normal + d;
      const int n0 = nx * (NDIMS + 1);
      store_periodic(base, (ia - 1) * (NDIMS + 1), n0, ny + 2, jb, PLANE1(kc), ny, nz, 1, nv[d]);
    }
  }
}

__device__ __forceinline__ void k_iface_fused_body(const VBlock & vb, Grid g, const double * __restrict__ vof,
                                                   double * __restrict__ normal, double * __restrict__ curv) {
  k_iface_fused_fz_body<0>(vb, g, vof, normal, curv, Freeze{});
}

__global__ __launch_bounds__(kIThreads) void k_iface_fused(Grid g, const double * __restrict__ vof,
                                                           double * __restrict__ normal,
                                                           double * __restrict__ curv) {
  k_iface_fused_body(hw_block(), g, vof, normal, curv);
}

#if NDIMS == 2
template <int FZ>
__global__ __launch_bounds__(kIThreads) void k_iface_fused_fz(Grid g, const double * __restrict__ vof,
                                                              double * __restrict__ normal,
                                                              double * __restrict__ curv, Freeze fz) {
  k_iface_fused_fz_body<FZ>(hw_block(), g, vof, normal, curv, fz);
}
#endif

// Work-group barrier that waits for LDS operations only: global loads issued
// before it stay in flight (a __syncthreads fence would wait for them).  The
// release / acquire fences are scoped to the work-group and to LDS ("local"),
// so the compiler orders the LDS accesses around the barrier itself and
// waits on the LDS counter only (s_waitcnt lgkmcnt(0); s_barrier).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

#if NDIMS == 3
// k_iface_fused marching in z: a 32 x 8 tile of (x, y) walks KZ planes.
// Vof planes live in a 3-plane LDS ring, corner normals in a 2-plane ring;
// the vof plane two ahead is loaded into registers while the current plane
// is computed.  For cell plane k: the registers (plane k+1) go to LDS, the
// loads of plane k+2 are issued, the corner plane between k and k+1 is
// formed (the one between k-1 and k is kept from the previous plane), then
// the cells of plane k.  Same arithmetic as k_iface_fused.  The intercepts
// (a Newton solve at each interface cell) are not solved in the march, where
// a wavefront would run them with only its few interface lanes active: the
// cells' vof and normals go to a list in LDS, which the work-group's lanes
// solve one cell each after the last plane.  THINC fringes make interfaces
// several cells thick, so a work-group can meet more interface cells than
// the list holds: those wait in their own entries of the output (normal and
// vof), and after the list their (a, b, k - kb) are compacted into the
// corner ring, free by then, and solved one cell per lane as well.  (Solving
// them in the march holds the Newton solve's registers across the whole
// plane loop and costs a wave of occupancy.)
constexpr int kMX = 32, kMY = 8;
constexpr int kMThreads = kMX * kMY;
constexpr int kMList = 256;
template <int KZ>
__device__ __forceinline__ void k_iface_march_body(const VBlock & vb, Grid g, const double * __restrict__ vof,
                                                   double * __restrict__ normal, double * __restrict__ curv) {
  constexpr int VX = kMX + 2, VY = kMY + 2, VP = VX * VY;
  constexpr int CX = kMX + 1, CY = kMY + 1, CP = CX * CY;
  constexpr int PF = (VP + kMThreads - 1) / kMThreads;   // prefetched values per thread
  __shared__ double sv[3][VP];
  __shared__ double sc[2][3][CP];
  __shared__ double lv[4][kMList];   // listed interface cells: vof, normal x, y, z
  __shared__ int lc[kMList];         // their (a, b, k - kb)
  __shared__ int ln;
  __shared__ int lnov;               // cells beyond the list
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int i0 = 1 + (int)vb.x * kMX, j0 = 1 + (int)vb.y * kMY, kb = 1 + (int)vb.z * KZ;
  const int ke = min(kb + KZ - 1, nz);
  const int tid = threadIdx.x;
  double pf[PF];
  // vof plane kk (cells i0-1 .. i0+kMX, j0-1 .. j0+kMY) into registers
  auto fetch = [&](int kk) {
#pragma unroll
    for (int r = 0; r < PF; r++) {
      const int e = tid + r * kMThreads;
      const int i = i0 - 1 + e % VX, j = j0 - 1 + e / VX;
      pf[r] = e < VP && i <= nx + 1 && j <= ny + 2 ? vof[I_VOF(i, j, kk)] : 0.;
    }
  };
  auto put = [&](int kk) {
#pragma unroll
    for (int r = 0; r < PF; r++) {
      const int e = tid + r * kMThreads;
      if (e < VP) sv[kk % 3][e] = pf[r];
    }
  };
  // corner normals between cell planes kk-1 and kk (compute_gradient) into
  // ring slot kk % 2
  auto corners = [&](int kk) {
    const double * lo = sv[(kk + 2) % 3];   // plane kk - 1
    const double * hi = sv[kk % 3];
    double * o0 = sc[kk & 1][0];
    double * o1 = sc[kk & 1][1];
    double * o2 = sc[kk & 1][2];
    for (int e = tid; e < CP; e += kMThreads) {
      const int a = e % CX, b = e / CX;
      const int i = i0 + a;
      const double rdx = g.xt[kXT * min(i, nx + 1) + XT_RXC];   // 1 / DXC
      const int p = a + VX * b;
      const double a0 = lo[p], b0 = lo[p + 1], c0 = lo[p + VX], d0 = lo[p + VX + 1];
      const double a1 = hi[p], b1 = hi[p + 1], c1 = hi[p + VX], d1 = hi[p + VX + 1];
      const double sx0 = -a0 + b0 - c0 + d0, sx1 = -a1 + b1 - c1 + d1;
      const double sy0 = -a0 - b0 + c0 + d0, sy1 = -a1 - b1 + c1 + d1;
      const double sz0 = a0 + b0 + c0 + d0, sz1 = a1 + b1 + c1 + d1;
      const double gx = rdx * (sx0 + sx1);
      const double gy = g.rdy * (sy0 + sy1);
      const double gz = g.rdz * (-sz0 + sz1);
      const double norm = sqrt(gx * gx + gy * gy + gz * gz);
      const double ninv = 1. / fmax(norm, DBL_EPSILON);
      o0[e] = gx * ninv;
      o1[e] = gy * ninv;
      o2[e] = gz * ninv;
    }
  };
  if (tid == 0) {
    ln = 0;
    lnov = 0;
  }
  fetch(kb - 1);
  put(kb - 1);
  fetch(kb);
  put(kb);
  fetch(kb + 1);
  lds_barrier();
  corners(kb);
  const int a = tid % kMX, b = tid / kMX;
  const int i = i0 + a, j = j0 + b;
  const bool own = i <= nx && j <= ny;
  const double dx = own ? g.dxf[i - 1] : 1.;
  const double rdx = own ? g.xt[kXT * i + XT_RXF] : 0.;   // 1 / DXF
  // first of the 4 entries of cell (ii, jj, kk) in `normal`
  auto park = [&](int ii, int jj, int kk) {
    return (size_t)(ii - 1) * 4 + (size_t)nx * 4 * ((size_t)jj + (size_t)(ny + 2) * (size_t)kk);
  };
  unsigned ovm = 0u;   // planes (bit k - kb) whose cell of this lane waits outside the list
  for (int k = kb; k <= ke; k++) {
    put(k + 1);
    if (k + 2 <= ke + 1) fetch(k + 2);
    lds_barrier();
    corners(k + 1);
    lds_barrier();
    if (own) {
      const double * L0 = sc[k & 1][0];
      const double * L1 = sc[k & 1][1];
      const double * L2 = sc[k & 1][2];
      const double * H0 = sc[(k + 1) & 1][0];
      const double * H1 = sc[(k + 1) & 1][1];
      const double * H2 = sc[(k + 1) & 1][2];
      const int q = a + CX * b;
      double cv;
      {
        const double dnxdx = rdx * ((-L0[q] + L0[q + 1] - L0[q + CX] + L0[q + CX + 1])
                                    + (-H0[q] + H0[q + 1] - H0[q + CX] + H0[q + CX + 1]));
        const double dnydy = g.rdy * ((-L1[q] - L1[q + 1] + L1[q + CX] + L1[q + CX + 1])
                                      + (-H1[q] - H1[q + 1] + H1[q + CX] + H1[q + CX + 1]));
        const double dnzdz = g.rdz * (-(L2[q] + L2[q + 1] + L2[q + CX] + L2[q + CX + 1])
                                      + (H2[q] + H2[q + 1] + H2[q + CX] + H2[q + CX + 1]));
        cv = 0.25 * (-dnxdx - dnydy - dnzdz);
      }
      store_periodic(curv, i, nx + 2, ny + 2, j, PLANE1(k), ny, nz, 1, cv);
      const double lvof = sv[k % 3][(a + 1) + VX * (b + 1)];
      if (!(lvof < kVofMin || 1. - kVofMin < lvof)) {
        double nv[4];
        const double * Ld[3] = {L0, L1, L2};
        const double * Hd[3] = {H0, H1, H2};
        for (int d = 0; d < 3; d++) {
          nv[d] = (Ld[d][q] + Ld[d][q + 1] + Ld[d][q + CX] + Ld[d][q + CX + 1])
                + (Hd[d][q] + Hd[d][q + 1] + Hd[d][q + CX] + Hd[d][q + CX + 1]);
        }
        nv[0] /= dx;
        nv[1] /= g.dy;
        nv[2] /= g.dz;
        {
          const double norm = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
          const double ninv = 1. / fmax(norm, DBL_EPSILON);
          nv[0] *= ninv;
          nv[1] *= ninv;
          nv[2] *= ninv;
        }
        const int slot = atomicAdd(&ln, 1);
        if (slot < kMList) {
          lv[0][slot] = lvof;
          lv[1][slot] = nv[0];
          lv[2][slot] = nv[1];
          lv[3][slot] = nv[2];
          lc[slot] = a | (b << 8) | ((k - kb) << 16);
        } else {
          // list full: the cell's normal and vof wait in its own entry of
          // `normal` (x, y, z, intercept) and bit k - kb of ovm
          double * ent = normal + park(i, j, k);
          ent[0] = nv[0];
          ent[1] = nv[1];
          ent[2] = nv[2];
          ent[3] = lvof;
          ovm |= 1u << (k - kb);
        }
      }
    }
  }
  lds_barrier();
  const int nl = min(ln, kMList);
  for (int e = tid; e < nl; e += kMThreads) {
    const int c = lc[e];
    const int ci = i0 + (c & 255), cj = j0 + ((c >> 8) & 255), ck = kb + (c >> 16);
    double nv[4] = {lv[1][e], lv[2][e], lv[3][e], 0.};
    nv[3] = intercept(lv[0][e], nv);
    for (int d = 0; d <= 3; d++) {
      double * base = normal + d;
      store_periodic(base, (ci - 1) * 4, nx * 4, ny + 2, cj, PLANE1(ck), ny, nz, 1, nv[d]);
    }
  }
  // the cells beyond the list: lane offsets by atomic add, then windows of
  // kOv cells compacted into the corner ring and solved a lane each
  constexpr int kOv = (int)(sizeof(sc) / sizeof(int));
  const int cnt = __builtin_popcount(ovm);
  const int off = cnt > 0 ? atomicAdd(&lnov, cnt) : 0;
  __syncthreads();   // the parked entries (global) and lnov (LDS)
  const int tot = lnov;
  int * ovl = reinterpret_cast<int *>(&sc[0][0][0]);
  for (int w = 0; w < tot; w += kOv) {
    unsigned m = ovm;
    for (int e = off; m != 0u; e++) {
      const int kk = __builtin_ctz(m);
      m &= m - 1u;
      if (e >= w && e < w + kOv) ovl[e - w] = a | (b << 8) | (kk << 16);
    }
    lds_barrier();
    const int n = min(tot - w, kOv);
    for (int x = tid; x < n; x += kMThreads) {
      const int c = ovl[x];
      const int ci = i0 + (c & 255), cj = j0 + ((c >> 8) & 255), ck = kb + (c >> 16);
      const double * ent = normal + park(ci, cj, ck);
      double nv[4] = {ent[0], ent[1], ent[2], 0.};
      nv[3] = intercept(ent[3], nv);
      for (int d = 0; d <= 3; d++) {
        double * base = normal + d;
        store_periodic(base, (ci - 1) * 4, nx * 4, ny + 2, cj, PLANE1(ck), ny, nz, 1, nv[d]);
      }
    }
    lds_barrier();
  }
}

template <int KZ>
__global__ __launch_bounds__(kMThreads) void k_iface_march(Grid g, const double * __restrict__ vof,
                                                           double * __restrict__ normal, double * __restrict__ curv) {
  k_iface_march_body<KZ>(hw_block(), g, vof, normal, curv);
}
#endif

// interface_update_vof fused: the 2 x NDIMS face fluxes of each cell
// (compute_flux_*), the flux divergence and the RK update, written to the
// other VOF buffer with its halo images (walls stay zero).

// face_flux's THINC integral with the face direction D chosen at run time
// (the same quadrature points and operation order as face_flux<D>)
__device__ inline double thinc_face(const double * n, int D, double c) {
  double flux = 0.;
#if NDIMS == 2
  for (int q = 0; q < NGAUSS; q++) {
    const double pt[2] = {D == 0 ? c : kGaussP[q], D == 0 ? kGaussP[q] : c};
    flux += kGaussW[q] * indicator(n, pt);
  }
#else
  for (int q = 0; q < NGAUSS; q++) {
    double sr = 0.;
    for (int r = 0; r < NGAUSS; r++) {
      const double pt[3] = {D == 0 ? c : kGaussP[q], D == 0 ? kGaussP[q] : D == 1 ? c : kGaussP[r],
                            D == 2 ? c : kGaussP[r]};
      sr += kGaussW[r] * indicator(n, pt);
    }
    flux += kGaussW[q] * sr;
  }
#endif
  return flux;
}

// k_vof_fused without decision record / replay: the THINC integrals of the
// work-group's interface faces (upwind cell holding an interface) are not
// formed by each cell's lane, where a wavefront runs them with its few
// interface lanes active, but one face per lane from a list in LDS.  Every
// face of the work-group has a slot (lane * NF + face) holding its flux, or
// for an interface face its velocity until the listed pass replaces it by
// velocity * integral.  Fluxes and their sum are those of face_flux, in the
// same order.  At most 256 lanes (launch bounds), so the slots never run out.
constexpr int kVLanes = 256;
__device__ __forceinline__ void vof_fused_listed(const VBlock & vb, Grid g, const double * __restrict__ ux,
                                                 const double * __restrict__ uy, const double * __restrict__ uz,
                                                 const double * __restrict__ vof, const double * __restrict__ normal,
                                                 double * __restrict__ srca, const double * __restrict__ srcb,
                                                 double * __restrict__ vof_new, double dta, double dtb, int use_b,
                                                 int keep_a) {
  constexpr int NC = NDIMS + 1;
  constexpr int NF = 2 * NDIMS;   // x lower, x upper, y lower, y upper (, z lower, z upper)
  __shared__ double fv[kVLanes * NF];           // flux (velocity of a listed face before the listed pass)
  __shared__ unsigned fi[kVLanes * NF];         // listed face: the upwind cell's normal (entry of I_NRM)
  __shared__ unsigned char fd[kVLanes * NF];    // listed face: 2 D + (upwind cell is the upper one)
  __shared__ unsigned short fl[kVLanes * NF];   // the listed faces' slots
  __shared__ int fn;
  const int nx = g.nx, ny = g.ny;
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nt = blockDim.x * blockDim.y;
  if (tid == 0) fn = 0;
  __syncthreads();
  THREAD_IJK(1, 1, 1)
  const bool own = i <= nx && j <= ny;
#if NDIMS == 3
  const int nz = g.nz;
#else
  const int nz = 1;
  (void)uz;
#endif
  double vc = 0., sb = 0.;
  size_t is = 0;
  if (own) {
    is = I_SC(i, j, k);
    const double uxm = ux[I_UX(i, j, k)], uxp = ux[I_UX(i + 1, j, k)];
    const double uym = uy[I_C1(i, j, k)], uyp = uy[I_C1(i, j + 1, k)];
    vc = vof[I_VOF(i, j, k)];
    const double vxm = vof[I_VOF(i - 1, j, k)], vxp = vof[I_VOF(i + 1, j, k)];
    const double vym = vof[I_VOF(i, j - 1, k)], vyp = vof[I_VOF(i, j + 1, k)];
    sb = use_b ? srcb[is] : 0.;
#if NDIMS == 3
    const double uzm = uz[I_C1(i, j, k)], uzp = uz[I_C1(i, j, k + 1)];
    const double vzm = vof[I_VOF(i, j, k - 1)], vzp = vof[I_VOF(i, j, k + 1)];
#endif
    const unsigned cc = (unsigned)I_NRM(i, j, k);
    // face f between its lower cell (vof vm, normal nm) and upper cell (vp, np)
    auto face = [&](int f, double vel, double vm, double vp, unsigned nm, unsigned np) {
      const int sl = tid * NF + f;
      if ((f == 0 && i == 1) || (f == 1 && i == nx)) {   // wall faces carry no flux
        fv[sl] = 0.;
        return;
      }
      const bool up = vel < 0.;
      const double lv = up ? vp : vm;
      if (lv < kVofMin || 1. - kVofMin < lv) {
        fv[sl] = vel * lv;
        return;
      }
      fv[sl] = vel;
      fi[sl] = up ? np : nm;
      fd[sl] = (unsigned char)(2 * (f >> 1) + (up ? 1 : 0));
      fl[atomicAdd(&fn, 1)] = (unsigned short)sl;
    };
    face(0, uxm, vxm, vc, (unsigned)I_NRM(i - 1, j, k), cc);
    face(1, uxp, vc, vxp, cc, (unsigned)I_NRM(i + 1, j, k));
    face(2, uym, vym, vc, (unsigned)I_NRM(i, j - 1, k), cc);
    face(3, uyp, vc, vyp, cc, (unsigned)I_NRM(i, j + 1, k));
#if NDIMS == 3
    face(4, uzm, vzm, vc, (unsigned)I_NRM(i, j, k - 1), cc);
    face(5, uzp, vc, vzp, cc, (unsigned)I_NRM(i, j, k + 1));
#endif
  }
  __syncthreads();
  const int n = fn;
  for (int e = tid; e < n; e += nt) {
    const int sl = fl[e];
    const int d = fd[sl];
    fv[sl] = fv[sl] * thinc_face(normal + NC * (size_t)fi[sl], d >> 1, (d & 1) ? -0.5 : +0.5);
  }
  __syncthreads();
  if (!own) return;
  const double * f = fv + tid * NF;
  double s = 0.;
  s += g.xt[kXT * i + XT_RXF] * (f[0] - f[1]);   // 1 / DXF(i)
  s += g.rdy * (f[2] - f[3]);
#if NDIMS == 3
  s += g.rdz * (f[4] - f[5]);
#endif
  if (keep_a) st_stream(&srca[is], s);
  double v = vc;
  v += dta * s;
  if (use_b) v += dtb * sb;
  store_periodic(vof_new, i, nx + 2, ny + 4, j + 1, PLANE2(k), ny, nz, 2, v);
}

// The same with each cell's lane forming its own faces' integrals, as the
// 2D adjoint's decision record (FZ = 1) and replay (FZ = 2) need.
template <int FZ>
__device__ __forceinline__ void k_vof_fused_fz_body(const VBlock & vb, Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                            const double * __restrict__ uz, const double * __restrict__ vof,
                            const double * __restrict__ normal, double * __restrict__ srca,
                            const double * __restrict__ srcb, double * __restrict__ vof_new,
                            double dta, double dtb, int use_b, int keep_a, const Freeze & fz) {
  static_assert(FZ == 0 || NDIMS == 2, "decision replay is 2D");
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
#if NDIMS == 3
  const int nz = g.nz;
#else
  const int nz = 1;
  (void)uz;
#endif
  // the six face velocities, the seven vof values and srcb are loaded before
  // any flux is formed, so they are in flight together (the upwind choices
  // branch; the walls' velocities at i = 1 and the face past i = nx are loaded
  // and unused); only interface cells read normals, inside face_flux
  constexpr int NC = NDIMS + 1;
  const size_t is = I_SC(i, j, k);
  const double uxm = ux[I_UX(i, j, k)], uxp = ux[I_UX(i + 1, j, k)];
  const double uym = uy[I_C1(i, j, k)], uyp = uy[I_C1(i, j + 1, k)];
  const double vc = vof[I_VOF(i, j, k)];
  const double vxm = vof[I_VOF(i - 1, j, k)], vxp = vof[I_VOF(i + 1, j, k)];
  const double vym = vof[I_VOF(i, j - 1, k)], vyp = vof[I_VOF(i, j + 1, k)];
  const double sb = use_b ? srcb[is] : 0.;
#if NDIMS == 3
  const double uzm = uz[I_C1(i, j, k)], uzp = uz[I_C1(i, j, k + 1)];
  const double vzm = vof[I_VOF(i, j, k - 1)], vzp = vof[I_VOF(i, j, k + 1)];
#endif
  const double * nc = normal + NC * I_NRM(i, j, k);
  unsigned char * dxm = FZ ? fz.fx + I_UX(i, j, k) : nullptr;
  unsigned char * dxp = FZ ? fz.fx + I_UX(i + 1, j, k) : nullptr;
  unsigned char * dym = FZ ? fz.fy + I_C1(i, j, k) : nullptr;
  unsigned char * dyp = FZ ? fz.fy + I_C1(i, j + 1, k) : nullptr;
  const double fxm = i == 1 ? 0. : face_flux<0, FZ>(uxm, vxm, vc, normal + NC * I_NRM(i - 1, j, k), nc, dxm);
  const double fxp = i == nx ? 0. : face_flux<0, FZ>(uxp, vc, vxp, nc, normal + NC * I_NRM(i + 1, j, k), dxp);
  const double fym = face_flux<1, FZ>(uym, vym, vc, normal + NC * I_NRM(i, j - 1, k), nc, dym);
  const double fyp = face_flux<1, FZ>(uyp, vc, vyp, nc, normal + NC * I_NRM(i, j + 1, k), dyp);
  double s = 0.;
  s += g.xt[kXT * i + XT_RXF] * (fxm - fxp);   // 1 / DXF(i)
  s += g.rdy * (fym - fyp);
#if NDIMS == 3
  const double fzm = face_flux<2>(uzm, vzm, vc, normal + NC * I_NRM(i, j, k - 1), nc);
  const double fzp = face_flux<2>(uzp, vc, vzp, nc, normal + NC * I_NRM(i, j, k + 1));
  s += g.rdz * (fzm - fzp);
#endif
  if (keep_a) st_stream(&srca[is], s);
  double v = vc;
  v += dta * s;
  if (use_b) v += dtb * sb;
  store_periodic(vof_new, i, nx + 2, ny + 4, j + 1, PLANE2(k), ny, nz, 2, v);
}

__device__ __forceinline__ void k_vof_fused_body(const VBlock & vb, Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                            const double * __restrict__ uz, const double * __restrict__ vof,
                            const double * __restrict__ normal, double * __restrict__ srca,
                            const double * __restrict__ srcb, double * __restrict__ vof_new,
                            double dta, double dtb, int use_b, int keep_a) {
  vof_fused_listed(vb, g, ux, uy, uz, vof, normal, srca, srcb, vof_new, dta, dtb, use_b, keep_a);
}

__global__ __launch_bounds__(256) void k_vof_fused(Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                            const double * __restrict__ uz, const double * __restrict__ vof,
                            const double * __restrict__ normal, double * __restrict__ srca,
                            const double * __restrict__ srcb, double * __restrict__ vof_new,
                            double dta, double dtb, int use_b, int keep_a) {
  k_vof_fused_body(hw_block(), g, ux, uy, uz, vof, normal, srca, srcb, vof_new, dta, dtb, use_b, keep_a);
}

#if NDIMS == 2
template <int FZ>
__global__ __launch_bounds__(256) void k_vof_fused_fz(Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                            const double * __restrict__ vof, const double * __restrict__ normal,
                            double * __restrict__ srca, const double * __restrict__ srcb,
                            double * __restrict__ vof_new, double dta, double dtb, int use_b, int keep_a, Freeze fz) {
  k_vof_fused_fz_body<FZ>(hw_block(), g, ux, uy, nullptr, vof, normal, srca, srcb, vof_new, dta, dtb, use_b, keep_a,
                          fz);
}
#endif

// fluid_predict_field sources of all fields at once, with the CSF force
// (force.c:11-66) formed in place from vof and curvature; writes srca and the
// increment of every DOF
struct FusedRhs {
  RhsArgs base;
  const double * vof;
  const double * curv;
  double tension;
  const double * lapx[4];
  double dif[4];
  double * srca[4];
  const double * srcb[4];
  double * delta[4];
  // use_b = 0 (first RK stage, beta = 0): srcb is not read and its term is
  // b dt * 0.  keep_a = 0 (last RK stage): srca is not stored, as no later
  // stage reads it (the next step's first stage has beta = 0).
  int use_b, keep_a;
};

__device__ __forceinline__ void k_rhs_fused_body(const VBlock & vb, Grid g, FusedRhs f) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
  const double * vof = f.vof;
  const double * curv = f.curv;
  const double v = vof[I_VOF(i, j, k)];
  const double cv = curv[I_C1(i, j, k)];
  RhsArgs a = f.base;
  const double ca = a.ca, cb = a.cb, cg = a.cg;
  // every source first, every store last: no store sits between the loads
  // of the four stencils, so shared neighbours are loaded once
  double sa0 = 0., sg0 = 0., sa1, sg1, sa3, sg3;
  if (i >= 2) {
    const double grad = g.xt[kXT * i + XT_RXC] * (-vof[I_VOF(i - 1, j, k)] + v);   // 1/DXC(i) * ...
    const double kappa = 0.5 * (curv[I_C1(i - 1, j, k)] + cv);
    a.lapx = f.lapx[0];
    a.dif = f.dif[0];
    rhs_ux_at(g, a, i, j, k, f.tension * grad * kappa, sa0, sg0);
  }
  {
    const double grad = g.rdy * (-vof[I_VOF(i, j - 1, k)] + v);
    const double kappa = 0.5 * (curv[I_C1(i, j - 1, k)] + cv);
    a.lapx = f.lapx[1];
    a.dif = f.dif[1];
    rhs_uy_at(g, a, i, j, k, f.tension * grad * kappa, sa1, sg1);
  }
#if NDIMS == 3
  double sa2, sg2;
  {
    const double grad = g.rdz * (-vof[I_VOF(i, j, k - 1)] + v);
    const double kappa = 0.5 * (curv[I_C1(i, j, k - 1)] + cv);
    a.lapx = f.lapx[2];
    a.dif = f.dif[2];
    rhs_uz_at(g, a, i, j, k, f.tension * grad * kappa, sa2, sg2);
  }
#endif
  {
    a.lapx = f.lapx[3];
    a.dif = f.dif[3];
    rhs_t_at(g, a, i, j, k, 0., sa3, sg3);
  }
  const size_t is = I_SC(i, j, k);
  const bool ub = f.use_b, ka = f.keep_a;
  if (i >= 2) {
    const size_t ix = I_SX(i, j, k);
    if (ka) f.srca[0][ix] = sa0;
    f.delta[0][ix] = +ca * sa0 + cb * (ub ? f.srcb[0][ix] : 0.) + cg * sg0;
  }
  if (ka) f.srca[1][is] = sa1;
  f.delta[1][is] = +ca * sa1 + cb * (ub ? f.srcb[1][is] : 0.) + cg * sg1;
#if NDIMS == 3
  if (ka) f.srca[2][is] = sa2;
  f.delta[2][is] = +ca * sa2 + cb * (ub ? f.srcb[2][is] : 0.) + cg * sg2;
#endif
  if (ka) f.srca[3][is] = sa3;
  f.delta[3][is] = +ca * sa3 + cb * (ub ? f.srcb[3][is] : 0.) + cg * sg3;
}

__global__ __launch_bounds__(256) void k_rhs_fused(Grid g, FusedRhs f) {
  k_rhs_fused_body(hw_block(), g, f);
}

// k_rhs_fused with each thread marching R consecutive rows j of one (i, k)
// column: the rows j-1, j, j+1 of the stencils' centre columns (and the
// other columns one row apart) stay in registers from one DOF to the next,
// so each DOF loads one new row per column (35 loads instead of ~70) and
// the next row's loads are independent of the current row's arithmetic.
// Values and arithmetic are those of k_rhs_fused (the same rhs_*_v).
//
// rhs_rows marches rows jb .. jb+nr-1 of column (i, k): it stores the
// explicit sources (keep_a) and hands each row's increments
// ca sa + cb sb + cg sg of ux, uy, uz, t to sink(r, d0, d1, d2, d3) (d0 is
// meaningful only for i >= 2, the ux DOFs; d2 only in 3D).
template <int R, class Sink>
__device__ __forceinline__ void rhs_rows(const Grid & g, const FusedRhs & f, int i, int jb, int k, int nr,
                                         Sink && sink) {
  const int nx = g.nx, ny = g.ny;
  const ptrdiff_t sx = nx + 1, sc = nx + 2;   // row strides of ux and cell arrays
#if NDIMS == 3
  const ptrdiff_t pc = sc * (ny + 2), pv = sc * (ny + 4);
#endif
  const RhsArgs & b = f.base;
  const bool has_ux = i >= 2;
  // column pointers at row j (advanced by one row per DOF)
  const double * ux = b.ux + I_UX(i, jb, k);
  const double * uy = b.uy + I_C1(i, jb, k);
  const double * t = b.t + I_C1(i, jb, k);
  const double * pp = b.p + I_C1(i, jb, k);
  const double * vf = f.vof + I_VOF(i, jb, k);
  const double * cu = f.curv + I_C1(i, jb, k);
#if NDIMS == 3
  const ptrdiff_t px = sx * (ny + 2);
  const double * uz = b.uz + I_C1(i, jb, k);
#endif
  // rows carried from the previous DOF: centre columns at j-1 and j, and the
  // off-centre columns read at two rows
  double a0m = ux[-sx], a0c = ux[0], axpm = ux[1 - sx];
  double b0m = uy[-sc], b0c = uy[0], bxmc = uy[-1];
  double t0m = t[-sc], t0c = t[0];
  double p0m = pp[-sc], v0m = vf[-sc], k0m = cu[-sc];
#if NDIMS == 3
  double bzmc = uy[-pc];
  double c0m = uz[-sc], c0c = uz[0], czpm = uz[pc - sc];
#endif
  size_t is = I_SC(i, jb, k);
  size_t ix = has_ux ? I_SX(i, jb, k) : 0;
  for (int r = 0; r < nr; r++) {
    // the new row of every column
    const double a0p = ux[sx], axpc = ux[1];
    const double axm = has_ux ? ux[-1] : 0.;
    const double b0p = uy[sc], bxmp = uy[sc - 1], bxp = uy[1];
    const double t0p = t[sc], txm = t[-1], txp = t[1];
    const double p0c = pp[0], pxm = pp[-1];
    const double v = vf[0], vxm = vf[-1];
    const double cv = cu[0], kxm = cu[-1];
#if NDIMS == 3
    const double azm = ux[-px], azp = ux[px], axpzm = ux[1 - px];
    const double bzmp = uy[sc - pc], bzp = uy[pc];
    const double c0p = uz[sc], cxm = uz[-1], cxp = uz[1], czpc = uz[pc], czm = uz[-pc], cxmzp = uz[pc - 1];
    const double tzm = t[-pc], tzp = t[pc];
    const double pzm = pp[-pc], vzm = vf[-pv], kzm = cu[-pc];
#endif
    double sb0 = 0., sb1 = 0., sb2 = 0., sb3 = 0.;
    if (f.use_b) {
      sb1 = f.srcb[1][is];
      sb3 = f.srcb[3][is];
      if (has_ux) sb0 = f.srcb[0][ix];
#if NDIMS == 3
      sb2 = f.srcb[2][is];
#endif
    }
    (void)sb2;
    RhsArgs a = b;
    double sa0 = 0., sg0 = 0., sa1, sg1, sa3, sg3;
    if (has_ux) {
      UxSt s;
      s.c = a0c; s.xm = axm; s.xp = axpc; s.ym = a0m; s.yp = a0p;
      s.vlm = bxmc; s.vlc = b0c; s.vum = bxmp; s.vuc = b0p;
#if NDIMS == 3
      s.zm = azm; s.zp = azp;
      s.wlm = cxm; s.wlc = c0c; s.wum = cxmzp; s.wuc = czpc;
#else
      s.zm = s.zp = s.wlm = s.wlc = s.wum = s.wuc = 0.;
#endif
      s.pm = pxm; s.pc = p0c; s.tm = txm; s.tc = t0c;
      const double grad = g.xt[kXT * i + XT_RXC] * (-vxm + v);   // 1/DXC(i) * ...
      const double kappa = 0.5 * (kxm + cv);
      a.lapx = f.lapx[0];
      a.dif = f.dif[0];
      rhs_ux_v(g, a, i, s, f.tension * grad * kappa, sa0, sg0);
    }
    {
      UySt s;
      s.c = b0c; s.xm = bxmc; s.xp = bxp; s.ym = b0m; s.yp = b0p;
      s.ulm = a0m; s.ulc = a0c; s.uum = axpm; s.uuc = axpc;
#if NDIMS == 3
      s.zm = bzmc; s.zp = bzp;
      s.wlm = c0m; s.wlc = c0c; s.wum = czpm; s.wuc = czpc;
#else
      s.zm = s.zp = s.wlm = s.wlc = s.wum = s.wuc = 0.;
#endif
      s.pm = p0m; s.pc = p0c;
      const double grad = g.rdy * (-v0m + v);
      const double kappa = 0.5 * (k0m + cv);
      a.lapx = f.lapx[1];
      a.dif = f.dif[1];
      rhs_uy_v(g, a, i, s, f.tension * grad * kappa, sa1, sg1);
    }
    double sa2 = 0., sg2 = 0.;
#if NDIMS == 3
    {
      UzSt s;
      s.c = c0c; s.xm = cxm; s.xp = cxp; s.ym = c0m; s.yp = c0p; s.zm = czm; s.zp = czpc;
      s.ulm = azm; s.ulc = a0c; s.uum = axpzm; s.uuc = axpc;
      s.vlm = bzmc; s.vlc = b0c; s.vum = bzmp; s.vuc = b0p;
      s.pm = pzm; s.pc = p0c;
      const double grad = g.rdz * (-vzm + v);
      const double kappa = 0.5 * (kzm + cv);
      a.lapx = f.lapx[2];
      a.dif = f.dif[2];
      rhs_uz_v(g, a, i, s, f.tension * grad * kappa, sa2, sg2);
    }
#endif
    {
      TSt s;
      s.c = t0c; s.xm = txm; s.xp = txp; s.ym = t0m; s.yp = t0p;
      s.ul = a0c; s.uu = axpc; s.vl = b0c; s.vu = b0p;
#if NDIMS == 3
      s.zm = tzm; s.zp = tzp; s.wl = c0c; s.wu = czpc;
#else
      s.zm = s.zp = s.wl = s.wu = 0.;
#endif
      a.lapx = f.lapx[3];
      a.dif = f.dif[3];
      rhs_t_v(g, a, i, s, sa3, sg3);
    }
    const double ca = b.ca, cb = b.cb, cg = b.cg;
    const bool ka = f.keep_a;
    if (ka) {
      if (has_ux) st_stream(&f.srca[0][ix], sa0);
      st_stream(&f.srca[1][is], sa1);
#if NDIMS == 3
      st_stream(&f.srca[2][is], sa2);
#endif
      st_stream(&f.srca[3][is], sa3);
    }
    sink(r, is, ix, +ca * sa0 + cb * sb0 + cg * sg0, +ca * sa1 + cb * sb1 + cg * sg1,
         +ca * sa2 + cb * sb2 + cg * sg2, +ca * sa3 + cb * sb3 + cg * sg3);
    // shift the carried rows and advance one row
    a0m = a0c; a0c = a0p; axpm = axpc;
    b0m = b0c; b0c = b0p; bxmc = bxmp;
    t0m = t0c; t0c = t0p;
    p0m = p0c; v0m = v; k0m = cv;
    ux += sx; uy += sc; t += sc; pp += sc; vf += sc; cu += sc;
    is += nx;
    ix += nx - 1;
#if NDIMS == 3
    bzmc = bzmp;
    c0m = c0c; c0c = c0p; czpm = czpc;
    uz += sc;
#endif
  }
}

template <int R>
__device__ __forceinline__ void k_rhs_march_body(const VBlock & vb, Grid g, FusedRhs f) {
  const int nx = g.nx, ny = g.ny;
#if NDIMS == 3
  const Tile3 tl = g.bands ? band_tile(vb) : plain_tile(vb);
  const int k = (int)tl.z + 1;
#else
  const Tile3 tl = plain_tile(vb);
  const int k = 1;
#endif
  const int i = (int)(tl.x * blockDim.x + threadIdx.x) + 1;
  const int jb = (int)(tl.y * blockDim.y + threadIdx.y) * R + 1;
  if (i > nx || jb > ny) return;
  const int nr = min(R, ny + 1 - jb);
  const bool has_ux = i >= 2;
  rhs_rows<R>(g, f, i, jb, k, nr, [&](int, size_t is, size_t ix, double d0, double d1, double d2, double d3) {
    if (has_ux) st_stream(&f.delta[0][ix], d0);
    st_stream(&f.delta[1][is], d1);
#if NDIMS == 3
    st_stream(&f.delta[2][is], d2);
#else
    (void)d2;
#endif
    st_stream(&f.delta[3][is], d3);
  });
}

template <int R>
__global__ __launch_bounds__(256) void k_rhs_march(Grid g, FusedRhs f) {
  k_rhs_march_body<R>(hw_block(), g, f);
}

// ---------------------------------------------------------------------------
// The whole predictor of one RK stage in one kernel (predict/main.c:62-79;
// ux.c:306-388 and its uy, uz, t counterparts): a work-group owns complete x
// lines -- R consecutive rows j of one plane k, every x -- and
//   1. marches the sources of ux, uy, uz, t over its rows (rhs_rows, the same
//      arithmetic as k_rhs_march) and keeps the increments in LDS,
//   2. solves (I - prefactor Lx) x = increment along each line there (one
//      wavefront per field, one lane per row: the recurrences of gtsv with
//      the shared factors, as k_lines_lds),
//   3. writes old field + solution into a second buffer of each field (with
//      the x walls and the periodic halo images), because the stencils of the
//      neighbouring rows, run by other work-groups, still read the old field.
// The increments never travel through HBM.  Values are those of
// k_rhs_march followed by k_lines_lds, bit for bit.
constexpr int kPNF = NDIMS + 1;   // fields solved in x: ux, uy, (uz,) t

struct PredLines {
  FusedRhs rhs;              // sources; old fields in rhs.base
  const double * tab[4];     // Thomas table per field kind (0 ux, 1 uy, 2 uz, 3 t)
  double * out[4];           // new buffer per field kind
  int xw;                    // lanes along x per row group: 64, 128 or 256
};

// LDS slot of a field kind: ux 0, uy 1, uz 2 (3D), t NDIMS
__device__ __forceinline__ int pred_kind(int slot) { return slot == NDIMS ? 3 : slot; }
// element `kind` of a four-entry argument array, selected without indexing
// (an indexed kernel argument is copied to scratch memory)
template <class T>
__device__ __forceinline__ T pick4(const T (&a)[4], int kind) {
  return kind == 0 ? a[0] : kind == 1 ? a[1] : kind == 2 ? a[2] : a[3];
}

template <int RT>
__device__ __forceinline__ void k_pred_lines_body(const VBlock & vb, Grid g, const PredLines & P) {
  extern __shared__ double S[];
  const int nx = g.nx, ny = g.ny;
#if NDIMS == 3
  const Tile3 tl = g.bands ? band_tile(vb) : plain_tile(vb);
  const int k = (int)tl.z + 1;
  const int nz = g.nz;
#else
  const Tile3 tl = plain_tile(vb);
  const int k = 1;
  const int nz = 1;
#endif
  const int XW = P.xw;
  const int G = 256 / XW;
  const int R = G * RT;
  const int ld = nx;                    // row stride of the LDS lines
  const size_t fs = (size_t)R * ld;     // LDS stride of one field's rows
  const int j0 = (int)tl.y * R + 1;     // first row of this work-group
  if (j0 > ny) return;                  // uniform over the work-group
  const int nrows = min(R, ny + 1 - j0);
  const FusedRhs & f = P.rhs;
  // 1. increments of the rows into LDS
  {
    const int tx = (int)threadIdx.x % XW, ty = (int)threadIdx.x / XW;
    const int jb = j0 + ty * RT;
    if (jb <= ny) {
      const int nr = min(RT, ny + 1 - jb);
      double * srow = S + (size_t)(ty * RT) * ld;
      for (int c = 0; c < nx; c += XW) {
        const int i = c + tx + 1;
        if (i > nx) break;
        rhs_rows<RT>(g, f, i, jb, k, nr, [&](int r, size_t, size_t, double d0, double d1, double d2, double d3) {
          double * row = srow + (size_t)r * ld;
          if (i >= 2) row[i - 2] = d0;
          row[fs + i - 1] = d1;
#if NDIMS == 3
          row[2 * fs + i - 1] = d2;
          row[3 * fs + i - 1] = d3;
#else
          (void)d2;
          row[2 * fs + i - 1] = d3;
#endif
          // keep the next row's loads behind these stores (as the global
          // stores of k_rhs_march do): hoisting them spills registers
          __asm__ __volatile__("" ::: "memory");
        });
      }
    }
  }
  __syncthreads();
  // 3a. the update is flattened over the LDS lines: element e = tid + 256 q
  // (q < 32; the LDS holds at most 8192 values) is row r of field slot,
  // position m.  The old field values of this thread's elements are loaded
  // before the sweeps, so that their latency passes during them.
  constexpr int kPE = 32;
  const int n1 = ny + 2;
  const float inv_ld = 1.f / (float)ld;
  const int rsh = 31 - __clz(R);   // R is a power of two
  // e / ld through a float product: e < 2^13 and (e + 1/2) / ld stays at
  // least 1/(2 ld) away from an integer, far beyond the product's rounding
  auto decode = [&](int e, int & slot, int & r, int & m) {
    const int line = (int)(((float)e + 0.5f) * inv_ld);
    m = e - line * ld;
    slot = line >> rsh;
    r = line & (R - 1);
  };
  auto old_of = [&](int kind) -> const double * {
    return kind == 0 ? f.base.ux : kind == 1 ? f.base.uy : kind == 2 ? f.base.uz : f.base.t;
  };
  auto row_of = [&](int kind, int j) -> size_t {
    const size_t n0 = kind == 0 ? (size_t)nx + 1 : (size_t)nx + 2;
#if NDIMS == 3
    return n0 * ((size_t)j + (size_t)n1 * (size_t)k);
#else
    return n0 * (size_t)j;
#endif
  };
  double ov[kPE];
#pragma unroll
  for (int q = 0; q < kPE; q++) {
    int slot, r, m;
    decode((int)threadIdx.x + 256 * q, slot, r, m);
    ov[q] = 0.;
    if (slot < kPNF && r < nrows) {
      const int kind = pred_kind(slot);
      if (m < (kind == 0 ? nx - 1 : nx)) ov[q] = old_of(kind)[row_of(kind, j0 + r) + 1 + m];
    }
  }
  // 2. one wavefront per field, one lane per row
  const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = (int)threadIdx.x & 63;
  if (w < kPNF) {
    const int kind = pred_kind(w);
    lines_sweep_wave(pick4(P.tab, kind), kind == 0 ? nx - 1 : nx, S + (size_t)w * fs, ld, nrows, lane);
  }
  __syncthreads();
  // 3b. new field values, the x walls (constants the predictor never
  // changes) and, in rows next to the periodic halo, the images of both (the
  // reference's halo copy of the updated field).  The element positions are
  // decoded again rather than kept from 3a (an empty asm hides that the
  // thread index is the same): holding 32 addresses across the sweeps would
  // cost the registers of a second wave per SIMD.
  int tid3 = (int)threadIdx.x;
  __asm__ __volatile__("" : "+v"(tid3));
#pragma unroll
  for (int q = 0; q < kPE; q++) {
    const int e = tid3 + 256 * q;
    int slot, r, m;
    decode(e, slot, r, m);
    if (slot < kPNF && r < nrows) {
      const int kind = pred_kind(slot);
      const int n = kind == 0 ? nx - 1 : nx;
      const int n0 = kind == 0 ? nx + 1 : nx + 2;
      const int j = j0 + r;
      const size_t roff = row_of(kind, j);
      double * const out = pick4(P.out, kind);
      const bool edge = j < 2 || j >= ny
#if NDIMS == 3
                        || k < 2 || k >= nz
#endif
          ;
      if (m < n) {
        const double v = ov[q] + S[e];
        out[roff + 1 + m] = v;
        if (edge) store_images(out, 1 + m, n0, n1, j, PLANE1(k), ny, nz, 1, v);
      }
      if (m == 0 || m == n - 1) {
        const int xw = m == 0 ? 0 : n + 1;
        const double wv = old_of(kind)[roff + xw];
        out[roff + xw] = wv;
        if (edge) store_images(out, xw, n0, n1, j, PLANE1(k), ny, nz, 1, wv);
      }
    }
  }
}

template <int RT>
__global__ __launch_bounds__(256) void k_pred_lines(Grid g, PredLines P) {
  k_pred_lines_body<RT>(hw_block(), g, P);
}

__device__ inline double prefactor_div(double a, double b) { return a / b; }

// fluid_correct_velocity + fluid_update_pressure fused, with halo images
// and the Neumann walls of p
__device__ __forceinline__ void k_corr_press_body(const VBlock & vb, Grid g, double pref_c, double pref_p, const double * __restrict__ psi,
                             double * __restrict__ ux, double * __restrict__ uy,
                             double * __restrict__ uz, double * __restrict__ p) {
  const int nx = g.nx, ny = g.ny;
  THREAD_IJK(1, 1, 1)
  if (i > nx || j > ny) return;
#if NDIMS == 3
  const int nz = g.nz;
#else
  const int nz = 1;
  (void)uz;
#endif
  // every load first (the image stores below are loops the compiler does
  // not move loads across): one memory round trip per thread instead of four
  const double pc = psi[I_C1(i, j, k)];
  const double pm = psi[I_C1(i - 1, j, k)];
  const double pp = psi[I_C1(i + 1, j, k)];
  const double pym = psi[I_C1(i, j - 1, k)];
  const double u0 = ux[I_UX(i, j, k)];   // i = 1: the wall entry, loaded and unused
  const double v0 = uy[I_C1(i, j, k)];
  const double p0 = p[I_C1(i, j, k)];
  const double dxc_m = g.dxc[i - 1], dxc_p = g.dxc[i], dxf = g.dxf[i - 1];
#if NDIMS == 3
  const double pzm = psi[I_C1(i, j, k - 1)];
  const double w0 = uz[I_C1(i, j, k)];
#endif
  if (i >= 2) {
    const double v = u0 - prefactor_div(pref_c, dxc_m) * (+pc - pm);
    store_periodic(ux, i - 1, nx + 1, ny + 2, j, PLANE1(k), ny, nz, 1, v);
  }
  {
    const double v = v0 - prefactor_div(pref_c, g.dy) * (+pc - pym);
    store_periodic(uy, i, nx + 2, ny + 2, j, PLANE1(k), ny, nz, 1, v);
  }
#if NDIMS == 3
  {
    const double v = w0 - prefactor_div(pref_c, g.dz) * (+pc - pzm);
    store_periodic(uz, i, nx + 2, ny + 2, j, PLANE1(k), ny, nz, 1, v);
  }
#endif
  {
    double v = p0;
    v += pc;
    const double dpsidx_xm = (-pm + pc) / dxc_m;
    const double dpsidx_xp = (-pc + pp) / dxc_p;
    v -= pref_p / dxf * (-dpsidx_xm + dpsidx_xp);
    store_periodic(p, i, nx + 2, ny + 2, j, PLANE1(k), ny, nz, 1, v);
    neumann_walls(p, i, nx, ny + 2, j, PLANE1(k), ny, nz, v);
  }
}

__global__ __launch_bounds__(256) void k_corr_press(Grid g, double pref_c, double pref_p, const double * __restrict__ psi,
                             double * __restrict__ ux, double * __restrict__ uy,
                             double * __restrict__ uz, double * __restrict__ p) {
  k_corr_press_body(hw_block(), g, pref_c, pref_p, psi, ux, uy, uz, p);
}

}  // namespace vofk
