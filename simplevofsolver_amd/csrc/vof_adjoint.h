// vof_adjoint.h -- reverse-mode (adjoint) kernels of the 2D fused time step.
//
// Each kernel is the transpose of the linearisation of one forward stage,
// written in scatter form: a thread recomputes the forward quantities of the
// outputs it owns and adds (output adjoint) x (partial derivative) into the
// adjoints of the inputs it read, with FP64 atomic adds.  Adjoint arrays use
// the forward arrays' padded layouts but hold values only on canonical
// entries: a contribution to a periodic halo image is added to the interior
// entry it images, one to a Dirichlet wall value (a constant) is dropped,
// and one to a Neumann image (p, psi) is added to the adjacent interior
// cell.  Corner normals and curvatures of halo rows are images as well.
//
// The temperature is not carried (the differentiated configuration has no
// buoyancy).  Corner normals whose gradient is tiny (|grad vof| dx below
// `floor`, single-phase regions and the fringe of the interface) are treated
// as constants: the normalisation of a near-zero gradient is not smooth on
// any scale a perturbation can resolve, and its linearisation amplifies
// roundoff without bound.
#pragma once

#if NDIMS == 2

namespace vofa {

using vofk::Grid;
using vofk::VBlock;


__device__ inline int wrapj(int j, int ny) { return j < 1 ? j + ny : (j > ny ? j - ny : j); }

__device__ inline void aadd(double * p, double v) {
  if (v != 0.) unsafeAtomicAdd(p, v);
}

// ux DOFs i = 2..nx (faces 1 and nx+1 are walls)
__device__ inline void add_ux(double * a, int nx, int ny, int i, int j, double v) {
  if (i < 2 || i > nx) return;
  aadd(a + I_UX(i, wrapj(j, ny), 1), v);
}
// cell-centred DOFs with Dirichlet walls (uy)
__device__ inline void add_cd(double * a, int nx, int ny, int i, int j, double v) {
  if (i < 1 || i > nx) return;
  aadd(a + I_C1(i, wrapj(j, ny), 1), v);
}
// cell-centred with Neumann walls (p, psi)
__device__ inline void add_cn(double * a, int nx, int ny, int i, int j, double v) {
  i = i < 1 ? 1 : (i > nx ? nx : i);
  aadd(a + I_C1(i, wrapj(j, ny), 1), v);
}
// vof (walls are constants)
__device__ inline void add_vof(double * a, int nx, int ny, int i, int j, double v) {
  if (i < 1 || i > nx) return;
  aadd(a + I_VOF(i, wrapj(j, ny), 1), v);
}
// corner normals m(i, j), i = 1..nx+1, 2 components
__device__ inline void add_m(double * a, int nx, int ny, int i, int j, double vx, double vy) {
  double * p = a + 2 * I_DVOF(i, wrapj(j, ny), 1);
  aadd(p, vx);
  aadd(p + 1, vy);
}

// ---------------------------------------------------------------------------
// R1: correction and pressure update (correct_velocity/*.c, update_pressure.c)
//   u = u* - pc / dx (psi_c - psi_m),  p += psi - pp / dxf (Lx psi)
// lambda_u* = lambda_u (in place), lambda_p unchanged; lambda_psi accumulated.
__global__ void ka_corr(Grid g, double pc, double pp, const double * __restrict__ lux,
                        const double * __restrict__ luy, const double * __restrict__ lp,
                        double * __restrict__ lpsi) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  if (i >= 2) {
    const double a = pc / g.dxc[i - 1] * lux[I_UX(i, j, 1)];
    add_cn(lpsi, nx, ny, i, j, -a);
    add_cn(lpsi, nx, ny, i - 1, j, +a);
  }
  {
    const double a = pc / g.dy * luy[I_C1(i, j, 1)];
    add_cn(lpsi, nx, ny, i, j, -a);
    add_cn(lpsi, nx, ny, i, j - 1, +a);
  }
  {
    const double l = lp[I_C1(i, j, 1)];
    const double a = pp / g.dxf[i - 1] * l;
    const double rm = 1. / g.dxc[i - 1], rp = 1. / g.dxc[i];
    add_cn(lpsi, nx, ny, i, j, l + a * (rm + rp));
    add_cn(lpsi, nx, ny, i - 1, j, -a * rm);
    add_cn(lpsi, nx, ny, i + 1, j, -a * rp);
  }
}

// D^-1 lambda_psi into the interior right-hand side b[(i-1) + nx (j-1)]
__global__ void ka_psi_rhs(Grid g, const double * __restrict__ lpsi, double * __restrict__ b) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  b[(size_t)(i - 1) + (size_t)nx * (size_t)(j - 1)] = lpsi[I_C1(i, j, 1)] / g.dxf[i - 1];
}

// R2: lambda_r = D x (x = G D^-1 lambda_psi); r = pref div u*:
// lambda_u* += pref div^T lambda_r
__global__ void ka_div_t(Grid g, double pref, const double * __restrict__ x, double * __restrict__ lux,
                         double * __restrict__ luy) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  const double lr = g.dxf[i - 1] * x[I_C1(i, j, 1)];
  const double ax = pref * lr / g.dxf[i - 1];
  add_ux(lux, nx, ny, i + 1, j, +ax);
  add_ux(lux, nx, ny, i, j, -ax);
  const double ay = pref * lr / g.dy;
  add_cd(luy, nx, ny, i, j + 1, +ay);
  add_cd(luy, nx, ny, i, j, -ay);
}

// R3 input: v = D^-1 lambda_u* per DOF, line-major slabs for the x solver
// (ux: n = nx - 1 per line, D = DXC; uy: n = nx, D = DXF)
__global__ void ka_lines_in(Grid g, const double * __restrict__ lux, const double * __restrict__ luy,
                            double * __restrict__ vx, double * __restrict__ vy) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  if (i >= 2) vx[(size_t)(j - 1) * (nx - 1) + (i - 2)] = lux[I_UX(i, j, 1)] / g.dxc[i - 1];
  vy[(size_t)(j - 1) * nx + (i - 1)] = luy[I_C1(i, j, 1)] / g.dxf[i - 1];
}

struct AdjRhs {
  const double * ux;      // forward velocity at the stage
  const double * uy;
  const double * lapx_ux; // x Laplacians (3 per point)
  const double * lapx_c;
  const double * qx;      // M^-1 D^-1 lambda_u* (ux- and uy-shaped arrays)
  const double * qy;
  double * lux;           // lambda_u (holds lambda_u*, accumulates)
  double * luy;
  double * lp;            // lambda_p (accumulates)
  double * csx;           // beta carry of lambda_sa (in: from the next stage, out: to the previous)
  double * csy;
  double * lfx;           // lambda of the interface force
  double * lfy;
  double ca, cb, cg, dif;
};

// R4 + R5: lambda_delta = D q; lambda_sa = ca lambda_delta + carry, new carry
// cb lambda_delta; lambda_sg = cg lambda_delta; then the transposed source
// stencils of compute_rhs_ux / compute_rhs_uy (predict/ux.c, predict/uy.c)
__global__ void ka_rhs(Grid g, AdjRhs a) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  const double * ux = a.ux;
  const double * uy = a.uy;
  if (i >= 2) {
    const double dxci = g.dxc[i - 1];
    const size_t is = I_SX(i, j, 1);
    const double ld = dxci * a.qx[I_UX(i, j, 1)];
    const double lsa = a.ca * ld + a.csx[is];
    a.csx[is] = a.cb * ld;
    const double lsg = a.cg * ld;
    a.lfx[is] = lsa;
    const double u_c = ux[I_UX(i, j, 1)];
    const double u_xm = ux[I_UX(i - 1, j, 1)], u_xp = ux[I_UX(i + 1, j, 1)];
    const double u_ym = ux[I_UX(i, j - 1, 1)], u_yp = ux[I_UX(i, j + 1, 1)];
    double d_xm = 0., d_c = 0., d_xp = 0., d_ym = 0., d_yp = 0.;
    {  // advection_x
      const double l = +0.5 / dxci * (+0.5 * u_xm + 0.5 * u_c);
      const double u = -0.5 / dxci * (+0.5 * u_c + 0.5 * u_xp);
      d_xm += l + (u_xm - u_c) * (0.25 / dxci);
      d_c += -l - u + (u_xm - u_c) * (0.25 / dxci) - (u_xp - u_c) * (0.25 / dxci);
      d_xp += u - (u_xp - u_c) * (0.25 / dxci);
    }
    const double w_xm = 0.5 * g.dxf[i - 2] / dxci;
    const double w_xp = 0.5 * g.dxf[i - 1] / dxci;
    {  // advection_y
      const double uy_l = w_xm * uy[I_C1(i - 1, j, 1)] + w_xp * uy[I_C1(i, j, 1)];
      const double uy_u = w_xm * uy[I_C1(i - 1, j + 1, 1)] + w_xp * uy[I_C1(i, j + 1, 1)];
      const double l = +0.5 / g.dy * uy_l;
      const double u = -0.5 / g.dy * uy_u;
      d_ym += l;
      d_c += -l - u;
      d_yp += u;
      const double bl = lsa * (u_ym - u_c) * (0.5 / g.dy);
      const double bu = lsa * (u_yp - u_c) * (-0.5 / g.dy);
      add_cd(a.luy, nx, ny, i - 1, j, bl * w_xm);
      add_cd(a.luy, nx, ny, i, j, bl * w_xp);
      add_cd(a.luy, nx, ny, i - 1, j + 1, bu * w_xm);
      add_cd(a.luy, nx, ny, i, j + 1, bu * w_xp);
    }
    {  // diffusion_y
      const double ly = +1. / g.dy / g.dy, cy = -2. / g.dy / g.dy;
      d_ym += a.dif * ly;
      d_c += a.dif * cy;
      d_yp += a.dif * ly;
    }
    const double * L = a.lapx_ux + 3 * (i - 2);
    add_ux(a.lux, nx, ny, i - 1, j, lsa * d_xm + lsg * a.dif * L[0]);
    add_ux(a.lux, nx, ny, i, j, lsa * d_c + lsg * a.dif * L[1]);
    add_ux(a.lux, nx, ny, i + 1, j, lsa * d_xp + lsg * a.dif * L[2]);
    add_ux(a.lux, nx, ny, i, j - 1, lsa * d_ym);
    add_ux(a.lux, nx, ny, i, j + 1, lsa * d_yp);
    // sg -= (p_i - p_{i-1}) / dxc
    add_cn(a.lp, nx, ny, i, j, -lsg / dxci);
    add_cn(a.lp, nx, ny, i - 1, j, +lsg / dxci);
  }
  {
    const double dxfi = g.dxf[i - 1];
    const size_t is = I_SC(i, j, 1);
    const double ld = dxfi * a.qy[I_C1(i, j, 1)];
    const double lsa = a.ca * ld + a.csy[is];
    a.csy[is] = a.cb * ld;
    const double lsg = a.cg * ld;
    a.lfy[is] = lsa;
    const double v_c = uy[I_C1(i, j, 1)];
    const double v_xm = uy[I_C1(i - 1, j, 1)], v_xp = uy[I_C1(i + 1, j, 1)];
    const double v_ym = uy[I_C1(i, j - 1, 1)], v_yp = uy[I_C1(i, j + 1, 1)];
    double d_xm = 0., d_c = 0., d_xp = 0., d_ym = 0., d_yp = 0.;
    {  // advection_x
      const double ux_l = +0.5 * ux[I_UX(i, j - 1, 1)] + 0.5 * ux[I_UX(i, j, 1)];
      const double ux_u = +0.5 * ux[I_UX(i + 1, j - 1, 1)] + 0.5 * ux[I_UX(i + 1, j, 1)];
      const double l = +0.5 / dxfi * ux_l;
      const double u = -0.5 / dxfi * ux_u;
      d_xm += l;
      d_c += -l - u;
      d_xp += u;
      const double bl = lsa * (v_xm - v_c) * (0.25 / dxfi);
      const double bu = lsa * (v_xp - v_c) * (-0.25 / dxfi);
      add_ux(a.lux, nx, ny, i, j - 1, bl);
      add_ux(a.lux, nx, ny, i, j, bl);
      add_ux(a.lux, nx, ny, i + 1, j - 1, bu);
      add_ux(a.lux, nx, ny, i + 1, j, bu);
    }
    {  // advection_y
      const double l = +0.5 / g.dy * (+0.5 * v_ym + 0.5 * v_c);
      const double u = -0.5 / g.dy * (+0.5 * v_c + 0.5 * v_yp);
      d_ym += l + (v_ym - v_c) * (0.25 / g.dy);
      d_c += -l - u + (v_ym - v_c) * (0.25 / g.dy) - (v_yp - v_c) * (0.25 / g.dy);
      d_yp += u - (v_yp - v_c) * (0.25 / g.dy);
    }
    {
      const double ly = +1. / g.dy / g.dy, cy = -2. / g.dy / g.dy;
      d_ym += a.dif * ly;
      d_c += a.dif * cy;
      d_yp += a.dif * ly;
    }
    const double * L = a.lapx_c + 3 * (i - 1);
    add_cd(a.luy, nx, ny, i - 1, j, lsa * d_xm + lsg * a.dif * L[0]);
    add_cd(a.luy, nx, ny, i, j, lsa * d_c + lsg * a.dif * L[1]);
    add_cd(a.luy, nx, ny, i + 1, j, lsa * d_xp + lsg * a.dif * L[2]);
    add_cd(a.luy, nx, ny, i, j - 1, lsa * d_ym);
    add_cd(a.luy, nx, ny, i, j + 1, lsa * d_yp);
    add_cn(a.lp, nx, ny, i, j, -lsg / g.dy);
    add_cn(a.lp, nx, ny, i, j - 1, +lsg / g.dy);
  }
}

// R6: CSF force (force.c:11-66) from the stage's vof and curvature
__global__ void ka_force(Grid g, double tension, const double * __restrict__ vof,
                         const double * __restrict__ curv, const double * __restrict__ lfx,
                         const double * __restrict__ lfy, double * __restrict__ lvof,
                         double * __restrict__ lcurv) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  const double v = vof[I_VOF(i, j, 1)];
  const double cv = curv[I_C1(i, j, 1)];
  if (i >= 2) {
    const double l = lfx[I_SX(i, j, 1)] * tension;
    const double rdx = 1. / g.dxc[i - 1];
    const double grad = rdx * (-vof[I_VOF(i - 1, j, 1)] + v);
    const double kappa = 0.5 * (curv[I_C1(i - 1, j, 1)] + cv);
    add_vof(lvof, nx, ny, i, j, l * rdx * kappa);
    add_vof(lvof, nx, ny, i - 1, j, -l * rdx * kappa);
    add_cd(lcurv, nx, ny, i, j, l * grad * 0.5);
    add_cd(lcurv, nx, ny, i - 1, j, l * grad * 0.5);
  }
  {
    const double l = lfy[I_SC(i, j, 1)] * tension;
    const double rdy = 1. / g.dy;
    const double grad = rdy * (-vof[I_VOF(i, j - 1, 1)] + v);
    const double kappa = 0.5 * (curv[I_C1(i, j - 1, 1)] + cv);
    add_vof(lvof, nx, ny, i, j, l * rdy * kappa);
    add_vof(lvof, nx, ny, i, j - 1, -l * rdy * kappa);
    add_cd(lcurv, nx, ny, i, j, l * grad * 0.5);
    add_cd(lcurv, nx, ny, i, j - 1, l * grad * 0.5);
  }
}

// R7: vof_new = vof + dta s + dtb s_prev: lambda_s = dta lambda_vof + carry,
// new carry dtb lambda_vof (lambda_vof itself passes through unchanged)
__global__ void ka_vof_src(Grid g, double dta, double dtb, const double * __restrict__ lvof,
                           double * __restrict__ ls, double * __restrict__ cs) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  const size_t is = I_SC(i, j, 1);
  const double l = lvof[I_VOF(i, j, 1)];
  ls[is] = dta * l + cs[is];
  cs[is] = dtb * l;
}

// indicator and its derivative in z = n.x + d
__device__ inline double thinc_h(const double * n, double x, double y) {
  return 1. / (1. + exp(-2. * vofk::kVofBeta * (n[0] * x + n[1] * y + n[2])));
}

// R8: THINC fluxes (update/flxx.c, flxy.c): a thread owns the x face i and
// the y face j of cell (i, j); lambda_F from lambda_s, then the upwind
// cell's vof or (normal, intercept) and the face velocity
__global__ void ka_flux(Grid g, const double * __restrict__ ux, const double * __restrict__ uy,
                        const double * __restrict__ vof, const double * __restrict__ normal,
                        const double * __restrict__ ls, double * __restrict__ lux,
                        double * __restrict__ luy, double * __restrict__ lvof, double * __restrict__ lnrm) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  const double beta2 = 2. * vofk::kVofBeta;
  if (i >= 2) {
    // s(i) += F(i) / dxf(i); s(i-1) -= F(i) / dxf(i-1)
    const double lF = ls[I_SC(i, j, 1)] / g.dxf[i - 1] - ls[I_SC(i - 1, j, 1)] / g.dxf[i - 2];
    const double vel = ux[I_UX(i, j, 1)];
    const int ii = vel < 0. ? i : i - 1;
    const double x = vel < 0. ? -0.5 : +0.5;
    const double lv = vof[I_VOF(ii, j, 1)];
    if (lv < vofk::kVofMin || 1. - vofk::kVofMin < lv) {
      add_ux(lux, nx, ny, i, j, lF * lv);
      add_vof(lvof, nx, ny, ii, j, lF * vel);
    } else {
      const double * n = normal + 3 * I_NRM(ii, j, 1);
      double hs = 0., gx = 0., gy = 0., gd = 0.;
      for (int q = 0; q < NGAUSS; q++) {
        const double y = vofk::kGaussP[q];
        const double h = thinc_h(n, x, y);
        hs += vofk::kGaussW[q] * h;
        const double c = lF * vel * vofk::kGaussW[q] * beta2 * h * (1. - h);
        gx += c * x;
        gy += c * y;
        gd += c;
      }
      add_ux(lux, nx, ny, i, j, lF * hs);
      double * o = lnrm + 3 * I_NRM(ii, wrapj(j, ny), 1);
      aadd(o, gx);
      aadd(o + 1, gy);
      aadd(o + 2, gd);
    }
  }
  {
    const double lF = (ls[I_SC(i, j, 1)] - ls[I_SC(i, wrapj(j - 1, ny), 1)]) / g.dy;
    const double vel = uy[I_C1(i, j, 1)];
    const int jj = vel < 0. ? j : j - 1;
    const double y = vel < 0. ? -0.5 : +0.5;
    const double lv = vof[I_VOF(i, jj, 1)];
    if (lv < vofk::kVofMin || 1. - vofk::kVofMin < lv) {
      add_cd(luy, nx, ny, i, j, lF * lv);
      add_vof(lvof, nx, ny, i, jj, lF * vel);
    } else {
      const double * n = normal + 3 * I_NRM(i, jj, 1);
      double hs = 0., gx = 0., gy = 0., gd = 0.;
      for (int q = 0; q < NGAUSS; q++) {
        const double x = vofk::kGaussP[q];
        const double h = thinc_h(n, x, y);
        hs += vofk::kGaussW[q] * h;
        const double c = lF * vel * vofk::kGaussW[q] * beta2 * h * (1. - h);
        gx += c * x;
        gy += c * y;
        gd += c;
      }
      add_cd(luy, nx, ny, i, j, lF * hs);
      double * o = lnrm + 3 * I_NRM(i, wrapj(jj, ny), 1);
      aadd(o, gx);
      aadd(o + 1, gy);
      aadd(o + 2, gd);
    }
  }
}

// R9 + R10: per cell, the curvature (compute_curvature) and, for interface
// cells, the intercept (Newton root, implicit-function derivative) and the
// renormalised cell normal (compute_normal), back to the corner normals
__global__ void ka_normal_curv(Grid g, const double * __restrict__ vof, const double * __restrict__ dvof,
                               const double * __restrict__ normal, const double * __restrict__ lnrm,
                               const double * __restrict__ lcurv, double * __restrict__ lvof,
                               double * __restrict__ lm) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx || j > ny) return;
  const double rdx = 1. / g.dxf[i - 1], rdy = 1. / g.dy;
  double mx[4] = {0., 0., 0., 0.}, my[4] = {0., 0., 0., 0.};   // corners 00 10 01 11
  {
    // kappa = 0.5 (-dnx/dx - dny/dy)
    const double lk = lcurv[I_C1(i, j, 1)];
    const double ax = 0.5 * lk * rdx, ay = 0.5 * lk * rdy;
    mx[0] += ax; mx[1] -= ax; mx[2] += ax; mx[3] -= ax;
    my[0] += ay; my[1] += ay; my[2] -= ay; my[3] -= ay;
  }
  const double lv = vof[I_VOF(i, j, 1)];
  if (!(lv < vofk::kVofMin || 1. - vofk::kVofMin < lv)) {
    const double * d00 = dvof + 2 * I_DVOF(i, j, 1);
    const double * d10 = dvof + 2 * I_DVOF(i + 1, j, 1);
    const double * d01 = dvof + 2 * I_DVOF(i, j + 1, 1);
    const double * d11 = dvof + 2 * I_DVOF(i + 1, j + 1, 1);
    const double rx = (d00[0] + d10[0] + d01[0] + d11[0]) * rdx;
    const double ry = (d00[1] + d10[1] + d01[1] + d11[1]) * rdy;
    const double rn = sqrt(rx * rx + ry * ry);
    const double * nn = normal + 3 * I_NRM(i, j, 1);
    const double * ln = lnrm + 3 * I_NRM(i, j, 1);
    double lnx = ln[0], lny = ln[1];
    const double ld = ln[2];
    if (ld != 0.) {
      // d = -log(D) / (2 beta), F(D) = sum_g w_g / (1 + e_g D) - vof = 0
      const double D = exp(-2. * vofk::kVofBeta * nn[2]);
      double fD = 0., fnx = 0., fny = 0.;
      for (int qj = 0; qj < NGAUSS; qj++)
        for (int qi = 0; qi < NGAUSS; qi++) {
          const double w = vofk::kGaussW[qi] * vofk::kGaussW[qj];
          const double px = vofk::kGaussP[qi], py = vofk::kGaussP[qj];
          const double e = exp(-2. * vofk::kVofBeta * (nn[0] * px + nn[1] * py));
          const double den = 1. / (1. + e * D);
          fD -= w * e * den * den;
          const double fe = -w * D * den * den;          // dF/de
          fnx += fe * e * (-2. * vofk::kVofBeta * px);
          fny += fe * e * (-2. * vofk::kVofBeta * py);
        }
      const double lD = ld * (-0.5 / (vofk::kVofBeta * D));
      const double mu = -lD / fD;
      add_vof(lvof, nx, ny, i, j, -mu);
      lnx += mu * fnx;
      lny += mu * fny;
    }
    if (rn > 0.) {
      const double ux_ = nn[0], uy_ = nn[1];
      const double dot = ux_ * lnx + uy_ * lny;
      const double gx = (lnx - ux_ * dot) / rn * rdx;
      const double gy = (lny - uy_ * dot) / rn * rdy;
      for (int c = 0; c < 4; c++) {
        mx[c] += gx;
        my[c] += gy;
      }
    }
  }
  add_m(lm, nx, ny, i, j, mx[0], my[0]);
  add_m(lm, nx, ny, i + 1, j, mx[1], my[1]);
  add_m(lm, nx, ny, i, j + 1, mx[2], my[2]);
  add_m(lm, nx, ny, i + 1, j + 1, mx[3], my[3]);
}

// R11: corner normals (compute_gradient) back to vof, corners i = 1..nx+1,
// j = 1..ny (row ny + 1 is the image of row 1)
__global__ void ka_gradient(Grid g, double floor, const double * __restrict__ vof, const double * __restrict__ lm,
                            double * __restrict__ lvof) {
  const int nx = g.nx, ny = g.ny;
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
  const int j = (int)(blockIdx.y * blockDim.y + threadIdx.y) + 1;
  if (i > nx + 1 || j > ny) return;
  const double * l = lm + 2 * I_DVOF(i, j, 1);
  const double lx = l[0], ly = l[1];
  if (lx == 0. && ly == 0.) return;
  const double rdx = 1. / g.dxc[i - 1], rdy = 1. / g.dy;
  const double a = vof[I_VOF(i - 1, j - 1, 1)], b = vof[I_VOF(i, j - 1, 1)];
  const double c = vof[I_VOF(i - 1, j, 1)], d = vof[I_VOF(i, j, 1)];
  const double gx = rdx * (-a + b - c + d);
  const double gy = rdy * (-a - b + c + d);
  const double norm = sqrt(gx * gx + gy * gy);
  if (norm < floor * rdx) return;
  const double mx = gx / norm, my = gy / norm;
  const double dot = mx * lx + my * ly;
  const double lgx = (lx - mx * dot) / norm, lgy = (ly - my * dot) / norm;
  add_vof(lvof, nx, ny, i - 1, j - 1, -lgx * rdx - lgy * rdy);
  add_vof(lvof, nx, ny, i, j - 1, +lgx * rdx - lgy * rdy);
  add_vof(lvof, nx, ny, i - 1, j, -lgx * rdx + lgy * rdy);
  add_vof(lvof, nx, ny, i, j, +lgx * rdx + lgy * rdy);
}

// FFT in y of an interior right-hand side b (scaled by `scale`), as
// k_div_rfft_y does for the divergence
template <int ELEMS>
__global__ __launch_bounds__(vofk::kFftThreads) void ka_vec_rfft_y(
    Grid g, vofk::FftPlan P, const double2 * __restrict__ tw, const double * __restrict__ b, double scale,
    double2 * __restrict__ spec) {
  __shared__ double2 sa[ELEMS], sb[ELEMS];
  const int nx = g.nx;
  const int N = g.ny;
  const int NC = ELEMS / N;
  const int x0 = blockIdx.x * 2 * NC;
  for (int e = threadIdx.x; e < 2 * NC * N; e += blockDim.x) {
    const int xl = e % (2 * NC), j = e / (2 * NC);
    const int i = x0 + xl;
    const double r = i < nx ? scale * b[(size_t)i + (size_t)nx * (size_t)j] : 0.;
    double * slot = reinterpret_cast<double *>(&sa[(xl >> 1) * N + j]);
    slot[xl & 1] = r;
  }
  __syncthreads();
  double2 * res = vofk::lds_fft<double2>(sa, sb, NC, P, tw, -1.);
  const int nh = N / 2 + 1;
  for (int e = threadIdx.x; e < NC * nh; e += blockDim.x) {
    const int c = e / nh, ky = e % nh;
    const double2 z = res[c * N + ky];
    const double2 zc = res[c * N + (N - ky) % N];
    const double2 A = make_double2(0.5 * (z.x + zc.x), 0.5 * (z.y - zc.y));
    const double2 B = make_double2(0.5 * (z.y + zc.y), -0.5 * (z.x - zc.x));
    const int xa = x0 + 2 * c, xb = xa + 1;
    if (xa < nx) spec[(size_t)xa * nh + ky] = A;
    if (xb < nx) spec[(size_t)xb * nh + ky] = B;
  }
}

}  // namespace vofa

#endif
