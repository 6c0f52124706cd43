// 2D weakly compressible SPH on gfx950.  SOPHIA_single comments out
// KERNEL_interaction2D (WCSPH_Calc.cuh:187), so the 2D step follows
// SOPHIA_multi (WCSPH_Calc.cuh:368-563) on one GPU's share: no
// KERNEL_set_alpha (Lagrangian particles, re-sorted every step), KGC,
// prep2D, prep2D_coupling, boundary2D, continuity, EOS, interaction2D,
// KERNEL_time_update and xsph2D.  The 2D kernels of the reference differ
// from the 3D ones in their formulas, not only in the dimension, so they
// are separate kernels here.  Particles are one row of every field (no
// SP1/P1 pair): the corrector keeps the predicted x, y and rho in xp, yp,
// rhop for XSPH, which reads them as the reference reads SP1.
// Arithmetic is FP64 in every kernel; R is the storage type of the
// velocity-like fields.
#include "common.hpp"
#include "nbr.hpp"

namespace sophia {

namespace {

constexpr int kMoving = 9;  // MOVING (Parameters.cuh:12)

__device__ inline double dist2d(double dx, double dy) {
#pragma clang fp contract(off)
  const double a = dx * dx;
  const double b = dy * dy;
  return sqrt(a + b);
}

// KERNEL_clc_correction_KGC_2D (functions_PREP2.cuh:73-159): inverse of the
// 2x2 moment matrix sum_j -m_j dW (x_ij x_ij^T) / (rho_j r) over r > 0,
// identity when |det| <= Min_det.  Stored in the symmetric 3x3 slots
// (xx, xy, 0, yy, 0, 0) so the 3D correction applies it to (gx, gy, 0).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_kgc2d(DevParams P, SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2 || v.ptype[i] > 1000) return;
  const int n = v.n;
  const double h = v.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h);
  double xx = 0, xy = 0, yy = 0;
  for_nbrs4(v, i, false, [&](const Nb<R>& q, bool on) {
    const double dx = q.dx, dy = q.dy;
    const double r = dist2d(dx, dy);
    const bool use = on && q.pt <= 1000 && r > 0;
    const double dw = knl_dw<double>(P.kernel_type, A, h, r);
    const double mtd = q.m * dw;
    const double rtd = 1.0 / (q.rho * r);
    double t;
    t = -mtd * dx * dx; t *= rtd; xx += use ? t : 0.0;
    t = -mtd * dy * dx; t *= rtd; xy += use ? t : 0.0;
    t = -mtd * dy * dy; t *= rtd; yy += use ? t : 0.0;
  });
  const double det = xx * yy - xy * xy;
  double a = 1, b = 0, d = 1;
  if (fabs(det) > 1e-4) {  // Min_det (Parameters.cuh:125)
    const double rd = 1.0 / det;
    a = yy * rd;
    b = -xy * rd;
    d = xx * rd;
  }
  v.cm[i] = (R)a;
  v.cm[n + i] = (R)b;
  v.cm[2 * n + i] = (R)0;
  v.cm[3 * n + i] = (R)d;
  v.cm[4 * n + i] = (R)0;
  v.cm[5 * n + i] = (R)0;
}

// KERNEL_clc_prep2D (functions_PREP2.cuh:1247-1470) with
// KERNEL_clc_prep2D_coupling (:2944-3046): Shepard sum when the filter is
// due, reference density m / (h / 1.6)^2 every step, the Antuono density
// gradient -(rho_j - rho_i) m_j / rho_j grad~W; porosity 1 (no spheres in
// 2D).  Turbulence, p_type switching and surface-tension normals are off.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_prep2d(DevParams P, SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2) return;
  v.por[i] = 1.0;
  if (v.ptype[i] > 1000) return;
  const int n = v.n;
  const double h = v.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h);
  const double rhoi = v.rho[i];
  const bool anto = P.delsph == 2;
  double flt = 0, rx = 0, ry = 0;
  for_nbrs4(v, i, true, [&](const Nb<R>& q, bool on) {
    const bool use = on && q.pt <= 1000;
    const double dx = q.dx, dy = q.dy;
    const double r = dist2d(dx, dy) + 1e-20;
    const double mj = q.m, rhoj = q.rho;
    if (P.filt_step) flt += use ? mj / rhoj * knl_w<double>(P.kernel_type, A, h, r) : 0.0;
    if (anto) {
      const double dw = knl_dw<double>(P.kernel_type, A, h, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = 0.0;
      kgc_apply(P, v.cm, n, i, gx, gy, gz);
      rx += use ? -(rhoj - rhoi) * (mj / rhoj) * gx : 0.0;
      ry += use ? -(rhoj - rhoi) * (mj / rhoj) * gy : 0.0;
    }
  });
  const double hv = h / P.h_over_dx;
  v.rho_ref[i] = v.m[i] / (hv * hv);
  if (P.filt_step) v.flt_s[i] = (R)flt;
  if (anto) {
    v.grx[i] = (R)rx;
    v.gry[i] = (R)ry;
    v.grz[i] = (R)0;
  }
}

// KERNEL_boundary2D (functions_BC.cuh:1234-1328), noslip_bc = 1: walls
// (p_type 0) and moving walls (9) take 2 u_i [moving] minus the Shepard
// average of the velocity of every other p_type (the reference compares
// p_type as unsigned, so negative types count).  Two passes, so every read
// sees the velocities before the update.
__device__ inline bool noslip_wall(int pt) {
  const uint32_t u = (uint32_t)pt;
  return !(u > 0 && u != (uint32_t)kMoving);
}
template <typename R>
__global__ __launch_bounds__(kBlock) void k_boundary2d(DevParams P, SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] > 2 || pti > 1000 || !noslip_wall(pti)) return;
  const double h = v.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h);
  double tx = 0, ty = 0, fl = 1.0e-10;
  for_nbrs4(v, i, true, [&](const Nb<R>& q, bool on) {
    const bool use = on && q.pt <= 1000 && (uint32_t)q.pt > 0;
    const double r = dist2d(q.dx, q.dy) + 1e-20;
    const double w = knl_w<double>(P.kernel_type, A, h, r);
    tx += use ? (double)q.ux * w : 0.0;
    ty += use ? (double)q.uy * w : 0.0;
    fl += use ? w : 0.0;
  });
  const double mv = pti == kMoving ? 1.0 : 0.0;
  v.xsx[i] = (R)(2 * (double)v.ux[i] * mv - tx / fl);
  v.xsy[i] = (R)(2 * (double)v.uy[i] * mv - ty / fl);
}
template <typename R>
__global__ void k_boundary2d_apply(SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] > 2 || pti > 1000 || !noslip_wall(pti)) return;
  v.ux[i] = v.xsx[i];
  v.uy[i] = v.xsy[i];
}

// KERNEL_clc_continuity_norm2D (functions_MASS.cuh:670-769) -> drho, walls
// included with their velocity taken as 0: the divergence
// sum m_j (rho_i / rho_j) (u_i - u_j) . grad~W times rho_i / rho_ref_i, and
// delta-SPH on rho_j - rho_i (Molteni, or Antuono with the gradients).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_continuity2d(DevParams P, SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] > 2 || pti > 1000) return;
  const int n = v.n;
  const double h = v.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h);
  const double uxi = (double)v.ux[i] * (pti > 0), uyi = (double)v.uy[i] * (pti > 0);
  const double rhoi = v.rho[i], gxi = (double)v.grx[i], gyi = (double)v.gry[i];
  const bool dsph = P.delsph > 0, anto = P.delsph == 2;
  double tx = 0, ty = 0, tR = 0;
  for_nbrs4(v, i, true, [&](const Nb<R>& q, bool on) {
    const bool use = on && q.pt <= 1000;
    const double dx = q.dx, dy = q.dy;
    const double r = dist2d(dx, dy) + 1e-20;
    const double dw = knl_dw<double>(P.kernel_type, A, h, r);
    double gx = dw * dx / r, gy = dw * dy / r, gz = 0.0;
    if (P.kgc > 0) kgc_apply(P, v.cm, n, i, gx, gy, gz);
    const double mj = q.m, rhoj = q.rho;
    const double uxj = (double)q.ux * (q.pt > 0), uyj = (double)q.uy * (q.pt > 0);
    const double tmprho = mj * (rhoi / rhoj);
    tx += use ? (uxi - uxj) * tmprho * gx : 0.0;
    ty += use ? (uyi - uyj) * tmprho * gy : 0.0;
    if (dsph) {
      double phi = (gxi + (double)q.grx) * -dx;
      phi += (gyi + (double)q.gry) * -dy;
      phi = (-0.5 * phi) * (anto ? 1.0 : 0.0) + (rhoj - rhoi);
      double tr = -2 * P.delta_sph * h * P.soundspeed * (mj / rhoj);
      tr *= phi * dw / r;
      tR += use ? tr : 0.0;
    }
  });
  v.drho[i] = (R)((tx + ty) * (rhoi / v.rho_ref[i]) + tR * (dsph ? 1.0 : 0.0));
}

// KERNEL_interaction2D (functions_INTERACTION.cuh:64-411), i_type 1:
// pressure -m_j (p_i + p_j) / (eps_i rho_i eps_j rho_j), laminar viscosity
// (no eps factor), artificial viscosity, gravity on y and conduction
// divided by m_i, as the reference writes them; interface, repulsive,
// surface-tension, Boussinesq and concentration terms are off.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_interaction2d(DevParams P, SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] != 1 || pti > 1000) return;
  const int n = v.n;
  const double h = v.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h);
  const double uxi = v.ux[i], uyi = v.uy[i], tempi = v.temp[i];
  const double pi = v.pres[i], mi = v.m[i], rhoi = v.rho[i], pori = v.por[i];
  const bool tab = P.prop_table != 0;
  const double visi = tab ? (double)v.visp[i] : P.mu_fluid;
  double kci = 0, cpi = 1, eta = 0;
  if (P.con_solve) {
    kci = tab ? (double)v.condp[i] : (pti > 1000 || pti < -1000 ? P.k_dem : P.k_fluid);
    cpi = pti > 1000 || pti < -1000 ? P.cp_dem : P.cp_fluid;
    eta = 0.001 * h;
  }
  double tx = 0, ty = 0, Rc = 0, pgx = 0, pgy = 0;
  for_nbrs4(v, i, true, [&](const Nb<R>& q, bool on) {
    const bool use = on && q.pt <= 1000;
    const double dx = q.dx, dy = q.dy;
    const double r = dist2d(dx, dy) + 1e-20;
    const double dw = knl_dw<double>(P.kernel_type, A, h, r);
    double gx = dw * dx / r, gy = dw * dy / r, gz = 0.0;
    if (P.kgc > 0) kgc_apply(P, v.cm, n, i, gx, gy, gz);
    const double mj = q.m, rhoj = q.rho, pj = q.pres, porj = q.por;
    const double uxj = q.ux, uyj = q.uy;
    if (P.fp_solve) {
      const double pp = P.pgrad_form ? pj - pi : pi + pj;
      const double Cp = -mj * pp / (pori * rhoi * porj * rhoj);
      const double tpx = Cp * gx, tpy = Cp * gy;
      tx += use ? tpx : 0.0;
      ty += use ? tpy : 0.0;
      pgx += use ? rhoi * pori * tpx : 0.0;
      pgy += use ? rhoi * pori * tpy : 0.0;
    }
    if (P.fv_solve) {
      const double visj = tab ? (double)v.visp[q.j] : P.mu_fluid;
      double Cv = dx * gx + dy * gy;
      Cv *= 4 * (mj / (rhoi * rhoj));
      Cv *= (visi * visj) / (visi + visj);
      Cv /= r;
      Cv /= r;
      tx += use ? Cv * (uxi - uxj) : 0.0;
      ty += use ? Cv * (uyi - uyj) : 0.0;
    }
    if (P.fva_solve) {
      const double uij_xij = (uxi - uxj) * dx + (uyi - uyj) * dy;
      const double hij = (h + q.h) * 0.5;
      double phi = hij * uij_xij;
      phi /= (r * r + 0.01 * hij * hij);
      double Pij = mi * phi * (-0.1 * P.soundspeed + 0.02 * phi);
      Pij /= (rhoi + rhoj);
      Pij *= 0.5;
      const bool av = use && uij_xij < 0;
      tx += av ? -Pij * gx : 0.0;
      ty += av ? -Pij * gy : 0.0;
    }
    if (P.con_solve) {
      const int ptj = q.pt;
      const double kcj = tab ? (double)v.condp[q.j] : (ptj > 1000 || ptj < -1000 ? P.k_dem : P.k_fluid);
      double H = 4.0 * mj * kcj * kci * (tempi - (double)q.temp) * dw;
      H /= mi * cpi * (r + eta * eta) * rhoi * rhoj * (kci + kcj);
      Rc += use ? H : 0.0;
    }
  });
  if (P.fg_solve) ty += -P.gravity;
  v.fx[i] = (R)tx;
  v.fy[i] = (R)ty;
  v.fz[i] = (R)0;
  v.fmag[i] = (R)sqrt(tx * tx + ty * ty);
  v.pgx[i] = (R)pgx;
  v.pgy[i] = (R)pgy;
  v.pgz[i] = (R)0;
  if (P.con_solve) v.dtemp[i] = (R)Rc;
}

// KERNEL_time_update (functions_TIME.cuh:523-675), i_type 1: u = u0 + a dt
// for fluid (0 for p_type <= 0, u0 kept past u_limit), x = x0 + u dt
// without XSPH, x0 + u dt for moving walls, rho = rho0 + drho dt for every
// p_type; the temperature keeps its predicted value (the reference's
// enthalpy path carries heat).  With XSPH the predicted x, y and rho of
// every particle are kept first.
template <typename R>
__global__ void k_time_update2d(DevParams P, SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  if (P.xsph_solve) {
    v.xp[i] = v.x[i];
    v.yp[i] = v.y[i];
    v.rhop[i] = v.rho[i];
  }
  if (v.itype[i] != 1) return;
  const int t = v.ptype[i];
  const double dt = P.dt_sph;
  if (t == kMoving) {
    v.x[i] = v.x0[i] + (double)v.ux[i] * dt;
    v.y[i] = v.y0[i] + (double)v.uy[i] * dt;
    v.z[i] = v.z0[i] + (double)v.uz[i] * dt;
  } else {
    double ux0 = 0, uy0 = 0, uz0 = 0, ax = 0, ay = 0, az = 0;
    if (t > 0) {
      ux0 = v.ux0[i]; uy0 = v.uy0[i]; uz0 = v.uz0[i];
      ax = v.fx[i]; ay = v.fy[i]; az = v.fz[i];
    }
    double uxc = ux0 + ax * dt, uyc = uy0 + ay * dt, uzc = uz0 + az * dt;
    if (uxc * uxc + uyc * uyc + uzc * uzc >= P.u_limit * P.u_limit) {
      uxc = ux0;
      uyc = uy0;
      uzc = uz0;
    }
    if (!P.xsph_solve) {
      v.x[i] = v.x0[i] + uxc * dt;
      v.y[i] = v.y0[i] + uyc * dt;
      v.z[i] = v.z0[i] + uzc * dt;
    }
    v.ux[i] = (R)uxc;
    v.uy[i] = (R)uyc;
    v.uz[i] = (R)uzc;
  }
  if (P.rho_type == 1) v.rho[i] = v.rho0[i] + (double)v.drho[i] * dt;
}

// KERNEL_xsph2D (functions_TIME.cuh:1520-1607): x = x0 + (u_i + c sum_j
// m_j / rho_j (u_j - u_i) W / flt_s_i) dt for fluid, x0 for walls, x0 + u dt
// for moving walls; corrected velocities, predicted positions and
// densities; r < kappa h without the 1e-20 (flagged entries count).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_xsph2d(DevParams P, SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] != 1 || pti > 1000) return;
  const int n = v.n;
  const double h = v.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h), Rs = P.kappa * h;
  const double xi = v.xp[i], yi = v.yp[i];
  const double uxi = v.ux[i], uyi = v.uy[i], fl = v.flt_s[i], c = P.c_xsph;
  double tx = 0, ty = 0;
  const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
  for (int k = 0; k < cnt; k++) {
    const int j = (int)(v.nbr[nbr_slot(k, i, n)] & ~kNbrFlag);
    if (v.ptype[j] > 1000) continue;
    const double r = dist2d(xi - v.xp[j], yi - v.yp[j]);
    if (!(r < Rs)) continue;
    const double w = knl_w<double>(P.kernel_type, A, h, r);
    const double mj = v.m[j], rhoj = v.rhop[j];
    tx += c * mj / rhoj * (-uxi + (double)v.ux[j]) * w / fl;
    ty += c * mj / rhoj * (-uyi + (double)v.uy[j]) * w / fl;
  }
  const double dt = P.dt_sph;
  if (pti == kMoving) {
    v.x[i] = v.x0[i] + uxi * dt;
    v.y[i] = v.y0[i] + uyi * dt;
  } else {
    const double on = pti > 0 ? 1.0 : 0.0;
    v.x[i] = v.x0[i] + (uxi + tx) * dt * on;
    v.y[i] = v.y0[i] + (uyi + ty) * dt * on;
  }
}

}  // namespace

#define W2_LAUNCH(kern, ...)                                 \
  do {                                                       \
    if (v.n > 0) {                                           \
      kern<R><<<grid_xcd(v.n), kBlock, 0, s>>>(__VA_ARGS__); \
      HIP_TRY(hipGetLastError());                            \
    }                                                        \
    return 0;                                                \
  } while (0)

template <typename R> int wc2_kgc(hipStream_t s, const DevParams& P, const SphView<R>& v) { W2_LAUNCH(k_kgc2d, P, v); }
template <typename R> int wc2_prep(hipStream_t s, const DevParams& P, const SphView<R>& v) { W2_LAUNCH(k_prep2d, P, v); }
template <typename R> int wc2_boundary(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  if (v.n <= 0) return 0;
  k_boundary2d<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
  k_boundary2d_apply<R><<<grid_xcd(v.n), kBlock, 0, s>>>(v);
  HIP_TRY(hipGetLastError());
  return 0;
}
template <typename R> int wc2_continuity(hipStream_t s, const DevParams& P, const SphView<R>& v) { W2_LAUNCH(k_continuity2d, P, v); }
template <typename R> int wc2_interaction(hipStream_t s, const DevParams& P, const SphView<R>& v) { W2_LAUNCH(k_interaction2d, P, v); }
template <typename R> int wc2_time_update(hipStream_t s, const DevParams& P, const SphView<R>& v) { W2_LAUNCH(k_time_update2d, P, v); }
template <typename R> int wc2_xsph(hipStream_t s, const DevParams& P, const SphView<R>& v) { W2_LAUNCH(k_xsph2d, P, v); }

#define W2_INST(R)                                                                \
  template int wc2_kgc<R>(hipStream_t, const DevParams&, const SphView<R>&);        \
  template int wc2_prep<R>(hipStream_t, const DevParams&, const SphView<R>&);       \
  template int wc2_boundary<R>(hipStream_t, const DevParams&, const SphView<R>&);   \
  template int wc2_continuity<R>(hipStream_t, const DevParams&, const SphView<R>&); \
  template int wc2_interaction<R>(hipStream_t, const DevParams&, const SphView<R>&); \
  template int wc2_time_update<R>(hipStream_t, const DevParams&, const SphView<R>&); \
  template int wc2_xsph<R>(hipStream_t, const DevParams&, const SphView<R>&);
W2_INST(float)
W2_INST(double)

}  // namespace sophia
