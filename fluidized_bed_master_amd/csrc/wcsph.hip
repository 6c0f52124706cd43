// Weakly compressible SPH step (SOPHIA_single, WCSPH_Calc.cuh:4-340) on
// gfx950: the stages that differ from the ISPH step.  KGC, porosity, the
// Shepard / delta-SPH prep, the wall Neumann interpolation and XSPH are the
// ISPH kernels (the WCSPH reference kernels are the same text).
// KERNEL_set_alpha (functions_ALE.cuh:2-20) makes every SPH particle
// Eulerian, so the neighbour list of step 0 stays valid.
#include "common.hpp"
#include "nbr.hpp"

namespace sophia {

namespace {

constexpr int kMoving = 9;  // MOVING (Parameters.cuh:12)

// KERNEL_clc_predictor (functions_TIME.cuh:77-292): half step of x
// (Lagrangian only), u, rho (continuity) and T with the previous step's
// ftotal, drho and dtemp; saves the start-of-step state.
template <typename R>
__global__ void k_wc_predictor(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] == 3) return;
  const int t = v.ptype[i];
  const double dt = P.dt_sph;
  const double el = P.eulerian ? 0.0 : 1.0;
  const double bz = (v.btype[i] == 0);
  if (t != kMoving) {
    const double x0 = v.x[i], y0 = v.y[i], z0 = v.z[i];
    const double ux0 = (double)v.ux[i], uy0 = (double)v.uy[i], uz0 = (double)v.uz[i];
    if (t >= 0) {
      v.x[i] = x0 + ux0 * (dt * 0.5) * (t > 0) * el;
      v.y[i] = y0 + uy0 * (dt * 0.5) * (t > 0) * el;
      v.z[i] = z0 + uz0 * (dt * 0.5) * (t > 0) * el;
      v.ux[i] = (R)(ux0 + (double)v.fx[i] * bz * (dt * 0.5));
      v.uy[i] = (R)(uy0 + (double)v.fy[i] * bz * (dt * 0.5));
      v.uz[i] = (R)(uz0 + (double)v.fz[i] * bz * (dt * 0.5));
    }
    v.x0[i] = x0; v.y0[i] = y0; v.z0[i] = z0;
    v.ux0[i] = (R)ux0; v.uy0[i] = (R)uy0; v.uz0[i] = (R)uz0;
  }
  const double rho = v.rho[i];
  if (P.first_step) v.rho_ref[i] = rho;
  if (P.rho_type == 1 && t > 0) {
    if (P.first_step) v.vol0[i] = v.m[i] / rho;
    v.rho0[i] = rho;
    v.rho[i] = rho + (double)v.drho[i] * (dt * 0.5);
  } else {
    const double hv = v.h[i] / P.h_over_dx;
    if (P.first_step) v.vol0[i] = hv * hv * hv;
    v.rho0[i] = rho;
  }
  if (P.con_solve == 1) {
    v.temp0[i] = v.temp[i];
    v.temp[i] = (R)((double)v.temp[i] + (double)v.dtemp[i] * (dt * 0.5));
  }
}

// P3 reset after the predictor (WCSPH_Calc.cuh:59-60)
template <typename R>
__global__ void k_wc_reset(SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  const int n = v.n;
  v.fx[i] = 0; v.fy[i] = 0; v.fz[i] = 0; v.fmag[i] = 0;
  v.fpx[i] = 0; v.fpy[i] = 0; v.fpz[i] = 0;
  v.dtemp[i] = 0; v.drho[i] = 0;
  for (int c = 0; c < 6; c++) v.cm[c * n + i] = 0;
}

// KERNEL_clc_volume_update3D (functions_MASS.cuh:148-222): Shepard-corrected
// volume over the +-3 cells (Eulerian particles at count 0 only), then
// KERNEL_clc_mass_update (functions_MASS.cuh:45-60): m = rho vol.
template <typename R>
__global__ void k_volume_update(DevParams P, SphView<R> v, CellView cv, int count) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2) return;
  if (P.eulerian && count > 0) return;
  if (v.ptype[i] >= 1000) return;
  const double h = v.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h), Rs = P.kappa * h;
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  double tR = 0, filt = 0;
  for_cells_r(P, xi, yi, zi, cv, 3, [&](int j) {
    if (!(v.ptype[j] < 1000) || v.itype[j] == 4) return;
    const double r = dist_exact(xi - v.x[j], yi - v.y[j], zi - v.z[j]) + 1e-20;
    if (!(r < Rs)) return;
    const double w = knl_w<double>(P.kernel_type, A, h, r);
    tR += w;
    filt += v.m[j] / v.rho[j] / v.vol0[j] * w;
  });
  v.vol[i] = v.vol0[i] * 1.0 / tR * filt;
}

template <typename R>
__global__ void k_mass_update(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.ptype[i] > 1000 || v.itype[i] > 2) return;
  if (!P.eulerian) return;  // elix == eliy == 1
  v.m[i] = v.rho[i] * v.vol[i];
}

// KERNEL_clc_continuity_norm3D (functions_MASS.cuh:771-914) -> drho: the
// velocity divergence, the eps-advection term of Eulerian particles and
// the delta-SPH diffusion (Molteni, or Antuono with the density gradient).
// FP64 sums: drho integrates into rho, which the Tait EOS raises to gamma.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_continuity(DevParams P, SphView<R> v) {
  typedef double T;
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2 || v.ptype[i] > 1000) return;
  if (P.open_boundary > 0 && v.btype[i] > 0) return;
  if (v.ptype[i] <= 0) return;
  const int n = v.n;
  const T h = v.h[i], A = knl_tmpA<T>(P.kernel_type, P.dim, h);
  const T xi = v.x[i], yi = v.y[i], zi = v.z[i];
  const T uxi = v.ux[i], uyi = v.uy[i], uzi = v.uz[i];
  const T rhoi = v.rho[i], pori = v.por[i];
  const T gxi = v.grx[i], gyi = v.gry[i], gzi = v.grz[i];
  const T rri = v.rho_ref[i];
  const T el = P.eulerian ? 0.0 : 1.0;
  const bool dsph = P.delsph > 0, anto = P.delsph == 2;
  T tx = 0, ty = 0, tz = 0, px = 0, py = 0, pz = 0, tR = 0;
  for_nbrs4<true>(v, i, true, [&](const Nb<R>& q, bool on) {
    const bool use = on && q.pt <= 1000;
    const T dx = (T)q.dx, dy = (T)q.dy, dz = (T)q.dz;
    const T r = sqrt(dx * dx + dy * dy + dz * dz) + 1e-20;
    const T dw = knl_dw<T>(P.kernel_type, A, h, r);
    T gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
    if (P.kgc > 0) kgc_apply(P, v.cm, n, i, gx, gy, gz);
    const T mj = q.m, porj = q.por, rhoj = q.rho;
    const T tmprho = mj * (rhoi / rhoj);
    tx += use ? (uxi - (T)q.ux) * tmprho * gx : 0.0;
    ty += use ? (uyi - (T)q.uy) * tmprho * gy : 0.0;
    tz += use ? (uzi - (T)q.uz) * tmprho * gz : 0.0;
    px += use ? uxi * mj / rhoj * (rhoi / pori - rhoj / porj) * gx : 0.0;
    py += use ? uyi * mj / rhoj * (rhoi / pori - rhoj / porj) * gy : 0.0;
    pz += use ? uzi * mj / rhoj * (rhoi / pori - rhoj / porj) * gz : 0.0;
    if (dsph) {
      const T rrj = q.rref;
      T phi = (gxi + (T)q.grx) * -dx;
      phi += (gyi + (T)q.gry) * -dy;
      phi += (gzi + (T)q.grz) * -dz;
      phi = (-0.5 * phi) * (anto ? 1.0 : 0.0) + (rhoj / rrj - rhoi / rri);
      T trr = -2 * P.delta_sph * h * rri * P.soundspeed * (mj / rhoj);
      trr *= phi * dw / r;
      tR += use ? trr : 0.0;
    }
  });
  v.drho[i] = (R)((tx + ty + tz) + pori * (px + py + pz) * (1 - el) + tR * (dsph ? 1.0 : 0.0));
}

// KERNEL_EOS (functions_PROP.cuh:7-28): p = B ((rho / rho_ref)^gamma - 1)
// with B = b c0^2 / gamma; the reference's b is 1.186 (an air density: it
// applies no rho0 factor), eos_b_factor sets it
template <typename R>
__global__ void k_eos(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2 || v.ptype[i] > 1000 || v.btype[i] != 0) return;
  const double B = P.soundspeed * P.soundspeed * (P.eos_b) / P.gamma;
  v.pres[i] = B * (pow(v.rho[i] / v.rho_ref[i], P.gamma) - 1.0);
}

// KERNEL_interaction3D (functions_INTERACTION.cuh:413-832), i_type 1:
// pressure, laminar and artificial viscosity, gravity, conduction and the
// Eulerian advection term (interface, repulsive, surface-tension,
// Boussinesq and concentration terms are off in this build).  Geometry and
// the pressure term in FP64, as in the ISPH force kernels.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_wc_interaction(DevParams P, SphView<R> v) {
  typedef R T;
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] != 1 || v.ptype[i] > 1000) return;
  const int pti = v.ptype[i];
  const int n = v.n;
  const double hd = v.h[i];
  const T h = (T)hd;
  const double Ad = knl_tmpA<double>(P.kernel_type, P.dim, hd);
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  const T uxi = v.ux[i], uyi = v.uy[i], uzi = v.uz[i], tempi = v.temp[i];
  const double pid = v.pres[i], rhoid = v.rho[i], porid = v.por[i];
  const T mi = (T)v.m[i], rhoi = (T)rhoid, pori = (T)porid;
  const bool tab = P.prop_table != 0;
  const T visi = tab ? v.visp[i] : (T)P.mu_fluid;
  const T eta = (T)(0.001 * hd);
  const T kci = tab ? v.condp[i] : (T)(pti > 1000 || pti < -1000 ? P.k_dem : P.k_fluid);
  const T cpi = (T)(pti > 1000 || pti < -1000 ? P.cp_dem : P.cp_fluid);
  const T el = P.eulerian ? (T)0 : (T)1;
  const bool euler = el < (T)1;
  T tx = 0, ty = 0, tz = 0, Rc = 0, ex = 0, ey = 0, ez = 0, et = 0;
  double ptx = 0, pty = 0, ptz = 0, px = 0, py = 0, pz = 0;
  for_nbrs4(v, i, true, [&](const Nb<R>& q, bool on) {
    const bool use = on && q.pt <= 1000;
    const double dxd = q.dx, dyd = q.dy, dzd = q.dz;
    const double rd = sqrt(dxd * dxd + dyd * dyd + dzd * dzd) + 1e-20;
    const double dwd = knl_dw<double>(P.kernel_type, Ad, hd, rd);
    double gxd = dwd * dxd / rd, gyd = dwd * dyd / rd, gzd = dwd * dzd / rd;
    if (P.kgc > 0) kgc_apply(P, v.cm, n, i, gxd, gyd, gzd);
    if (P.fp_solve) {
      const double pp = P.pgrad_form ? q.pres - pid : pid + q.pres;
      const double Cp = -porid * q.m * pp / (rhoid * q.rho);
      ptx += use ? Cp * gxd : 0.0;
      pty += use ? Cp * gyd : 0.0;
      ptz += use ? Cp * gzd : 0.0;
      px += use ? rhoid / porid * Cp * gxd : 0.0;
      py += use ? rhoid / porid * Cp * gyd : 0.0;
      pz += use ? rhoid / porid * Cp * gzd : 0.0;
    }
    const T dx = (T)dxd, dy = (T)dyd, dz = (T)dzd, r = (T)rd, dw = (T)dwd;
    const T gx = (T)gxd, gy = (T)gyd, gz = (T)gzd;
    const T mj = (T)q.m, rhoj = (T)q.rho;
    const T uxj = q.ux, uyj = q.uy, uzj = q.uz;
    if (P.fv_solve && pti > 0) {
      T visj = tab ? v.visp[q.j] : (T)P.mu_fluid;
      T Cv = dx * gx + dy * gy + dz * gz;
      Cv *= (visi * visj) / (visi + visj);
      Cv *= 4 * pori * (mj / (rhoi * rhoj));
      Cv /= r;
      Cv /= r;
      tx += use ? Cv * (uxi - uxj) : (T)0;
      ty += use ? Cv * (uyi - uyj) : (T)0;
      tz += use ? Cv * (uzi - uzj) : (T)0;
    }
    if (P.fva_solve) {
      T uij_xij = (uxi - uxj) * dx + (uyi - uyj) * dy + (uzi - uzj) * dz;
      T hij = (h + (T)q.h) * (T)0.5;
      T phi = hij * uij_xij;
      phi /= (r * r + (T)0.01 * hij * hij);
      T Pij = mi * phi * ((T)(-0.1 * P.soundspeed) + (T)0.02 * phi);
      Pij /= (rhoi + rhoj);
      Pij *= (T)0.5;
      const bool av = use && uij_xij < 0;
      tx += av ? -Pij * gx : (T)0;
      ty += av ? -Pij * gy : (T)0;
      tz += av ? -Pij * gz : (T)0;
    }
    if (P.con_solve) {
      const int ptj = q.pt;
      T kcj = tab ? v.condp[q.j] : (T)(ptj > 1000 || ptj < -1000 ? P.k_dem : P.k_fluid);
      T H = (T)4.0 * mj * kcj * kci * (tempi - (T)q.temp) * dw;
      H /= cpi * (r + eta * eta) * rhoi * rhoj * (kci + kcj);
      Rc += use ? H : (T)0;
    }
    if (euler) {
      T c = mj / rhoj;
      ex += use ? (uxi * (uxj - uxi) * gx + uyi * (uxj - uxi) * gy + uzi * (uxj - uxi) * gz) * c * (1 - el) : (T)0;
      ey += use ? (uxi * (uyj - uyi) * gx + uyi * (uyj - uyi) * gy + uzi * (uyj - uyi) * gz) * c * (1 - el) : (T)0;
      ez += use ? (uxi * (uzj - uzi) * gx + uyi * (uzj - uzi) * gy + uzi * (uzj - uzi) * gz) * c * (1 - el) : (T)0;
      if (P.con_solve) {
        T tj = q.temp;
        et += use ? (uxi * (tj - tempi) * gx + uyi * (tj - tempi) * gy + uzi * (tj - tempi) * gz) * c * (1 - el)
                  : (T)0;
      }
    }
  });
  // pressure first, then the other terms, in the reference's per-pair order
  // of one accumulator; FP32 rounding of the sum of both is below the tests'
  // tolerance
  double fx = ptx + (double)tx, fy = pty + (double)ty, fz = ptz + (double)tz;
  if (P.fg_solve) fz += -P.gravity;
  v.fx[i] = (R)(fx - (double)ex);
  v.fy[i] = (R)(fy - (double)ey);
  v.fz[i] = (R)(fz - (double)ez);
  v.pgx[i] = (R)px;
  v.pgy[i] = (R)py;
  v.pgz[i] = (R)pz;
  v.fmag[i] = (R)sqrt(fx * fx + fy * fy + fz * fz);
  if (P.con_solve) v.dtemp[i] = Rc * (T)(pti != -3) - et;
}

// KERNEL_time_update_single (functions_TIME.cuh:676-877): full step from the
// start-of-step state with the half-step forces.  The temperature branch
// reads `k_con_solve==1 * p_type_i>0`, i.e. k_con_solve == (p_type_i > 0):
// with conduction off, walls take temp0 + dtemp dt and fluid keeps T.
template <typename R>
__global__ void k_wc_corrector(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] == 3) return;
  const int t = v.ptype[i];
  const double dt = P.dt_sph;
  const double el = P.eulerian ? 0.0 : 1.0;
  const double bz = (v.btype[i] == 0);
  if (t == kMoving) {
    v.x[i] = v.x0[i] + (double)v.ux[i] * dt;
    v.y[i] = v.y0[i] + (double)v.uy[i] * dt;
    v.z[i] = v.z0[i] + (double)v.uz[i] * dt;
  } else {
    const double ux0 = (double)v.ux0[i], uy0 = (double)v.uy0[i], uz0 = (double)v.uz0[i];
    double ax = 0, ay = 0, az = 0;
    if (t > 0) {
      ax = (double)v.fx[i] * bz;
      ay = (double)v.fy[i] * bz;
      az = (double)v.fz[i] * bz;
    }
    double uxc = ux0 + ax * dt, uyc = uy0 + ay * dt, uzc = uz0 + az * dt;
    if (uxc * uxc + uyc * uyc + uzc * uzc >= P.u_limit * P.u_limit) {
      uxc = ux0;
      uyc = uy0;
      uzc = uz0;
    }
    if (!P.xsph_solve) {
      v.x[i] = v.x0[i] + uxc * dt * (t > 0) * el;
      v.y[i] = v.y0[i] + uyc * dt * (t > 0) * el;
      v.z[i] = v.z0[i] + uzc * dt * (t > 0) * el;
    }
    v.ux[i] = (R)uxc;
    v.uy[i] = (R)uyc;
    v.uz[i] = (R)uzc;
  }
  if (P.rho_type == 1 && t > 0) v.rho[i] = v.rho0[i] + (double)v.drho[i] * dt;
  if (P.con_solve == (t > 0 ? 1 : 0)) v.temp[i] = (R)((double)v.temp0[i] + (double)v.dtemp[i] * dt);
}

}  // namespace

#define WC_LAUNCH(kern, ...)                                        \
  do {                                                              \
    if (v.n > 0) {                                                  \
      kern<R><<<grid_xcd(v.n), kBlock, 0, s>>>(__VA_ARGS__);        \
      HIP_TRY(hipGetLastError());                                   \
    }                                                               \
    return 0;                                                       \
  } while (0)

template <typename R> int wc_predictor(hipStream_t s, const DevParams& P, const SphView<R>& v) { WC_LAUNCH(k_wc_predictor, P, v); }
template <typename R> int wc_reset_p3(hipStream_t s, const SphView<R>& v) { WC_LAUNCH(k_wc_reset, v); }
template <typename R>
int wc_volume_mass(hipStream_t s, const DevParams& P, const SphView<R>& v, const CellView& cv, int count) {
  if (v.n <= 0) return 0;
  k_volume_update<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v, cv, count);
  k_mass_update<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
  HIP_TRY(hipGetLastError());
  return 0;
}
template <typename R> int wc_continuity(hipStream_t s, const DevParams& P, const SphView<R>& v) { WC_LAUNCH(k_continuity, P, v); }
template <typename R> int wc_eos(hipStream_t s, const DevParams& P, const SphView<R>& v) { WC_LAUNCH(k_eos, P, v); }
template <typename R> int wc_interaction(hipStream_t s, const DevParams& P, const SphView<R>& v) { WC_LAUNCH(k_wc_interaction, P, v); }
template <typename R> int wc_corrector(hipStream_t s, const DevParams& P, const SphView<R>& v) { WC_LAUNCH(k_wc_corrector, P, v); }

#define WC_INST(R)                                                                                           \
  template int wc_predictor<R>(hipStream_t, const DevParams&, const SphView<R>&);                           \
  template int wc_reset_p3<R>(hipStream_t, const SphView<R>&);                                               \
  template int wc_volume_mass<R>(hipStream_t, const DevParams&, const SphView<R>&, const CellView&, int);   \
  template int wc_continuity<R>(hipStream_t, const DevParams&, const SphView<R>&);                          \
  template int wc_eos<R>(hipStream_t, const DevParams&, const SphView<R>&);                                 \
  template int wc_interaction<R>(hipStream_t, const DevParams&, const SphView<R>&);                         \
  template int wc_corrector<R>(hipStream_t, const DevParams&, const SphView<R>&);
WC_INST(float)
WC_INST(double)

}  // namespace sophia
