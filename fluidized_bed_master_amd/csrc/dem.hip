// DEM and SPH-DEM coupling kernels on gfx950.
//
// DEM spheres are a few percent of the particles; their kernels are
// latency bound, run in FP64 on every path, and walk the 27 neighbour cells
// of the other species directly (no list).  The shear history is 16 slots
// per sphere in slot-major SoA streams (slot k of 64 consecutive spheres is
// contiguous), read from the start-of-step copy and written to a second
// buffer, which is the reference's SP1 -> TP1 split
// (ISPH_Calc_seperate.cuh:218).
#include "common.hpp"

namespace sophia {

namespace {

// KERNEL_clc_prep3D_coupling_dem (functions_PREP2.cuh:3367-3523): fluid
// Shepard sums seen by each sphere and the sphere-phase porosity.
template <typename R>
__global__ void k_porosity_dem(DevParams P, DemView d, SphView<R> v, CellView sc, CellView dc) {
  const int i = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);  // one wave per sphere
  if (i >= d.n) return;
  if (d.itype[i] > 2) return;
  const int pti = d.ptype[i];
  const double h = d.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h), Rs = P.kappa * h;
  const double xi = d.x[i], yi = d.y[i], zi = d.z[i];
  double fltd = 0, fltd2 = 0, flt = 0, por = 0;
  if (v.n > 0) {
    wave_for_cells(P, xi, yi, zi, sc, [&](int j) {
      double r = dist_exact(xi - v.x[j], yi - v.y[j], zi - v.z[j]) + 1e-20;
      if (!(r < Rs)) return;
      if (!(pti > 1000)) return;
      int t = v.ptype[j];
      if (t != 1 && t != 3) return;
      double w = knl_w<double>(P.kernel_type, A, h, r);
      double c = (double)v.m[j] / (double)v.rho[j] * w;
      if (t == 1) {
        fltd += c;
        fltd2 += c;
      } else {
        flt += c;
      }
    });
  }
  wave_for_cells(P, xi, yi, zi, dc, [&](int j) {
    double r = dist_exact(xi - d.x[j], yi - d.y[j], zi - d.z[j]) + 1e-20;
    if (!(r < Rs)) return;
    if (pti > 1000 && d.ptype[j] > 1000) {
      double rj = d.rad[j];
      por += knl_w<double>(P.kernel_type, A, h, r) * 4.18879 * rj * rj * rj;
    }
  });
  fltd = wave_sum(fltd);
  fltd2 = wave_sum(fltd2);
  flt = wave_sum(flt);
  por = wave_sum(por);
  if ((threadIdx.x & 63) != 0) return;
  if (pti > 1000) {
    d.flt_s[i] = flt;
    d.flt_sd[i] = fltd;
    d.flt_sd2[i] = fltd2;
    d.por[i] = 1.0 - (por / (flt + fltd + 1.0e-20));
  }
}

// emissivity of spheres (functions_PROP.cuh:420-440, graphite fit above 700 K)
__device__ inline double dem_emissivity(double T) {
  return T >= 700.0 ? 0.505586 + 7.92943e-4 * T - 5.228643e-7 * T * T + 1.10479e-10 * T * T * T : 0.842337;
}

// Radiative rate of sphere i from sphere j across the facing caps of a
// double cone (functions_DEM_INTERACTION2.cuh:1085-1152), per heat capacity
// of i's outer shell rc = rho_i V_shell c_p.  The cone radius is
// 0.56 (2 R*) (1 - 0.401038)^(-1/3); the reference's integer -1/3 makes the
// last factor 1 (SOPHIA_COMPAT_NUSSELT).
__device__ inline double dem_radiation(const DevParams& P, double radi, double radj, double dist, double Ti,
                                       double Tj, double rc) {
  const double radij = 1 / (1 / radi + 1 / radj);
  const double ai = radi / dist, aj = radj / dist, dA = aj - ai;
  const double cone = (P.compat & SOPHIA_COMPAT_NUSSELT) ? 1.0 : 1.1863156050445258;  // (1 - 0.401038)^(-1/3)
  const double dc_r = 0.56 * (2 * radij) * cone;
  const double dc_ar = 3.14159 * dc_r * dc_r;
  const double di_max = 0.5 * (sqrt(1 + 4.0 * dc_ar / (3.14159 * dist * dist * (1 - dA * dA))) - dA);
  const double cos_i = (1 - dA * dA) / (2 * di_max) - dA;
  const double cos_j = (1 - dA * dA) / (2 * (di_max + dA)) + dA;
  const double sin_i = sqrt(1 - cos_i * cos_i), sin_j = sqrt(1 - cos_j * cos_j);
  const double ri_cr = radi * sin_i, rj_cr = radj * sin_j;
  const double d_cr = dist - radi * cos_i - radj * cos_j;
  const double z_cr = 1 + (4 * rj_cr * rj_cr + d_cr * d_cr) / (ri_cr * ri_cr);
  const double F_cr = 0.5 * (z_cr - sqrt(z_cr * z_cr - 4.0 * rj_cr * rj_cr / (ri_cr * ri_cr)));
  const double F_ij = (1 + cos_i) * F_cr * 0.5;
  const double ar_i = 2.0 * 3.14159 * radi * radi * (1 - cos_i);
  const double ar_j = 2.0 * 3.14159 * radj * radj * (1 - cos_j);
  const double ar_ratio = ar_j / ar_i;
  const double Y_ij = F_ij + (1 - F_ij) * (ar_ratio / (1 + ar_ratio));
  const double emi = dem_emissivity(Ti), emj = dem_emissivity(Tj);
  const double em_inv = (1 - emi) / (emi * ar_i) + (1 - emj) / (emj * ar_j) + 1 / (Y_ij * ar_i);
  return -5.6704e-8 * (Ti * Ti * Ti * Ti - Tj * Tj * Tj * Tj) / (em_inv * rc);
}

// KERNEL_DEM_interaction3D_dem (functions_DEM_INTERACTION2.cuh:888-1766):
// Hertz normal spring + damping, Mindlin tangential spring + damping with a
// Coulomb cap, torque, gravity; shear displacement history keyed by dem_idx.
__global__ void k_dem_contact(DevParams P, DemView d, CellView dc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // linear: work concentrates near walls/spheres
  if (i >= d.n) return;
  const int n = d.n, S = P.dem_slots;
  // the new history starts as a copy of the old one
  for (int k = 0; k < S; k++) {
    d.hist_idx_new[k * n + i] = d.hist_idx[k * n + i];
    for (int c = 0; c < 3; c++) d.hist_ds_new[(k * 3 + c) * n + i] = d.hist_ds[(k * 3 + c) * n + i];
  }
  if (d.itype[i] != 1) return;
  if (d.ptype[i] <= 1000) return;
  const double xi = d.x[i], yi = d.y[i], zi = d.z[i], uxi = d.ux[i], uyi = d.uy[i], uzi = d.uz[i];
  const double wxi = d.wx[i], wyi = d.wy[i], wzi = d.wz[i], radi = d.rad[i], mi = d.m[i], rii = d.ri[i];
  const double G = P.dem_G, nu = P.dem_nu, E = G * (2 * (1 + nu));
  double tx = 0, ty = 0, tz = 0, qx = 0, qy = 0, qz = 0;
  // heat (con_solve): the outer shell (thickness rad/6) exchanges heat with
  // neighbouring spheres by radiation and through contact circles
  const bool heat = P.con_solve == 1;
  const double Ti = d.temp[i], dr2 = radi / 6.0;
  const double rc = d.rho[i] * ((4.0 / 3.0) * 3.141592 * (radi * radi * radi - (radi - dr2) * (radi - dr2) * (radi - dr2))) *
                    P.cp_dem;
  double h_rad = 0, h_con = 0;
  for_cells(P, xi, yi, zi, dc, [&](int j) {
    if (!(d.ptype[j] > 1000)) return;
    const double radj = d.rad[j];
    const int didx = d.dem_idx[j];
    const double dist = dist_exact(xi - d.x[j], yi - d.y[j], zi - d.z[j]);
    const double ovlp = radi + radj - dist;
    if (!(dist < 3.5 * radi && dist > 0.0)) return;
    if (heat) h_rad += dem_radiation(P, radi, radj, dist, Ti, d.temp[j], rc);
    if (!(dist < 2.01 * radi)) return;
    if (ovlp > 0) {
      if (heat) {
        // contact conduction (:1300-1315), harmonic-mean conductivity
        const double kc_ij = 1 / (1 / P.k_dem + 1 / P.k_dem);
        const double radij_h = 1 / (1 / radi + 1 / radj);
        h_con += -4.0 * kc_ij * sqrt((2 * ovlp) * (radij_h - 0.5 * ovlp)) * (Ti - d.temp[j]) / rc;
      }
      double dsx = 0, dsy = 0, dsz = 0;
      int slot = -1;
      for (int k = 0; k < S; k++)
        if (d.hist_idx[k * n + i] == didx) {
          dsx = d.hist_ds[(k * 3 + 0) * n + i];
          dsy = d.hist_ds[(k * 3 + 1) * n + i];
          dsz = d.hist_ds[(k * 3 + 2) * n + i];
          slot = k;
          break;
        }
      if (slot < 0)
        for (int k = 0; k < S; k++)
          if (d.hist_idx[k * n + i] < 1) {
            slot = k;
            break;
          }
      const double nx = (d.x[j] - xi) / dist, ny = (d.y[j] - yi) / dist, nz = (d.z[j] - zi) / dist;
      const double ov = ovlp;
      const double ai = radi - 0.5 * ov, aj = -(radj - 0.5 * ov);
      const double rcxi = ai * nx, rcyi = ai * ny, rczi = ai * nz;
      const double rcxj = aj * nx, rcyj = aj * ny, rczj = aj * nz;
      const double wxj = d.wx[j], wyj = d.wy[j], wzj = d.wz[j];
      const double urx = (d.ux[j] + (wyj * rczj - wzj * rcyj)) - (uxi + (wyi * rczi - wzi * rcyi));
      const double ury = (d.uy[j] + (wzj * rcxj - wxj * rczj)) - (uyi + (wzi * rcxi - wxi * rczi));
      const double urz = (d.uz[j] + (wxj * rcyj - wyj * rcxj)) - (uzi + (wxi * rcyi - wyi * rcxi));
      const double un = nx * urx + ny * ury + nz * urz;
      const double unx = un * nx, uny = un * ny, unz = un * nz;
      const double usx = urx - unx, usy = ury - uny, usz = urz - unz;
      dsx += usx * P.dt_dem;
      dsy += usy * P.dt_dem;
      dsz += usz * P.dt_dem;
      const double usm = sqrt(usx * usx + usy * usy + usz * usz) + 1e-20;
      const double sx = usx / usm, sy = usy / usm, sz = usz / usm;
      const double ymij = 1 / (((1 - nu * nu) / E) + ((1 - nu * nu) / E));
      const double Gij = 1 / (((2 - nu) / G) + ((2 - nu) / G));
      const double radij = 1 / (1 / radi + 1 / radj);
      const double mij = 1 / (1 / mi + 1 / d.m[j]);
      const double Cspr = P.dem_hertz * ymij * sqrt(radij) * sqrt(ov) * ov / mi;
      const double Csprs = 8.0 * Gij * sqrt(radij * ov) / mi;
      const double Kn = 2 * ymij * sqrt(radij * ov);
      const double Cd = P.dem_damping * sqrt(mij * Kn) / mi;
      const double Cds = P.dem_damping * sqrt(mij * Csprs * mi) / mi;
      tx += Cspr * (xi - d.x[j]) / dist;
      ty += Cspr * (yi - d.y[j]) / dist;
      tz += Cspr * (zi - d.z[j]) / dist;
      tx += Cd * unx;
      ty += Cd * uny;
      tz += Cd * unz;
      const double fssx = Csprs * dsx + Cds * usx, fssy = Csprs * dsy + Cds * usy, fssz = Csprs * dsz + Cds * usz;
      const double mfss = sqrt(fssx * fssx + fssy * fssy + fssz * fssz);
      const double ffx = P.dem_friction * Cspr * sx, ffy = P.dem_friction * Cspr * sy, ffz = P.dem_friction * Cspr * sz;
      const double mfsf = sqrt(ffx * ffx + ffy * ffy + ffz * ffz);
      const bool slip = !(mfss < mfsf);
      const double fsx = slip ? (fssx >= 0 ? 1.0 : -1.0) * fabs(ffx) : fssx;
      const double fsy = slip ? (fssy >= 0 ? 1.0 : -1.0) * fabs(ffy) : fssy;
      const double fsz = slip ? (fssz >= 0 ? 1.0 : -1.0) * fabs(ffz) : fssz;
      tx += fsx;
      ty += fsy;
      tz += fsz;
      qx += mi * (rcyi * fsz - rczi * fsy);
      qy += mi * (rczi * fsx - rcxi * fsz);
      qz += mi * (rcxi * fsy - rcyi * fsx);
      if (slot >= 0) {
        d.hist_idx_new[slot * n + i] = didx;
        d.hist_ds_new[(slot * 3 + 0) * n + i] = dsx;
        d.hist_ds_new[(slot * 3 + 1) * n + i] = dsy;
        d.hist_ds_new[(slot * 3 + 2) * n + i] = dsz;
      }
    } else {
      for (int k = 0; k < S; k++)
        if (d.hist_idx[k * n + i] == didx) {
          d.hist_idx_new[k * n + i] = 0;
          for (int c = 0; c < 3; c++) d.hist_ds_new[(k * 3 + c) * n + i] = 0.0;
          break;
        }
    }
  });
  if (P.fg_solve) tz += -P.gravity;
  if (heat) {
    // (:1725-1729) without the inner-shell terms: the shells temp1..3 are
    // never advanced on this path
    d.hrad[i] = h_rad;
    d.hcon[i] = h_con;
    d.dtemp[i] = h_rad + h_con;
  }
  d.fx[i] = tx;
  d.fy[i] = ty;
  d.fz[i] = tz;
  d.tqx[i] = qx / rii;
  d.tqy[i] = qy / rii;
  d.tqz[i] = qz / rii;
}

// KERNEL_DEM_coupling3D_dem (functions_SPH_DEM_COUPLING2.cuh:627-1129):
// kernel-averaged fluid velocity and porosity around each sphere, Di Felice
// drag (fluid regime) and the pressure-gradient force.
template <typename R>
__global__ void k_dem_coupling(DevParams P, DemView d, SphView<R> v, CellView sc) {
  const int i = blockIdx.x * (kBlock / 64) + (int)(threadIdx.x >> 6);  // one wave per sphere
  if (i >= d.n) return;
  if (d.itype[i] != 1 || d.ptype[i] <= 1000) return;
  const double xi = d.x[i], yi = d.y[i], zi = d.z[i], uxi = d.ux[i], uyi = d.uy[i], uzi = d.uz[i];
  const double radi = d.rad[i], di = 2 * radi, voli = (4.0 / 3.0) * 3.14159 * radi * radi * radi;
  const double mi = d.m[i], flt_air = d.flt_s[i], flt_fluid = d.flt_sd[i];
  const double h = d.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h), Rs = P.kappa * h;
  double tx = 0, ty = 0, tz = 0, ufx = 0, ufy = 0, ufz = 0, uax = 0, uay = 0, uaz = 0, por = 0;
  double tf = 0, ta = 0;  // kernel-weighted fluid / air temperature sums (:806-867)
  if (v.n > 0) {
    wave_for_cells(P, xi, yi, zi, sc, [&](int j) {
      const int t = v.ptype[j];
      const double porj = (double)v.por[j];
      if (!(t <= 1000 && porj < 1.0)) return;
      double r = dist_exact(xi - v.x[j], yi - v.y[j], zi - v.z[j]);
      if (!(r < Rs)) return;
      double w = knl_w<double>(P.kernel_type, A, h, r);
      double mj = v.m[j], rhoj = v.rho[j];
      if (t == 1 && flt_fluid > 0.000001) {
        ufx += (double)v.ux[j] * mj * w / rhoj;
        ufy += (double)v.uy[j] * mj * w / rhoj;
        ufz += (double)v.uz[j] * mj * w / rhoj;
        por += porj * mj * w / rhoj;
        if (P.con_solve == 1) tf += (double)v.temp[j] * mj * w / rhoj;
      }
      if (t == 3 && flt_air > 0.000001) {
        uax += (double)v.ux[j] * mj * w / rhoj;
        uay += (double)v.uy[j] * mj * w / rhoj;
        uaz += (double)v.uz[j] * mj * w / rhoj;
        por += porj * mj * w / rhoj;
        if (P.con_solve == 1) ta += (double)v.temp[j] * mj * w / rhoj;
      }
      if (t != 0 && t != SOPHIA_PT_MOVING && t != -3 && t != 3 && flt_fluid > 0.01) {
        double c = (mj / rhoj) * w * voli / (mi * flt_fluid + 1.0e-15);
        tx += (double)v.pgx[j] * c;
        ty += (double)v.pgy[j] * c;
        tz += (double)v.pgz[j] * c;
      }
    });
  }
  tx = wave_sum(tx); ty = wave_sum(ty); tz = wave_sum(tz);
  ufx = wave_sum(ufx); ufy = wave_sum(ufy); ufz = wave_sum(ufz);
  uax = wave_sum(uax); uay = wave_sum(uay); uaz = wave_sum(uaz);
  por = wave_sum(por);
  if (P.con_solve == 1) {
    tf = wave_sum(tf);
    ta = wave_sum(ta);
  }
  if ((threadIdx.x & 63) != 0) return;
  const double pori = por / (flt_fluid + 1.0e-15);
  const double sf = flt_fluid + 1.0e-15, sa = flt_air + 1.0e-15, sm = flt_fluid + flt_air + 1.0e-15;
  const double rfx = ufx / sf - uxi, rfy = ufy / sf - uyi, rfz = ufz / sf - uzi;
  const double rax = uax / sa - uxi, ray = uay / sa - uyi, raz = uaz / sa - uzi;
  const double rmx = (ufx + uax) / sm - uxi, rmy = (ufy + uay) / sm - uyi, rmz = (ufz + uaz) / sm - uzi;
  const double f_fluid = flt_fluid / sm, f_air = flt_air / sm;
  const double magf = sqrt(rfx * rfx + rfy * rfy + rfz * rfz);
  const double maga = sqrt(rax * rax + ray * ray + raz * raz);
  const double magm = sqrt(rmx * rmx + rmy * rmy + rmz * rmz);
  d.test6[i] = magf;
  d.por[i] = pori;
  const double nu_f = P.mu_fluid / P.rho_fluid;
  double dfx = 0, dfy = 0, dfz = 0, dax = 0, day = 0, daz = 0;
  const bool any = (flt_fluid + flt_air) > 0.00001;
  // convective heat (:996-1106): Ranz-Marshall Nu = 2 + 0.6 Re^0.5 Pr^(1/3)
  // (Pr^0 under SOPHIA_COMPAT_NUSSELT), h = Nu k / d over the specific
  // surface 6/d; the sphere's k and diffusivity by material (:708-718)
  const double Ti = d.temp[i], rhoi = d.rho[i];
  const int mat = radi < P.dem_mat_radius ? 1 : 0;
  const double ki = P.dem_k_mat[mat], cpi = ki / (P.dem_alpha_mat[mat] * rhoi), a_ht = 6.0 / di;
  const bool int_third = (P.compat & SOPHIA_COMPAT_NUSSELT) != 0;
  double dq_vol = 0;
  if (magf > 1.0e-8 && any && f_fluid > P.drag_switch) {
    double Re = di * magf * pori / (nu_f + 1.0e-15);
    double sq = 0.63 + 4.8 / sqrt(Re);
    double Cd = sq * sq;
    double lg = 1.5 - log10(Re);
    // ISPH (KERNEL_DEM_coupling3D_dem, :979-1009): Di Felice, eps^(2-beta);
    // WCSPH (KERNEL_DEM_coupling3D, :298-310): the Rong form, eps^(-1-beta)
    double beta = P.wcsph ? 2.65 * (1.0 + pori) - (5.3 - 3.5 * pori) * pori * pori * exp(-lg * lg / 2.0)
                          : 3.7 - 0.65 * exp(-lg * lg / 2.0);
    double k = (1.0 / 8.0) * Cd * P.rho_fluid * 3.14159 * di * di;
    double e = pow(pori, P.wcsph ? -1 - beta : 2 - beta);
    dfx = k * rfx * magf * e / mi;
    dfy = k * rfy * magf * e / mi;
    dfz = k * rfz * magf * e / mi;
    d.fdx[i] = dfx;
    d.fdy[i] = dfy;
    d.fdz[i] = dfz;
    const double Nu = 2.0 + 0.6 * pow(Re, 0.5) * (int_third ? 1.0 : cbrt(P.pr_fluid));
    dq_vol = Nu * P.k_fluid / di * a_ht * (tf / sf - Ti);
  }
  if (maga > 1.0e-8 && any && f_air > P.drag_switch) {
    double Re = di * pori * maga / (nu_f + 1.0e-15);
    double sq = 0.63 + 4.8 / sqrt(Re);
    double lg = 1.5 - log10(Re);
    double beta = 2.65 * (1.0 + pori) - (5.3 - 3.5 * pori) * pori * pori * exp(-lg * lg / 2.0);
    double k = (1.0 / 8.0) * sq * sq * P.rho_fluid * 3.14159 * di * di;
    double e = pow(pori, -1 - beta);
    dax = k * rax * maga * e / mi;
    day = k * ray * maga * e / mi;
    daz = k * raz * maga * e / mi;
    const double Nu = 2.0 + 0.6 * pow(Re, 0.5) * (int_third ? 1.0 : cbrt(P.pr_air));
    dq_vol = Nu * ki / di * a_ht * (ta / sa - Ti);
  }
  if ((maga > 1.0e-8 || magf > 1.0e-8) && any && f_air <= P.drag_switch && f_fluid <= P.drag_switch) {
    double rho_m = P.rho_fluid * f_air + P.rho_water * f_fluid;
    double vis_m = P.mu_fluid * f_air + P.mu_water * f_fluid;
    double nu_m = vis_m / (rho_m + 1.0e-15);
    double Re = di * pori * magm / (nu_m + 1.0e-15);
    double sq = 0.63 + 4.8 / sqrt(Re);
    double lg = 1.5 - log10(Re);
    double beta = 2.65 * (1.0 + pori) - (5.3 - 3.5 * pori) * pori * pori * exp(-lg * lg / 2.0);
    double k = (1.0 / 8.0) * sq * sq * rho_m * 3.14159 * di * di;
    double e = pow(pori, -1 - beta);
    dax = k * rmx * magm * e / mi;
    day = k * rmy * magm * e / mi;
    daz = k * rmz * magm * e / mi;
    const double Nu = 2.0 + 0.6 * pow(Re, 0.5) * (int_third ? 1.0 : cbrt(P.pr_air));
    dq_vol = Nu * ki / di * a_ht * ((tf + ta) / sm - Ti);
  }
  // the ISPH kernel divides the fluid drag by eps (:1099-1101), the WCSPH one does not (:338-340)
  const double dscale = P.wcsph ? 1.0 : pori;
  d.fx[i] = d.fx[i] + dfx / dscale + tx;
  d.fy[i] = d.fy[i] + dfy / dscale + ty;
  d.fz[i] = d.fz[i] + dfz / dscale + tz;
  d.fax[i] = dax;
  d.fay[i] = day;
  d.faz[i] = daz;
  d.fbx[i] = tx;
  d.fby[i] = ty;
  d.fbz[i] = tz;
  if (P.con_solve == 1) {
    const double conv = dq_vol / (rhoi * cpi);  // [K/s]
    d.dtemp[i] = (P.compat & SOPHIA_COMPAT_HEAT_OVERWRITE) ? conv : d.dtemp[i] + conv;
    d.qsd[i] = dq_vol * mi / rhoi;  // [W]
  }
}

// KERNEL_SPH_coupling3D_sph (functions_SPH_DEM_COUPLING2.cuh:1815-2019):
// drag reaction distributed onto the fluid with the sphere's Shepard sum.
template <typename R>
__global__ void k_sph_reaction(DevParams P, SphView<R> v, DemView d, CellView dc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // linear: work concentrates near walls/spheres
  if (i >= v.n) return;
  const int t = v.ptype[i];
  if (v.itype[i] != 1 || t > 1000) return;
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  const double hd = v.h[i], Rs = P.kappa * hd;
  double tx = 0, ty = 0, tz = 0, tt = 0;
  bool hit = false;
  if (d.n > 0 && t != SOPHIA_PT_MOVING && t != -3 && t != 0 && (t == 1 || t == 3)) {
    const double A = knl_tmpA<double>(P.kernel_type, P.dim, hd);
    const double rhoi = v.rho[i];
    // the WCSPH reaction carries the fluid particle's porosity
    // (KERNEL_SPH_coupling3D, :1677-1679); the ISPH one does not
    const double pfac = P.wcsph ? v.por[i] : 1.0;
    // heat_capacity(ptypei, tempi) at :1860 swaps its arguments: the
    // temperature is read as the particle type (SOPHIA_COMPAT_CP_ARGS)
    const double cpi = (P.compat & SOPHIA_COMPAT_CP_ARGS)
                           ? ((unsigned)(double)v.temp[i] > 1000u ? P.cp_dem : P.cp_fluid)
                           : (t > 1000 ? P.cp_dem : P.cp_fluid);
    for_cells(P, xi, yi, zi, dc, [&](int j) {
      if (!(d.ptype[j] > 1000)) return;
      double r = dist_exact(xi - d.x[j], yi - d.y[j], zi - d.z[j]);
      if (!(r < Rs)) return;
      double flt = d.flt_sd2[j] + d.flt_s[j];
      if (!(flt > 0.0)) return;
      double w = knl_w<double>(P.kernel_type, A, hd, r);
      double c = w / (flt + 1.0e-20);
      const double mr = P.wcsph ? d.m[j] * pfac / rhoi : d.m[j] / rhoi;
      tx += -mr * d.fdx[j] * c;
      ty += -mr * d.fdy[j] * c;
      tz += -mr * d.fdz[j] * c;
      tt += -d.qsd[j] / (rhoi * cpi) * c;
      hit = true;
    });
  }
  if (!hit) {
    // sum of nothing: ftotal and its magnitude are unchanged
    R fx = v.fx[i], fy = v.fy[i], fz = v.fz[i];
    v.fmag[i] = sqrt(fx * fx + fy * fy + fz * fz);
    return;
  }
  R fx = (R)((double)v.fx[i] + tx), fy = (R)((double)v.fy[i] + ty), fz = (R)((double)v.fz[i] + tz);
  v.fx[i] = fx;
  v.fy[i] = fy;
  v.fz[i] = fz;
  v.fmag[i] = sqrt(fx * fx + fy * fy + fz * fz);
  if (P.con_solve == 1) v.dtemp[i] = (R)((double)v.dtemp[i] + tt);
}

// KERNEL_clc_projection_dem (functions_TIME_ISPH2.cuh:214-308)
__global__ void k_dem_predictor(DevParams P, DemView d) {
  int i = particle_index();
  if (i >= d.n) return;
  if (d.itype[i] > 2) return;
  const double ux0 = d.ux[i], uy0 = d.uy[i], uz0 = d.uz[i];
  if (d.ptype[i] >= 0) {
    const double bz = (d.btype[i] == 0);
    d.ux[i] = ux0 + d.fx[i] * bz * P.dt_dem;
    d.uy[i] = uy0 + d.fy[i] * bz * P.dt_dem;
    d.uz[i] = uz0 + d.fz[i] * bz * P.dt_dem;
  }
  d.ux0[i] = ux0;
  d.uy0[i] = uy0;
  d.uz0[i] = uz0;
  if (P.con_solve == 1 && d.ptype[i] > 0) d.temp0[i] = d.temp[i];
}

// KERNEL_time_update_projection_dem (functions_TIME_ISPH2.cuh:624-776):
// spheres carry no pressure acceleration, so u = u*; x from the average.
__global__ void k_dem_corrector(DevParams P, DemView d) {
  int i = particle_index();
  if (i >= d.n) return;
  if (d.itype[i] > 2) return;
  const int t = d.ptype[i];
  double ux0, uy0, uz0;
  if (t > 0) {
    ux0 = d.ux[i]; uy0 = d.uy[i]; uz0 = d.uz[i];
  } else {
    ux0 = d.ux0[i]; uy0 = d.uy0[i]; uz0 = d.uz0[i];
  }
  double uxc = ux0, uyc = uy0, uzc = uz0;
  if ((uxc * uxc + uyc * uyc + uzc * uzc) >= P.u_limit * P.u_limit) {
    uxc = ux0; uyc = uy0; uzc = uz0;
  }
  const double el = t > 0 ? 1.0 : 0.0;
  d.x[i] = d.x[i] + (d.ux0[i] + uxc) / 2.0 * P.dt_dem * el;
  d.y[i] = d.y[i] + (d.uy0[i] + uyc) / 2.0 * P.dt_dem * el;
  d.z[i] = d.z[i] + (d.uz0[i] + uzc) / 2.0 * P.dt_dem * el;
  d.ux[i] = uxc;
  d.uy[i] = uyc;
  d.uz[i] = uzc;
  if (t > 1000 && !(P.compat & SOPHIA_COMPAT_NO_SPIN)) {
    // spin from the contact torque per inertia, as the WCSPH corrector does
    // (functions_TIME.cuh:753-755); the ISPH path leaves it out
    const double bz = (d.btype[i] == 0);
    d.wx[i] += d.tqx[i] * bz * P.dt_dem;
    d.wy[i] += d.tqy[i] * bz * P.dt_dem;
    d.wz[i] += d.tqz[i] * bz * P.dt_dem;
  }
  if (P.con_solve == 1 && t > 0) d.temp[i] = d.temp0[i] + d.dtemp[i] * P.dt_dem * (double)(t != -1);
}

// KERNEL_clc_predictor (functions_TIME.cuh:77-292) for the spheres in the
// WCSPH step: half step of position, velocity and spin from the previous
// step's accelerations; the start-of-step state goes to x0, ux0, wx0, temp0.
// Runs before the spheres are re-sorted.
__global__ void k_wc_dem_predictor(DevParams P, DemView d) {
  int i = particle_index();
  if (i >= d.n) return;
  if (d.itype[i] == 3) return;
  const double bz = (d.btype[i] == 0), hdt = P.dt_dem * 0.5;
  const double x0 = d.x[i], y0 = d.y[i], z0 = d.z[i], ux0 = d.ux[i], uy0 = d.uy[i], uz0 = d.uz[i];
  const double wx0 = d.wx[i], wy0 = d.wy[i], wz0 = d.wz[i];
  d.x[i] = x0 + ux0 * hdt;
  d.y[i] = y0 + uy0 * hdt;
  d.z[i] = z0 + uz0 * hdt;
  d.ux[i] = ux0 + d.fx[i] * bz * hdt;
  d.uy[i] = uy0 + d.fy[i] * bz * hdt;
  d.uz[i] = uz0 + d.fz[i] * bz * hdt;
  d.wx[i] = wx0 + d.tqx[i] * bz * hdt;
  d.wy[i] = wy0 + d.tqy[i] * bz * hdt;
  d.wz[i] = wz0 + d.tqz[i] * bz * hdt;
  d.x0[i] = x0; d.y0[i] = y0; d.z0[i] = z0;
  d.ux0[i] = ux0; d.uy0[i] = uy0; d.uz0[i] = uz0;
  d.wx0[i] = wx0; d.wy0[i] = wy0; d.wz0[i] = wz0;
  if (P.con_solve == 1) {
    d.temp0[i] = d.temp[i];
    d.temp[i] = d.temp[i] + d.dtemp[i] * hdt;
  }
}

// KERNEL_time_update_single (functions_TIME.cuh:676-877) for the spheres:
// u = u0 + a dt (held at u0 past u_limit), w = w0 + torque/inertia dt,
// x = x0 + u dt, T = T0 + dT dt.  The reference moves nothing when XSPH is
// on (:788-792), which freezes the spheres; here they always move.
__global__ void k_wc_dem_corrector(DevParams P, DemView d) {
  int i = particle_index();
  if (i >= d.n) return;
  if (d.itype[i] == 3) return;
  const double bz = (d.btype[i] == 0), dt = P.dt_dem;
  double uxc = d.ux0[i] + d.fx[i] * bz * dt, uyc = d.uy0[i] + d.fy[i] * bz * dt, uzc = d.uz0[i] + d.fz[i] * bz * dt;
  if (uxc * uxc + uyc * uyc + uzc * uzc >= P.u_limit * P.u_limit) {
    uxc = d.ux0[i];
    uyc = d.uy0[i];
    uzc = d.uz0[i];
  }
  d.x[i] = d.x0[i] + uxc * dt;
  d.y[i] = d.y0[i] + uyc * dt;
  d.z[i] = d.z0[i] + uzc * dt;
  d.ux[i] = uxc;
  d.uy[i] = uyc;
  d.uz[i] = uzc;
  d.wx[i] = d.wx0[i] + d.tqx[i] * bz * dt;
  d.wy[i] = d.wy0[i] + d.tqy[i] * bz * dt;
  d.wz[i] = d.wz0[i] + d.tqz[i] * bz * dt;
  if (P.con_solve == 1) d.temp[i] = d.temp0[i] + d.dtemp[i] * dt;
}

// --- DEM wall boundaries (functions_DEM_BC.cuh) -----------------------------
// State of one sphere while the walls act on it in turn.
struct WallBody {
  double x, y, z, ux, uy, uz, wx, wy, wz;  // position, velocity, angular velocity
  double fx, fy, fz, tx, ty, tz;            // ftotal (per mass), torque
  double m, ri, rad;
};

// Contact response shared by the three walls (the common tail of
// KERNEL_treat_DEM_{cylinder,box,truncated_cone}_z_dem, e.g. :609-725):
// n is the unit normal pointing into the container.  Slow normal approach
// (|u.n| < 50 dt): sliding friction mu_s m (f.n) along the contact slip and
// rolling friction mu_rs m (f.n) along the spin slip, none when the wall
// pulls (f.n > 0); the normal velocity is removed.  Otherwise an impulse
// with restitution rest_n (normal) and rest_s (tangential).
__device__ inline void wall_response(const DemWallParams& W, double dt, double nx, double ny, double nz,
                                     WallBody& b) {
  const double rad = b.rad, m = b.m, ri = b.ri;
  const double un = b.ux * nx + b.uy * ny + b.uz * nz;
  // velocity of the contact point and its spin part
  const double cwx = -rad * (b.wy * nz - b.wz * ny);
  const double cwy = -rad * (b.wz * nx - b.wx * nz);
  const double cwz = -rad * (b.wx * ny - b.wy * nx);
  const double cx = b.ux + cwx, cy = b.uy + cwy, cz = b.uz + cwz;
  const double cn = cx * nx + cy * ny + cz * nz;
  const double cwn = cwx * nx + cwy * ny + cwz * nz;
  double sx = cx - cn * nx, sy = cy - cn * ny, sz = cz - cn * nz;       // slip
  double qx = cwx - cwn * nx, qy = cwy - cwn * ny, qz = cwz - cwn * nz;  // spin slip
  sx = fabs(sx) < 1e-10 ? 0.0 : sx;
  sy = fabs(sy) < 1e-10 ? 0.0 : sy;
  sz = fabs(sz) < 1e-10 ? 0.0 : sz;
  qx = fabs(qx) < 1e-10 ? 0.0 : qx;
  qy = fabs(qy) < 1e-10 ? 0.0 : qy;
  qz = fabs(qz) < 1e-10 ? 0.0 : qz;
  const double smag = sqrt(sx * sx + sy * sy + sz * sz);
  const double qmag = sqrt(qx * qx + qy * qy + qz * qz);
  const double ex = sx / (smag + 1e-20), ey = sy / (smag + 1e-20), ez = sz / (smag + 1e-20);
  const double gx = qx / (qmag + 1e-20), gy = qy / (qmag + 1e-20), gz = qz / (qmag + 1e-20);
  const double k = rad / ri;
  if (fabs(un) < 50.0 * dt) {
    const double fn = b.fx * nx + b.fy * ny + b.fz * nz;
    const bool on = !(fn > 0.0);
    const double a = on ? W.mu_s * m * fn : 0.0, r = on ? W.mu_rs * m * fn : 0.0;
    const double ffx = a * ex, ffy = a * ey, ffz = a * ez;
    const double tfx = ffx + r * gx, tfy = ffy + r * gy, tfz = ffz + r * gz;
    b.fx += ffx / m;
    b.fy += ffy / m;
    b.fz += ffz / m;
    const double mx = -k * (ny * tfz - nz * tfy), my = -k * (nz * tfx - nx * tfz), mz = -k * (nx * tfy - ny * tfx);
    b.tx += mx;
    b.ty += my;
    b.tz += mz;
    const double vx = b.ux + (ffx / m) * dt, vy = b.uy + (ffy / m) * dt, vz = b.uz + (ffz / m) * dt;
    const double vn = vx * nx + vy * ny + vz * nz;
    b.ux = vx - vn * nx;
    b.uy = vy - vn * ny;
    b.uz = vz - vn * nz;
    b.wx += mx * dt;
    b.wy += my * dt;
    b.wz += mz * dt;
  } else {
    const double jn = -m * (1.0 + W.rest_n) * cn;
    const double js = -1.0 / ((1.0 / m) + (rad * rad / ri)) * (1.0 + W.rest_s) * smag;
    const double jx = jn * nx + js * ex, jy = jn * ny + js * ey, jz = jn * nz + js * ez;
    b.ux += (1.0 / m) * jx;
    b.uy += (1.0 / m) * jy;
    b.uz += (1.0 / m) * jz;
    b.wx -= k * (ny * jz - nz * jy);
    b.wy -= k * (nz * jx - nx * jz);
    b.wz -= k * (nx * jy - ny * jx);
  }
}

// KERNEL_treat_DEM_cylinder_z_dem (:491-729): a sphere beyond the vertical
// cylinder of radius R about (xc, yc) is moved onto the wall radially.
__device__ inline void wall_cylinder(const DemWallParams& W, double dt, WallBody& b) {
  const double xc = W.cyl[0], yc = W.cyl[1], R = W.cyl[2], rad = b.rad;
  const double rx = b.x - xc, ry = b.y - yc;
  if (rx * rx + ry * ry - (R - rad) * (R - rad) <= 0.0) return;  // inside
  const double d = sqrt(rx * rx + ry * ry);
  const double cpx = xc + (R - rad) * rx / d, cpy = yc + (R - rad) * ry / d;
  b.x = cpx;
  b.y = cpy;
  const double nx = -(cpx - xc) / (R - rad), ny = -(cpy - yc) / (R - rad), nz = 0.0;
  if (sqrt(nx * nx + ny * ny + nz * nz) < 0.5) return;
  wall_response(W, dt, nx, ny, nz, b);
}

// KERNEL_treat_DEM_box_z_dem (:1770-2027): a sphere outside the box is
// clamped onto it; the normal is the sign pattern of the move.  The x sign
// is cleared on |cp_y - x| < 1e-12 as in the reference (:1877), so a box
// centred off x = 0 lets the rounding of xc + (x - xc) into the normal.
__device__ inline void wall_box(const DemWallParams& W, double dt, WallBody& b) {
  const double xc = W.box[0], yc = W.box[1], zc = W.box[2], rad = b.rad;
  const double hx = 0.5 * W.box[3] - rad, hy = 0.5 * W.box[4] - rad, hz = 0.5 * W.box[5] - rad;
  const double x = b.x, y = b.y, z = b.z;
  const double F = fmax(fmax(fabs(x - xc) - hx, fabs(y - yc) - hy), fabs(z - zc) - hz);
  if (F <= 0.0) return;  // inside
  const double cpx = xc + fmin(hx, fmax(-hx, x - xc));
  const double cpy = yc + fmin(hy, fmax(-hy, y - yc));
  const double cpz = zc + fmin(hz, fmax(-hz, z - zc));
  b.x = cpx;
  b.y = cpy;
  b.z = cpz;
  double sx = (cpx - x) / (fabs(cpx - x) + 1e-20);
  double sy = (cpy - y) / (fabs(cpy - y) + 1e-20);
  double sz = (cpz - z) / (fabs(cpz - z) + 1e-20);
  if (fabs(cpy - y) < 1e-12) sy = 0.0;
  if (fabs(cpy - x) < 1e-12) sx = 0.0;
  if (fabs(cpz - z) < 1e-12) sz = 0.0;
  const double sm = sqrt(sx * sx + sy * sy + sz * sz) + 1e-20;
  const double nx = sx / sm, ny = sy / sm, nz = sz / sm;
  if (sqrt(nx * nx + ny * ny + nz * nz) < 0.5) return;
  wall_response(W, dt, nx, ny, nz, b);
}

// KERNEL_treat_DEM_truncated_cone_z_dem (:4485-4725): the cone narrows from
// radius r1 to r2 over height h down to z = -h; a sphere above z = rad - h
// and outside the cone surface is moved onto it vertically.
__device__ inline void wall_cone(const DemWallParams& W, double dt, WallBody& b) {
  const double xc = W.cone[0], yc = W.cone[1], r1 = W.cone[2], r2 = W.cone[3], h = W.cone[4];
  const double rad = b.rad, x = b.x, y = b.y, z = b.z;
  if (z <= -h + rad) return;
  const double L = sqrt((r1 - r2) * (r1 - r2) + h * h);
  const double sn = h / L, cs = fabs(r1 - r2) / L, tn = sn / cs;
  const double rr = r2 + (h + z) / tn - rad / sn;
  if ((x - xc) * (x - xc) + (y - yc) * (y - yc) - rr * rr <= 0.0) return;  // inside
  const double d = sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc));
  b.z = tn * (d + rad / sn - r2) - h;
  const double nx = -sn * (x - xc) / d, ny = -sn * (y - yc) / d, nz = cs;
  wall_response(W, dt, nx, ny, nz, b);
}

// Walls in the reference's order, one thread per sphere (i_type 1 only).
__global__ void k_dem_walls(DevParams P, DemView d, DemWallParams W) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  if (d.itype[i] != 1 || d.ptype[i] <= 1000) return;
  WallBody b{d.x[i],  d.y[i],  d.z[i],  d.ux[i],  d.uy[i],  d.uz[i],  d.wx[i], d.wy[i], d.wz[i],
             d.fx[i], d.fy[i], d.fz[i], d.tqx[i], d.tqy[i], d.tqz[i], d.m[i],  d.ri[i], d.rad[i]};
  if (W.mask & 1) wall_cylinder(W, P.dt_dem, b);
  if (W.mask & 2) wall_box(W, P.dt_dem, b);
  if (W.mask & 4) wall_cone(W, P.dt_dem, b);
  d.x[i] = b.x; d.y[i] = b.y; d.z[i] = b.z;
  d.ux[i] = b.ux; d.uy[i] = b.uy; d.uz[i] = b.uz;
  d.wx[i] = b.wx; d.wy[i] = b.wy; d.wz[i] = b.wz;
  d.fx[i] = b.fx; d.fy[i] = b.fy; d.fz[i] = b.fz;
  d.tqx[i] = b.tx; d.tqy[i] = b.ty; d.tqz[i] = b.tz;
}

}  // namespace

int dem_walls(hipStream_t s, const DevParams& P, const DemView& d, const DemWallParams& W) {
  if (d.n <= 0 || W.mask == 0) return 0;
  k_dem_walls<<<grid_for(d.n), kBlock, 0, s>>>(P, d, W);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <typename R>
int dem_porosity(hipStream_t s, const DevParams& P, const DemView& d, const SphView<R>& v, const CellView& sc,
                 const CellView& dc) {
  if (d.n <= 0) return 0;
  k_porosity_dem<R><<<(d.n + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, s>>>(P, d, v, sc, dc);
  HIP_TRY(hipGetLastError());
  return 0;
}

int dem_contact(hipStream_t s, const DevParams& P, const DemView& d, const CellView& dc) {
  if (d.n <= 0) return 0;
  k_dem_contact<<<grid_xcd(d.n), kBlock, 0, s>>>(P, d, dc);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <typename R>
int dem_coupling(hipStream_t s, const DevParams& P, const DemView& d, const SphView<R>& v, const CellView& sc) {
  if (d.n <= 0) return 0;
  k_dem_coupling<R><<<(d.n + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, s>>>(P, d, v, sc);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <typename R>
int sph_dem_reaction(hipStream_t s, const DevParams& P, const SphView<R>& v, const DemView& d, const CellView& dc) {
  if (v.n <= 0) return 0;
  k_sph_reaction<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v, d, dc);
  HIP_TRY(hipGetLastError());
  return 0;
}

int dem_predictor(hipStream_t s, const DevParams& P, const DemView& d) {
  if (d.n <= 0) return 0;
  k_dem_predictor<<<grid_xcd(d.n), kBlock, 0, s>>>(P, d);
  HIP_TRY(hipGetLastError());
  return 0;
}

int wc_dem_predictor(hipStream_t s, const DevParams& P, const DemView& d) {
  if (d.n <= 0) return 0;
  k_wc_dem_predictor<<<grid_xcd(d.n), kBlock, 0, s>>>(P, d);
  HIP_TRY(hipGetLastError());
  return 0;
}
int wc_dem_corrector(hipStream_t s, const DevParams& P, const DemView& d) {
  if (d.n <= 0) return 0;
  k_wc_dem_corrector<<<grid_xcd(d.n), kBlock, 0, s>>>(P, d);
  HIP_TRY(hipGetLastError());
  return 0;
}
int dem_corrector(hipStream_t s, const DevParams& P, const DemView& d) {
  if (d.n <= 0) return 0;
  k_dem_corrector<<<grid_xcd(d.n), kBlock, 0, s>>>(P, d);
  HIP_TRY(hipGetLastError());
  return 0;
}

#define INST(R)                                                                                                    \
  template int dem_porosity<R>(hipStream_t, const DevParams&, const DemView&, const SphView<R>&, const CellView&,  \
                               const CellView&);                                                                   \
  template int dem_coupling<R>(hipStream_t, const DevParams&, const DemView&, const SphView<R>&, const CellView&); \
  template int sph_dem_reaction<R>(hipStream_t, const DevParams&, const SphView<R>&, const DemView&, const CellView&);
INST(float)
INST(double)

}  // namespace sophia
