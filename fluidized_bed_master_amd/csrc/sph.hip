// SPH stage kernels of the ISPH step on gfx950.
//
// One thread per particle.  Pair loops walk the neighbour list built at
// NNPS (slot-major, so the k-th neighbour of 64 consecutive particles is one
// coalesced 256 B read); the list keeps the reference's cell visiting order,
// so sums are accumulated in the same order as the reference kernels.
// Neighbour fields are gathered from SoA streams of cell-sorted particles;
// spatial locality of the sort keeps those gathers in L2.
//
// Each kernel cites the reference kernel it computes.
#include "common.hpp"
#include "nbr.hpp"

namespace sophia {

namespace {

// KERNEL_clc_mass_init_sph (functions_MASS.cuh:24-43): vol0 = vol =
// (h/1.6)^dim, rho0 <- rho
template <typename R>
__global__ void k_mass_init(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.ptype[i] > 1000 || v.itype[i] > 2) return;
  const double vol0 = pow(v.h[i] / P.h_over_dx, (double)P.dim);
  v.vol0[i] = vol0;
  v.vol[i] = vol0;
  v.rho0[i] = v.rho[i];
}

// rho_j / rho_ref_j - rho_i / rho_ref_i (the Antuono difference) in FP64:
// from the fields, or from the record's rho_j / rho_ref_j - 1 (FP32, exact
// to 2^-24 of the difference's own size, not of 1)
template <typename R>
__device__ inline double ratio_diff(const Nb<R>& q, double ratio_i, double) {
  return q.rho / q.rref - ratio_i;
}
__device__ inline double ratio_diff(const NbR& q, double, double ratio_i1) { return (double)q.aux - ratio_i1; }

// The 32 B records of the FP32 pair kernels (for_nbrs4_r32), and with VEL the
// velocity records; AUX: rho / rho_ref - 1 for the prep sums
template <bool AUX, bool VEL>
__global__ void k_pack_r32(SphView<float> v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  Rec32 r;
  r.p = v.pos4[i];
  r.m = (float)v.m[i];
  r.rho = (float)v.rho[i];
  r.pt = v.ptype[i];
  r.aux = AUX ? (float)(v.rho[i] / v.rho_ref[i] - 1.0) : 0.f;
  v.r32[i] = r;
  if constexpr (VEL) v.pkv[i] = make_float4(v.ux[i], v.uy[i], v.uz[i], v.temp[i]);
}

// before an FP32 pair kernel (the FP64 path gathers its fields directly)
template <typename R>
int pack_r32(hipStream_t s, const SphView<R>& v, bool aux, bool vel) {
  if constexpr (std::is_same<R, float>::value) {
    if (v.n <= 0) return 0;
    const int g = grid_for(v.n);
    if (vel) k_pack_r32<false, true><<<g, kBlock, 0, s>>>(v);
    else if (aux) k_pack_r32<true, false><<<g, kBlock, 0, s>>>(v);
    else k_pack_r32<false, false><<<g, kBlock, 0, s>>>(v);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}

// KERNEL_clc_correction_KGC_3D_sph (functions_PREP2.cuh:452-568)
template <typename R>
__global__ __launch_bounds__(kBlock) void k_kgc(DevParams P, SphView<R> v) {
  typedef R T;
  int i = particle_index();
  if (i >= v.n) return;
  const int n = v.n;
  if (v.itype[i] > 2 || v.ptype[i] > 1000) {
    for (int c = 0; c < 6; c++) v.cm[c * n + i] = 0;
    return;
  }
  const double hd = v.h[i];
  const T h = (T)hd, ih = (T)(1.0 / hd), A = knl_tmpA<T>(P.kernel_type, P.dim, h), Rs = (T)(P.kappa * hd);
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  T xx = 0, yy = 0, zz = 0, xy = 0, yz = 0, zx = 0;
  auto term = [&](const auto& q, bool on) {
    T dx = (T)q.dx, dy = (T)q.dy, dz = (T)q.dz;
    T r = sqrt(dx * dx + dy * dy + dz * dz);
    const bool use = on && q.pt <= 1000 && r > 0 && r < Rs;
    T dw = knl_dw_ih<T>(P.kernel_type, A, h, ih, r);
    T mtd = (T)q.m * dw;
    T rtd = (T)1 / ((T)q.rho * r);
    xx += use ? -mtd * dx * dx * rtd : (T)0;
    xy += use ? -mtd * dy * dx * rtd : (T)0;
    yy += use ? -mtd * dy * dy * rtd : (T)0;
    zx += use ? -mtd * dx * dz * rtd : (T)0;
    yz += use ? -mtd * dy * dz * rtd : (T)0;
    zz += use ? -mtd * dz * dz * rtd : (T)0;
  };
  if constexpr (std::is_same<R, float>::value) for_nbrs4_r32<false>(v, i, false, term);
  else for_nbrs4<true>(v, i, false, term);
  T det = xx * (yy * zz - yz * yz);
  det -= xy * (xy * zz - yz * zx);
  det += zx * (xy * yz - yy * zx);
  T c0, c1, c2, c3, c4, c5;
  if (fabs((double)det) > 1e-4) {
    T rr = (T)1 / det;
    c0 = (yy * zz - yz * yz) * rr;
    c1 = (zx * yz - xy * zz) * rr;
    c2 = (xy * yz - zx * yy) * rr;
    c3 = (xx * zz - zx * zx) * rr;
    c4 = (zx * xy - xx * yz) * rr;
    c5 = (xx * yy - xy * xy) * rr;
  } else {
    c0 = 1; c1 = 0; c2 = 0; c3 = 1; c4 = 0; c5 = 1;
  }
  v.cm[i] = c0; v.cm[n + i] = c1; v.cm[2 * n + i] = c2;
  v.cm[3 * n + i] = c3; v.cm[4 * n + i] = c4; v.cm[5 * n + i] = c5;
}

// KERNEL_clc_prep3D_coupling_sph (functions_PREP2.cuh:3190-3364): porosity
// eps_i = 1 - sum_a V_a W_ia / sum_{j fluid} V_j W_ij and its rate.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_porosity_sph(DevParams P, SphView<R> v, DemView d, CellView dc) {
  typedef R T;
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // linear: work concentrates near walls/spheres
  if (i >= v.n) return;
  if (v.itype[i] > 2) return;
  v.dpor_prev[i] = v.dpor[i];
  const double hd = v.h[i];
  const T h = (T)hd, A = knl_tmpA<T>(P.kernel_type, P.dim, h);
  const double Rs = P.kappa * hd;
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  const int pti = v.ptype[i];
  T flt = 0, por = 0, dfx = 0, dfy = 0, dfz = 0;
  auto term = [&](const auto& q, bool on) {
    T dx = (T)q.dx, dy = (T)q.dy, dz = (T)q.dz;
    T r = sqrt(dx * dx + dy * dy + dz * dz) + (T)1e-20;
    T w = knl_w<T>(P.kernel_type, A, h, r);
    int ptj = q.pt;
    const bool use = on && pti <= 1000 && ptj <= 1000 && ptj == 1;
    flt += use ? (T)q.m / (T)q.rho * w : (T)0;
  };
  if constexpr (std::is_same<R, float>::value) for_nbrs4_r32<false>(v, i, true, term);
  else for_nbrs4<true>(v, i, true, term);
  if (d.n > 0) {
    for_cells(P, xi, yi, zi, dc, [&](int j) {
      double rd = dist_exact(xi - d.x[j], yi - d.y[j], zi - d.z[j]) + 1e-20;
      if (!(rd < Rs)) return;
      if (!(pti <= 1000 && d.ptype[j] > 1000)) return;
      T dx = (T)(xi - d.x[j]), dy = (T)(yi - d.y[j]), dz = (T)(zi - d.z[j]);
      T r = (T)rd;
      T w = knl_w<T>(P.kernel_type, A, h, r);
      T dw = knl_dw<T>(P.kernel_type, A, h, r);
      T rj = (T)d.rad[j];
      T Vj = (T)4.18879 * rj * rj * rj;
      por += w * Vj;
      dfx += dw * dx / r * (T)d.ux[j] * Vj;
      dfy += dw * dy / r * (T)d.uy[j] * Vj;
      dfz += dw * dz / r * (T)d.uz[j] * Vj;
    });
  }
  if (pti <= 1000) {
    v.por[i] = (T)1 - (por / (flt + (T)1e-20));
    v.dpor[i] = (dfx + dfy + dfz) / (flt + (T)1e-20);
  }
}

// k_kgc and k_porosity_sph in one pass over the neighbour list (mixed path):
// the same 32 B records, each kernel's own terms, neighbour rules (KGC takes
// every list entry, the porosity only those without the eps flag) and
// arithmetic, so the results are those of the two kernels.
template <bool KGC>
__global__ __launch_bounds__(kBlock) void k_kgc_porosity(DevParams P, SphView<float> v, DemView d, CellView dc) {
  typedef float T;
  const int i = particle_index();
  if (i >= v.n) return;
  const int n = v.n;
  const int iti = v.itype[i], pti = v.ptype[i];
  const bool krow = KGC && !(iti > 2 || pti > 1000);
  if (KGC && !krow)
    for (int c = 0; c < 6; c++) v.cm[c * n + i] = 0;
  if (iti > 2) return;
  v.dpor_prev[i] = v.dpor[i];
  const double hd = v.h[i];
  const T h = (T)hd, ih = (T)(1.0 / hd), A = knl_tmpA<T>(P.kernel_type, P.dim, h), Rk = (T)(P.kappa * hd);
  const double Rs = P.kappa * hd;
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  T xx = 0, yy = 0, zz = 0, xy = 0, yz = 0, zx = 0;
  T flt = 0, por = 0, dfx = 0, dfy = 0, dfz = 0;
  for_nbrs4_r32<false>(v, i, false, [&](const NbR& q, bool on) {
    const T dx = q.dx, dy = q.dy, dz = q.dz;
    const T r = sqrt(dx * dx + dy * dy + dz * dz);
    if (krow) {
      const bool use = on && q.pt <= 1000 && r > 0 && r < Rk;
      T dw = knl_dw_ih<T>(P.kernel_type, A, h, ih, r);
      T mtd = q.m * dw;
      T rtd = (T)1 / (q.rho * r);
      xx += use ? -mtd * dx * dx * rtd : (T)0;
      xy += use ? -mtd * dy * dx * rtd : (T)0;
      yy += use ? -mtd * dy * dy * rtd : (T)0;
      zx += use ? -mtd * dx * dz * rtd : (T)0;
      yz += use ? -mtd * dy * dz * rtd : (T)0;
      zz += use ? -mtd * dz * dz * rtd : (T)0;
    }
    const T rp = r + (T)1e-20;
    const T w = knl_w<T>(P.kernel_type, A, h, rp);
    const bool usep = on && !q.flag && pti <= 1000 && q.pt <= 1000 && q.pt == 1;
    flt += usep ? q.m / q.rho * w : (T)0;
  });
  if (krow) {
    T det = xx * (yy * zz - yz * yz);
    det -= xy * (xy * zz - yz * zx);
    det += zx * (xy * yz - yy * zx);
    T c0, c1, c2, c3, c4, c5;
    if (fabs((double)det) > 1e-4) {
      T rr = (T)1 / det;
      c0 = (yy * zz - yz * yz) * rr;
      c1 = (zx * yz - xy * zz) * rr;
      c2 = (xy * yz - zx * yy) * rr;
      c3 = (xx * zz - zx * zx) * rr;
      c4 = (zx * xy - xx * yz) * rr;
      c5 = (xx * yy - xy * xy) * rr;
    } else {
      c0 = 1; c1 = 0; c2 = 0; c3 = 1; c4 = 0; c5 = 1;
    }
    v.cm[i] = c0; v.cm[n + i] = c1; v.cm[2 * n + i] = c2;
    v.cm[3 * n + i] = c3; v.cm[4 * n + i] = c4; v.cm[5 * n + i] = c5;
  }
  if (d.n > 0) {
    for_cells(P, xi, yi, zi, dc, [&](int j) {
      double rd = dist_exact(xi - d.x[j], yi - d.y[j], zi - d.z[j]) + 1e-20;
      if (!(rd < Rs)) return;
      if (!(pti <= 1000 && d.ptype[j] > 1000)) return;
      T dx = (T)(xi - d.x[j]), dy = (T)(yi - d.y[j]), dz = (T)(zi - d.z[j]);
      T r = (T)rd;
      T w = knl_w<T>(P.kernel_type, A, h, r);
      T dw = knl_dw<T>(P.kernel_type, A, h, r);
      T rj = (T)d.rad[j];
      T Vj = (T)4.18879 * rj * rj * rj;
      por += w * Vj;
      dfx += dw * dx / r * (T)d.ux[j] * Vj;
      dfy += dw * dy / r * (T)d.uy[j] * Vj;
      dfz += dw * dz / r * (T)d.uz[j] * Vj;
    });
  }
  if (pti <= 1000) {
    v.por[i] = (T)1 - (por / (flt + (T)1e-20));
    v.dpor[i] = (dfx + dfy + dfz) / (flt + (T)1e-20);
  }
}

// KERNEL_clc_prep3D_prep_sph (functions_PREP2.cuh:1623-1679)
template <typename R>
__global__ void k_prep_prep(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2 || v.ptype[i] > 1000) return;
  int t = v.ptype[i];
  double hv = v.h[i] / P.h_over_dx;
  double rt = (t == -3 || t == 0 || t == 3 || t == 1) ? P.rho_fluid : P.rho_water;
  if (t == 9 && v.z[i] < 0.05) rt = P.rho_fluid;
  double m_ref = rt * hv * hv * hv * (double)v.por[i];
  double V = P.dim == 2 ? hv * hv : hv * hv * hv;
  v.rho_ref[i] = m_ref / V;
  if (P.rho_type == 1) {
    v.rho0[i] = m_ref / (hv * hv * hv);
    v.rho[i] = m_ref / (hv * hv * hv);
    v.m[i] = m_ref;
  }
}

// KERNEL_clc_prep3D_sph (functions_PREP2.cuh:1985-2227): Shepard filter and
// the Antuono density gradient.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_prep(DevParams P, SphView<R> v) {
  typedef R T;
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2 || v.ptype[i] > 1000) return;
  const int n = v.n;
  const double hd = v.h[i];
  const T h = (T)hd, ih = (T)(1.0 / hd), A = knl_tmpA<T>(P.kernel_type, P.dim, h);
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  // rho / rho_ref differs from 1 by ~drho dt / rho: the difference of two
  // such ratios cancels to a few significant digits, so it is taken in FP64
  const double ratio_i = v.rho[i] / v.rho_ref[i], ratio_i1 = ratio_i - 1.0;
  const bool anto = P.delsph == 2;
  T flt = 0, rx = 0, ry = 0, rz = 0;
  auto term = [&](const auto& q, bool on) {
    const bool use = on && q.pt <= 1000;
    T dx = (T)q.dx, dy = (T)q.dy, dz = (T)q.dz;
    T r = sqrt(dx * dx + dy * dy + dz * dz) + (T)1e-20;
    const T Vj = (T)q.m / (T)q.rho;
    if (P.filt_step) flt += use ? Vj * knl_w_ih<T>(P.kernel_type, A, h, ih, r) : (T)0;
    if (anto) {
      T dw = knl_dw_ih<T>(P.kernel_type, A, h, ih, r);
      const T sd = dw / r;
      T gx = sd * dx, gy = sd * dy, gz = sd * dz;
      kgc_apply(P, v.cm, n, i, gx, gy, gz);
      T c = (T)(-ratio_diff(q, ratio_i, ratio_i1)) * Vj;
      rx += use ? c * gx : (T)0;
      ry += use ? c * gy : (T)0;
      rz += use ? c * gz : (T)0;
    }
  };
  if constexpr (std::is_same<R, float>::value) for_nbrs4_r32<false>(v, i, true, term);
  else for_nbrs4<true>(v, i, true, term);
  if (P.filt_step) v.flt_s[i] = flt;
  if (anto) {
    v.grx[i] = rx;
    v.gry[i] = ry;
    v.grz[i] = rz;
  }
}

// Flux walls (sophia_params.heat_flux): each wall particle inside the flux
// box sums V_k W(r, h_w) over the fluid of its neighbour list.  The fluid
// particle i then receives q A_w W_iw / S_w from every flux wall w in its
// list, so the walls' power q A_w reaches the fluid in full.
__device__ inline bool flux_receiver(const DevParams& P, int pt, int it, int bt) {
  return it <= 2 && !(P.open_boundary > 0 && bt > 0) && pt > 0 && pt < 1000;
}
template <typename R>
__global__ void k_flux_walls(DevParams P, SphView<R> v) {
  const int i = particle_index();
  if (i >= v.n) return;
  double S = 0;
  const double x = v.x[i], y = v.y[i], z = v.z[i];
  const double* b = P.heat_flux_box;
  if (v.ptype[i] == 0 && x >= b[0] && y >= b[1] && z >= b[2] && x <= b[3] && y <= b[4] && z <= b[5]) {
    const double h = v.h[i], A = knl_tmpA<double>(P.kernel_type, P.dim, h);
    for_nbrs4<false>(v, i, false, [&](const Nb<R>& q, bool on) {
      const bool use = on && flux_receiver(P, q.pt, v.itype[q.j], v.btype[q.j]);
      const double r = sqrt(q.dx * q.dx + q.dy * q.dy + q.dz * q.dz);
      S += use ? q.m / q.rho * knl_w<double>(P.kernel_type, A, h, r) : 0.0;
    });
  }
  v.flux_s[i] = S;
}

// viscosity() and conductivity() of every SPH particle at its temperature
// (functions_PROP.cuh:228-264) when property tables are on: one table
// interpolation per particle, read by the pair sums for both ends of a pair
template <typename R>
__global__ void k_props(DevParams P, SphView<R> v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  const double T = (double)v.temp[i];
  v.visp[i] = (R)prop_mu(P, v.ptype[i], T);
  v.condp[i] = (R)prop_k(P, v.ptype[i], T);
}

// The packed neighbour records of the FP64-geometry pair sums (for_nbrs4_pk),
// rebuilt from the SoA fields before each kernel that reads them.
template <typename R, bool VEL>
__global__ void k_pack_records(SphView<R> v, bool flux) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  v.pk[i] = make_double4(v.x[i], v.y[i], v.z[i], v.m[i]);
  if constexpr (VEL && std::is_same<R, float>::value) {
    PkB b;
    b.rho = v.rho[i];
    b.p = v.pres[i];
    b.ux = v.ux[i]; b.uy = v.uy[i]; b.uz = v.uz[i]; b.temp = v.temp[i];
    v.pkb[i] = b;
    PkC c;
    c.S = flux ? v.flux_s[i] : 0.0;
    c.h = (float)v.h[i];
    c.pt = v.ptype[i];
    v.pkc[i] = c;
    return;
  }
  v.pk2[i] = make_double2(v.rho[i], v.pres[i]);
  if constexpr (VEL) {
    typename Vec4Of<R>::type w;
    w.x = v.ux[i]; w.y = v.uy[i]; w.z = v.uz[i]; w.w = v.temp[i];
    v.pkv[i] = w;
    HRec<R> hr;
    hr.h = (R)v.h[i];
    hr.pt = v.ptype[i];
    v.pkh[i] = hr;
  }
}

// Pair sums in FP64 geometry read the packed records: the symmetric pressure
// gradient, and every sum of the FP64 path.  The FP32 difference form takes
// its geometry from the position records instead.
template <typename R>
__device__ __host__ inline bool packed_geometry(const DevParams& P) {
  return !P.pgrad_form || sizeof(R) == 8;
}

// KERNEL_advection_force3D_sph (functions_PPE2.cuh:1251-1607): viscous,
// artificial viscosity, Eulerian advection, conduction, gravity and the
// pressure-gradient force density used by the coupling.
template <typename R, bool FLUX>
__global__ __launch_bounds__(kBlock) void k_advection(DevParams P, SphView<R> v) {
  typedef R T;
  int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] > 2) return;
  if (P.open_boundary > 0 && v.btype[i] > 0) return;
  if (pti >= 1000 || pti <= 0) return;
  const int n = v.n;
  const double hd = v.h[i], ihd = 1.0 / hd;
  const T h = (T)hd;
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  const T uxi = v.ux[i], uyi = v.uy[i], uzi = v.uz[i];
  const double pid = v.pres[i];
  const T mi = (T)v.m[i], rhoi = (T)v.rho[i], pori = (T)v.por[i], tempi = v.temp[i];
  const bool tab = P.prop_table != 0;
  const T visi = tab ? v.visp[i] : (T)P.mu_fluid, kci = tab ? v.condp[i] : (T)(pti > 1000 ? P.k_dem : P.k_fluid);
  const T cpi = (T)(pti > 1000 ? P.cp_dem : P.cp_fluid);
  const T eli = P.eulerian ? (T)0 : (T)1;
  const bool euler = eli < (T)1;
  const double rhoid = v.rho[i];
  const double Ad = knl_tmpA<double>(P.kernel_type, P.dim, hd);
  // per-particle factors of the pair terms, hoisted: every neighbour in the
  // SPH list has p_type <= 1000, so without tables its viscosity and
  // conductivity are the fluid constants
  const T visf = (T)P.mu_fluid, kcf = (T)P.k_fluid;
  const T vis_const = (visi * visf) / (visi + visf);
  const T cv_i = 4 * pori / rhoi;
  const T kc_const = (T)4 * pori * kcf * kci / (rhoi * (kci + kcf));
  T tx = 0, ty = 0, tz = 0, Rc = 0, ex = 0, ey = 0, ez = 0, et = 0;
  double px = 0, py = 0, pz = 0, fl = 0;
  auto term = [&](const auto& q, bool on) {
    const bool use = on && q.pt < 1000;
    if constexpr (FLUX) {
      // wall heat flux from a flux wall j: q A_j W_ij / S_j
      double S;
      if constexpr (std::is_same<std::decay_t<decltype(q)>, NbP<float>>::value) S = q.S;
      else S = v.flux_s[q.j];
      if (use && S > 0.0) {
        const double rw = sqrt(q.dx * q.dx + q.dy * q.dy + q.dz * q.dz), a = v.h[q.j] / P.h_over_dx;
        fl += P.heat_flux * (a * a) * knl_w_ih<double>(P.kernel_type, Ad, hd, ihd, rw) / S;
      }
    }
    // geometry in FP64 arithmetic: the symmetric pressure-gradient term
    // cancels a hydrostatic offset of order H/h times its result.  Two
    // reciprocals per pair (1/r, 1/rho_j) replace the divisions of the
    // reference's expressions; the rounding moves by an ulp
    const double dxd = q.dx, dyd = q.dy, dzd = q.dz;
    const double rd = sqrt(dxd * dxd + dyd * dyd + dzd * dzd) + 1e-20;
    const double ird = 1.0 / rd, irhod = 1.0 / q.rho;
    const double dwd = knl_dw_ih<double>(P.kernel_type, Ad, hd, ihd, rd);
    const double sd = dwd * ird;
    double gxd = sd * dxd, gyd = sd * dyd, gzd = sd * dzd;
    kgc_apply(P, v.cm, n, i, gxd, gyd, gzd);
    if (P.fp_solve) {
      // (rho_i / eps_i) Cp with Cp = -eps_i m_j pp / (rho_i rho_j)
      const double pp = P.pgrad_form ? q.pres - pid : pid + q.pres;
      const double Vpp = q.m * pp * irhod;
      px -= use ? Vpp * gxd : 0.0;
      py -= use ? Vpp * gyd : 0.0;
      pz -= use ? Vpp * gzd : 0.0;
    }
    const T dx = (T)dxd, dy = (T)dyd, dz = (T)dzd, r = (T)rd, dw = (T)dwd, ir = (T)ird;
    const T gx = (T)gxd, gy = (T)gyd, gz = (T)gzd;
    const T mj = (T)q.m, irhoj = (T)irhod, Vj = mj * irhoj;
    const T uxj = q.ux, uyj = q.uy, uzj = q.uz;
    if (P.fv_solve) {
      const T visj = tab ? v.visp[q.j] : visf;
      const T vm = tab ? (visi * visj) / (visi + visj) : vis_const;
      // (x.g / r) / r: the self pair (x = 0, 1/r = 1e20) stays 0 in FP32,
      // where a product 1/r^2 would overflow
      T Cv = (dx * gx + dy * gy + dz * gz) * ir;
      Cv *= ir;
      Cv *= vm * cv_i * Vj;
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
      Pij /= (rhoi + (T)q.rho);
      Pij *= (T)0.5;
      const bool av = use && uij_xij < 0;
      tx += av ? -Pij * gx : (T)0;
      ty += av ? -Pij * gy : (T)0;
      tz += av ? -Pij * gz : (T)0;
    }
    if (euler) {
      const T c = Vj * (1 - eli);
      ex += use ? (uxi * (uxj - uxi) * gx + uyi * (uxj - uxi) * gy + uzi * (uxj - uxi) * gz) * c : (T)0;
      ey += use ? (uxi * (uyj - uyi) * gx + uyi * (uyj - uyi) * gy + uzi * (uyj - uyi) * gz) * c : (T)0;
      ez += use ? (uxi * (uzj - uzi) * gx + uyi * (uzj - uzi) * gy + uzi * (uzj - uzi) * gz) * c : (T)0;
      if (P.con_solve) {
        T tj = q.temp;
        et += use ? (uxi * (tj - tempi) * gx + uyi * (tj - tempi) * gy + uzi * (tj - tempi) * gz) * c : (T)0;
      }
    }
    if (P.con_solve) {
      // 4 eps_i m_j k_j k_i (T_i - T_j) dW / ((r + 1e-10) rho_i rho_j (k_i + k_j))
      T kf = kc_const;
      if (tab) {
        const T kcj = v.condp[q.j];
        kf = (T)4 * pori * kcj * kci / (rhoi * (kci + kcj));
      }
      const T ir10 = sizeof(T) == 4 ? ir : (T)(1.0 / (rd + 1e-10));
      const T H = kf * Vj * (tempi - (T)q.temp) * dw * ir10;
      Rc += use ? H : (T)0;
    }
  };
  // the difference-form pressure gradient carries no hydrostatic offset, so
  // with FP32 fields the position records give the geometry; the symmetric
  // form keeps FP64 geometry, read from the packed records
  if (packed_geometry<R>(P)) for_nbrs4_pk<true>(v, i, true, term);
  else for_nbrs4<true>(v, i, true, term);
  if (P.fg_solve) tz += -(T)P.gravity;
  v.fx[i] = tx - ex;
  v.fy[i] = ty - ey;
  v.fz[i] = tz - ez;
  if (P.con_solve) v.dtemp[i] = Rc / cpi * (T)(pti != -1) - et + (T)(fl / (rhoid * (double)cpi));
  v.pgx[i] = (R)px;
  v.pgy[i] = (R)py;
  v.pgz[i] = (R)pz;
  v.fmag[i] = sqrt(tx * tx + ty * ty + tz * tz);
}

// KERNEL_Neumann_boundary3D_sph (functions_BC.cuh:139-274): wall particles
// take a Shepard interpolation of fluid p, T and eps-weighted rho over the
// 27 cells within 2*kappa*h, with W(2h) normalised by A(h).
// the Neumann walls' candidate records (mixed path), before each pass
__global__ void k_pack_neu(SphView<float> v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  RecNeu r;
  r.p = v.pos4[i];
  r.pres = v.pres[i];
  r.rhopor = v.rho[i] / v.por[i];
  r.vol = (float)v.m[i] / (float)v.rho[i];
  r.temp = v.temp[i];
  r.pt = v.ptype[i];
  v.rn[i] = r;
}

constexpr int kNeumannLanes = 8;

template <typename R>
__global__ void k_neumann(DevParams P, SphView<R> v, CellView cv, double tol_abs) {
  typedef R T;
  // kNeumannLanes lanes per particle: wall particles are a minority that
  // clusters in cell order, one thread each leaves the chip nearly idle
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kNeumannLanes);
  if (i >= v.n) return;
  if (v.ptype[i] > 0) return;
  const double hd = v.h[i];
  const T h = (T)hd, A = knl_tmpA<T>(P.kernel_type, P.dim, h), H2 = (T)(2.0 * hd);
  const double R2 = 2.0 * (P.kappa * hd);
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  // weights in R, sums in FP64: the wall density is a weighted mean of
  // values within ~1e-4 of each other, and the delta-SPH term differences
  // density ratios at that level
  double tp = 0, fl = 0, tr = 0;
  T tt = 0;
  // FP32 fields: candidates are tested on the position records first, as in
  // the neighbour build; only the band |r - R2| <= tol takes the exact FP64
  // distance, so membership is that of the FP64 test, and the kernel weight
  // of the others takes the record distance (FP32 weights either way)
  constexpr bool REL = std::is_same<R, float>::value;
  const double tol = tol_abs + R2 * (1.0 / 1048576.0);
  const float lo2 = (float)((R2 - tol) * (R2 - tol)), hi2 = (float)((R2 + tol) * (R2 + tol));
  const float4 pi4 = REL ? v.pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  group_for_cells<kNeumannLanes>(P, xi, yi, zi, cv, [&](int j) {
    double rd;
    if constexpr (REL) {
      // one 64 B record (k_pack_neu): the range test's pos4 brings the
      // fields of a candidate in range with it
      const RecNeu q = v.rn[j];
      const float3 d = rel_diff(pi4, q.p, v.pgrid);
      const float d2 = d.x * d.x + d.y * d.y + d.z * d.z;
      if (d2 > hi2) return;  // r >= R2 for certain
      rd = d2 < lo2 ? (double)sqrtf(d2) + 1e-20 : dist_exact(xi - v.x[j], yi - v.y[j], zi - v.z[j]) + 1e-20;
      const int ptj = q.pt;
      if (!(ptj < 1000)) return;
      if (!(rd < R2)) return;
      if (!(ptj >= 1)) return;
      T w = knl_w<T>(P.kernel_type, A, H2, (T)rd);
      const T wv = (T)q.vol * w;
      tp += (double)wv * q.pres;
      fl += (double)wv;
      tt += wv * (T)q.temp;
      tr += q.rhopor * (double)wv;
    } else {
      rd = dist_exact(xi - v.x[j], yi - v.y[j], zi - v.z[j]) + 1e-20;
      int ptj = v.ptype[j];
      if (!(ptj < 1000)) return;
      if (!(rd < R2)) return;
      if (!(ptj >= 1)) return;
      T w = knl_w<T>(P.kernel_type, A, H2, (T)rd);
      T vj = (T)v.m[j] / (T)v.rho[j];
      const T wv = vj * w;
      tp += (double)wv * v.pres[j];
      fl += (double)wv;
      tt += wv * (T)v.temp[j];
      tr += v.rho[j] / v.por[j] * (double)wv;
    }
  });
  tp = group_sum<kNeumannLanes>(tp);
  fl = group_sum<kNeumannLanes>(fl);
  tt = group_sum<kNeumannLanes>(tt);
  tr = group_sum<kNeumannLanes>(tr);
  if ((threadIdx.x & (kNeumannLanes - 1)) != 0) return;
  if (fl < 1e-6) {
    v.rho[i] = v.rho_ref[i];
    v.pres[i] = 0.0;
  } else {
    v.pres[i] = tp / fl;
    v.temp[i] = tt / (T)fl;
    v.rho[i] = v.por[i] * tr / fl;
  }
}

// KERNEL_open_boundary_extrapolation3D_1_sph + _calc_sph (functions_BC.cuh:
// 787-984, 1204-1217), open_boundary = 1: each inlet/outlet buffer particle
// takes Shepard means of the interior fluid (p_type 1, buffer_type 0) around
// its ghost node, mirrored in z about L2 (inlet) or L3 (outlet), within
// kappa h of it over the 27 cells of the ghost node's cell.  Inlet: p = mean
// p, rho = eps rho_air, u = (0, 0, 2 u_in) - mean u, T = T_in.  Outlet:
// p = 0, rho = eps mean(rho/eps), u = mean u, T = mean T.  m = rho_old vol0
// (rho_old vol for Eulerian fluid).  The reference's gradient sums enter
// with a factor 0 and are left out.  A buffer particle reads only interior
// particles and its own fields, so the extrapolation and its copy-back
// (the _calc kernel) are one pass.
template <typename R>
__global__ void k_open_boundary(DevParams P, SphView<R> v, CellView cv, OpenBoundary ob) {
  // kNeumannLanes lanes per particle, as for the walls: buffer particles are
  // a thin band in cell order
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kNeumannLanes);
  if (i >= v.n) return;
  const int bt = v.btype[i];
  if (bt <= 0 || v.ptype[i] >= 1000) return;
  const double hd = v.h[i];
  const double A = knl_tmpA<double>(P.kernel_type, P.dim, hd);
  const double Rs = P.kappa * hd;
  const double xg = v.x[i], yg = v.y[i], zi = v.z[i];
  const double zg = bt == 1 ? 2 * ob.L[1] - zi : (bt == 2 ? 2 * ob.L[2] - zi : zi);
  double trho = 0, tpres = 0, tux = 0, tuy = 0, tuz = 0, ttemp = 0, tflt = 0;
  group_for_cells<kNeumannLanes>(P, xg, yg, zg, cv, [&](int j) {
    if (!(v.ptype[j] < 1000)) return;
    const double r = dist_exact(xg - v.x[j], yg - v.y[j], zg - v.z[j]) + 1e-20;
    if (!(r < Rs)) return;
    if (!(v.ptype[j] == 1 && v.btype[j] == 0)) return;
    const double w = knl_w<double>(P.kernel_type, A, hd, r);
    const double volj = v.vol[j];
    trho += v.rho[j] / v.por[j] * w * volj;
    tpres += volj * v.pres[j] * w;
    tuz += volj * (double)v.uz[j] * w;
    tux += volj * (double)v.ux[j] * w;
    tuy += volj * (double)v.uy[j] * w;
    ttemp += volj * (double)v.temp[j] * w;
    tflt += volj * w;
  });
  trho = group_sum<kNeumannLanes>(trho);
  tpres = group_sum<kNeumannLanes>(tpres);
  tuz = group_sum<kNeumannLanes>(tuz);
  tux = group_sum<kNeumannLanes>(tux);
  tuy = group_sum<kNeumannLanes>(tuy);
  ttemp = group_sum<kNeumannLanes>(ttemp);
  tflt = group_sum<kNeumannLanes>(tflt);
  if ((threadIdx.x & (kNeumannLanes - 1)) != 0) return;
  const double pori = v.por[i], rho_old = v.rho[i];
  double m = rho_old * v.vol0[i];
  if (P.eulerian && v.ptype[i] > 0) m = rho_old * v.vol[i];
  if (bt == 1) {
    v.pres[i] = tpres / tflt;
    v.rho[i] = pori * P.rho_fluid;
    v.uz[i] = (R)(ob.inlet_u * 2 - tuz / tflt);
    v.ux[i] = (R)(0 - tux / tflt);
    v.uy[i] = (R)(0 - tuy / tflt);
    v.temp[i] = (R)ob.inlet_temp;
  } else if (bt == 2) {
    v.pres[i] = 0.0;
    v.rho[i] = pori * (trho / tflt);
    v.uz[i] = (R)(tuz / tflt);
    v.ux[i] = (R)(tux / tflt);
    v.uy[i] = (R)(tuy / tflt);
    v.temp[i] = (R)(ttemp / tflt);
  } else {
    return;
  }
  v.m[i] = m;
}

// KERNEL_pressureforce3D_sph (functions_PPE2.cuh:2154-2335)
// KERNEL_clc_surface_detect3D (functions_PST.cuh:263-320), which the
// reference calls just before its PPE (commented out at
// ISPH_Calc_seperate.cuh:289-296): the mean of the unit vectors (x_i -
// x_j)/(r + 1e-10) over the neighbours within kappa h, divided by their count
// non (self included), marks surface particles where its length exceeds
// fs_threshold or non < 10.  free_surface = 2 counts neighbours of the
// particle's own p_type, as the reference does; free_surface = 1 counts every
// SPH particle, walls included, so that fluid against a wall is not surface.
// Surface particles leave the PPE rows and take the Dirichlet pressure p_ref
// (the commented-out rows of functions_PPE2.cuh:533-534) when the rows are
// marked.  FP64 sums in the list's (the reference's) order with no
// contraction, so labels match the oracle's exactly.  Labels depend on
// positions and types only and are computed after every neighbour-list build.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_surface(DevParams P, SphView<R> v) {
#pragma clang fp contract(off)
  int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  const bool fluid = v.itype[i] <= 2 && !(P.open_boundary > 0 && v.btype[i] > 0) && pti < 1000 && pti > 0;
  if (!fluid) {
    v.surf[i] = 0;
    return;
  }
  double tx = 0, ty = 0, tz = 0;
  int non = 0;
  for_nbrs4<false>(v, i, false, [&](const Nb<R>& q, bool on) {
#pragma clang fp contract(off)
    const bool use = on && (P.free_surface == 2 ? q.pt == pti : q.pt < 1000);
    const double rd = 1.0 / (dist_exact(q.dx, q.dy, q.dz) + 1e-10);
    const double ax = q.dx * rd, ay = q.dy * rd, az = q.dz * rd;
    tx += use ? ax : 0.0;
    ty += use ? ay : 0.0;
    tz += use ? az : 0.0;
    non += use ? 1 : 0;
  });
  const double rn = 1.0 / ((double)non + 1e-10);
  tx = tx * rn;
  ty = ty * rn;
  tz = tz * rn;
  const bool surf = dist_exact(tx, ty, tz) > P.fs_threshold || non < 10;
  v.surf[i] = surf ? 1 : 0;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_pressure_force(DevParams P, SphView<R> v) {
  typedef double T;  // FP64 pair sums: see the note on pressure-gradient cancellation
  int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] > 2) return;
  if (P.open_boundary > 0 && v.btype[i] > 0) return;
  if (pti >= 1000 || pti <= 0) return;
  const int n = v.n;
  const double hd = v.h[i];
  const T h = hd, ih = 1.0 / hd, A = knl_tmpA<T>(P.kernel_type, P.dim, h);
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i], pid = v.pres[i];
  const T rhoi = v.rho[i], pori = v.por[i];
  // Cp = -eps_i m_j pp / (rho_i rho_j): the per-particle factor hoisted and
  // 1/r, 1/rho_j taken once per pair (the rounding moves by an ulp)
  const T ci = -pori / rhoi;
  T tx = 0, ty = 0, tz = 0;
  auto term = [&](const auto& q, bool on) {
    const bool use = on && q.pt < 1000;
    T dx = (T)q.dx, dy = (T)q.dy, dz = (T)q.dz;
    T r = sqrt(dx * dx + dy * dy + dz * dz) + (T)1e-20;
    T dw = knl_dw_ih<T>(P.kernel_type, A, h, ih, r);
    const T sd = dw / r;
    T gx = sd * dx, gy = sd * dy, gz = sd * dz;
    kgc_apply(P, v.cm, n, i, gx, gy, gz);
    const T pp = P.pgrad_form ? q.pres - pid : pid + q.pres;
    const T Cp = ci * (q.m * pp / q.rho);
    tx += use ? Cp * gx : (T)0;
    ty += use ? Cp * gy : (T)0;
    tz += use ? Cp * gz : (T)0;
  };
  if (P.fp_solve) {
    // the difference form carries no hydrostatic offset, so geometry from
    // the position records is enough; the symmetric form keeps FP64
    // geometry, read from the packed records
    if (packed_geometry<R>(P)) for_nbrs4_pk<false>(v, i, true, term);
    else for_nbrs4<true>(v, i, true, term);
  }
  v.fpx[i] = (R)tx;
  v.fpy[i] = (R)ty;
  v.fpz[i] = (R)tz;
}

// KERNEL_clc_projection_sph (functions_TIME_ISPH2.cuh:104-212): predictor
template <typename R>
__global__ void k_predictor(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2) return;
  const int t = v.ptype[i];
  const R ux0 = v.ux[i], uy0 = v.uy[i], uz0 = v.uz[i];
  if (t >= 0) {
    const R bz = (R)(v.btype[i] == 0), dt = (R)P.dt_sph;
    v.ux[i] = ux0 + v.fx[i] * bz * dt;
    v.uy[i] = uy0 + v.fy[i] * bz * dt;
    v.uz[i] = uz0 + v.fz[i] * bz * dt;
  }
  v.ux0[i] = ux0;
  v.uy0[i] = uy0;
  v.uz0[i] = uz0;
  if (P.first_step) v.rho_ref[i] = v.rho[i];
  if (P.con_solve == 1 && t > 0) v.temp0[i] = v.temp[i];
}

// KERNEL_time_update_projection_sph (functions_TIME_ISPH2.cuh:467-621):
// corrector with velocity clamp; positions integrated in FP64
template <typename R>
__global__ void k_corrector(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  if (v.itype[i] > 2) return;
  const int t = v.ptype[i];
  const double dt = P.dt_sph;
  R ux0, uy0, uz0, ax, ay, az;
  if (t > 0) {
    ux0 = v.ux[i]; uy0 = v.uy[i]; uz0 = v.uz[i];
    R bz = (R)(v.btype[i] == 0);
    ax = v.fpx[i] * bz; ay = v.fpy[i] * bz; az = v.fpz[i] * bz;
  } else {
    ux0 = v.ux0[i]; uy0 = v.uy0[i]; uz0 = v.uz0[i];
    ax = ay = az = 0;
  }
  R uxc = ux0 + ax * (R)dt, uyc = uy0 + ay * (R)dt, uzc = uz0 + az * (R)dt;
  if ((double)(uxc * uxc + uyc * uyc + uzc * uzc) >= P.u_limit * P.u_limit) {
    uxc = ux0; uyc = uy0; uzc = uz0;
  }
  const double el = (t > 0 && !P.eulerian) ? 1.0 : 0.0;
  if (!P.xsph_solve || t > 1000) {
    double xn = v.x[i] + ((double)v.ux0[i] + (double)uxc) / 2.0 * dt * el;
    double yn = v.y[i] + ((double)v.uy0[i] + (double)uyc) / 2.0 * dt * el;
    double zn = v.z[i] + ((double)v.uz0[i] + (double)uzc) / 2.0 * dt * el;
    if (P.pmask) {
      if (P.pmask & 1) xn = per_pos(xn, P.x_min, P.plen[0]);
      if (P.pmask & 2) yn = per_pos(yn, P.y_min, P.plen[1]);
      if (P.pmask & 4) zn = per_pos(zn, P.z_min, P.plen[2]);
    }
    v.x[i] = xn;
    v.y[i] = yn;
    v.z[i] = zn;
  }
  v.ux[i] = uxc;
  v.uy[i] = uyc;
  v.uz[i] = uzc;
  if (P.con_solve == 1 && t > 0) v.temp[i] = v.temp0[i] + v.dtemp[i] * (R)dt * (R)(t != -1);
}

// KERNEL_xsph3D_sph + KERNEL_xsph3D_calc_sph (functions_TIME.cuh:1735-1850,
// 1985-1997): Shepard-normalised XSPH velocity and temperature smoothing
template <typename R>
__global__ __launch_bounds__(kBlock) void k_xsph(DevParams P, SphView<R> v) {
  typedef R T;
  int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] != 1 || pti > 1000 || pti == 0 || pti == 9) return;
  const double hd = v.h[i];
  const T h = (T)hd, ih = (T)(1.0 / hd), A = knl_tmpA<T>(P.kernel_type, P.dim, h);
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  const T uxi = v.ux[i], uyi = v.uy[i], uzi = v.uz[i], ti = v.temp[i], c = (T)P.c_xsph;
  T tx = 0, ty = 0, tz = 0, tt = 0, fl = 0;
  auto term = [&](const auto& q, bool on) {
    int ptj = q.pt;
    const bool use = on && (ptj == 1 || ptj == 3);
    T dx = (T)q.dx, dy = (T)q.dy, dz = (T)q.dz;
    T r = sqrt(dx * dx + dy * dy + dz * dz);
    T w = knl_w_ih<T>(P.kernel_type, A, h, ih, r);
    T vj = (T)q.m / (T)q.rho;
    tx += use ? c * vj * (-uxi + (T)q.ux) * w : (T)0;
    ty += use ? c * vj * (-uyi + (T)q.uy) * w : (T)0;
    tz += use ? c * vj * (-uzi + (T)q.uz) * w : (T)0;
    tt += use ? (T)0.01 * c * vj * ((T)q.temp - ti) * w : (T)0;
    fl += use ? vj * w : (T)0;
  };
  if constexpr (std::is_same<R, float>::value) for_nbrs4_r32<true>(v, i, false, term);
  else for_nbrs4<true>(v, i, false, term);
  const T on = (T)(pti > 0);
  v.xsx[i] = (uxi + tx / fl) * on;
  v.xsy[i] = (uyi + ty / fl) * on;
  v.xsz[i] = (uzi + tz / fl) * on;
  v.xst[i] = (ti + tt / fl) * on;
}

template <typename R>
__global__ void k_xsph_calc(DevParams P, SphView<R> v) {
  int i = particle_index();
  if (i >= v.n) return;
  const int pti = v.ptype[i];
  if (v.itype[i] != 1 || pti > 1000 || pti == 0 || pti == 9) return;
  v.ux[i] = v.xsx[i];
  v.uy[i] = v.xsy[i];
  v.uz[i] = v.xsz[i];
  v.temp[i] = v.xst[i];
  if (P.fix_xsph && pti > 0 && !P.eulerian) {
    // positions from the start-of-step position and the smoothed velocity
    const double dt = P.dt_sph;
    double xn = v.x[i] + (double)v.xsx[i] * dt;
    double yn = v.y[i] + (double)v.xsy[i] * dt;
    double zn = v.z[i] + (double)v.xsz[i] * dt;
    if (P.pmask) {
      if (P.pmask & 1) xn = per_pos(xn, P.x_min, P.plen[0]);
      if (P.pmask & 2) yn = per_pos(yn, P.y_min, P.plen[1]);
      if (P.pmask & 4) zn = per_pos(zn, P.z_min, P.plen[2]);
    }
    v.x[i] = xn;
    v.y[i] = yn;
    v.z[i] = zn;
  }
}

}  // namespace

// kernel_copy_max_sph/_dem + cub::DeviceReduce::Max (functions_OUTPUT.cuh:35-65,
// ISPH_Calc_seperate.cuh:928-969): max |u|, rho and ftotal over one species;
// walls (p_type 0) and i_type >= i_type_crt (2) count as 0.  Max is order
// independent, so the two-stage reduction is exact.
constexpr int kMonBlocks = 1024;
template <typename U>
__global__ __launch_bounds__(256) void k_monitor(int n, const int* ptype, const int* itype, const U* ux, const U* uy,
                                                 const U* uz, const double* rho, const U* ft, double* partials) {
  __shared__ double red[3][4];
  double m[3] = {0.0, 0.0, 0.0};
  for (int i = particle_index(); i < n; i += gridDim.x * blockDim.x) {
    if (ptype[i] == 0 || itype[i] >= 2) continue;
    const double a = (double)ux[i], b = (double)uy[i], c = (double)uz[i];
    m[0] = fmax(m[0], sqrt(a * a + b * b + c * c));
    m[1] = fmax(m[1], rho[i]);
    if (ft) m[2] = fmax(m[2], (double)ft[i]);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = 0; q < 3; q++) {
    double v = m[q];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if (lane == 0) red[q][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    partials[q * gridDim.x + blockIdx.x] = fmax(fmax(red[q][0], red[q][1]), fmax(red[q][2], red[q][3]));
  }
}

__global__ void k_monitor_fin(const double* partials, int G, double* out) {
  const int q = threadIdx.x;
  if (q >= 3) return;
  double v = 0.0;
  for (int b = 0; b < G; b++) v = fmax(v, partials[q * G + b]);
  out[q] = v;
}

template <typename U>
int monitor_max(hipStream_t s, int n, const int* ptype, const int* itype, const U* ux, const U* uy, const U* uz,
                const double* rho, const U* ft, double* scratch, double* out) {
  out[0] = out[1] = out[2] = 0.0;
  if (n <= 0) return 0;
  int G = grid_xcd(n);
  if (G > kMonBlocks) G = kMonBlocks;
  k_monitor<U><<<G, 256, 0, s>>>(n, ptype, itype, ux, uy, uz, rho, ft, scratch);
  k_monitor_fin<<<1, 64, 0, s>>>(scratch, G, scratch + 3 * kMonBlocks);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, scratch + 3 * kMonBlocks, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}
template int monitor_max<float>(hipStream_t, int, const int*, const int*, const float*, const float*, const float*,
                                const double*, const float*, double*, double*);
template int monitor_max<double>(hipStream_t, int, const int*, const int*, const double*, const double*,
                                 const double*, const double*, const double*, double*, double*);

#define LAUNCH(kern, ...)                                           \
  do {                                                              \
    if (v.n > 0) {                                                  \
      kern<R><<<grid_xcd(v.n), kBlock, 0, s>>>(__VA_ARGS__);        \
      HIP_TRY(hipGetLastError());                                   \
    }                                                               \
    return 0;                                                       \
  } while (0)

template <typename R> int sph_mass_init(hipStream_t s, const DevParams& P, const SphView<R>& v) { LAUNCH(k_mass_init, P, v); }
template <typename R> int sph_kgc(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  TRY(pack_r32(s, v, false, false));
  LAUNCH(k_kgc, P, v);
}
template <typename R>
int sph_porosity(hipStream_t s, const DevParams& P, const SphView<R>& v, const DemView& d, const CellView& dc) {
  TRY(pack_r32(s, v, false, false));
  LAUNCH(k_porosity_sph, P, v, d, dc);
}
template <typename R>
int sph_kgc_porosity(hipStream_t s, const DevParams& P, const SphView<R>& v, const DemView& d, const CellView& dc,
                     bool kgc) {
  if constexpr (std::is_same<R, float>::value) {
    if (v.n <= 0) return 0;
    TRY(pack_r32(s, v, false, false));
    if (kgc) k_kgc_porosity<true><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v, d, dc);
    else k_kgc_porosity<false><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v, d, dc);
    HIP_TRY(hipGetLastError());
    return 0;
  } else {
    if (kgc) TRY(sph_kgc(s, P, v));
    return sph_porosity(s, P, v, d, dc);
  }
}
template <typename R> int sph_prep(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  if (v.n > 0) {
    k_prep_prep<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
    HIP_TRY(hipGetLastError());
    TRY(pack_r32(s, v, P.delsph == 2, false));
    k_prep<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}
template <typename R> int sph_prep_prep(hipStream_t s, const DevParams& P, const SphView<R>& v) { LAUNCH(k_prep_prep, P, v); }
template <typename R> int sph_prep_sums(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  TRY(pack_r32(s, v, P.delsph == 2, false));
  LAUNCH(k_prep, P, v);
}
template <typename R> int sph_advection(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  if (v.n > 0) {
    if (packed_geometry<R>(P)) k_pack_records<R, true><<<grid_for(v.n), kBlock, 0, s>>>(v, P.heat_flux != 0.0);
    if (P.heat_flux != 0.0) k_advection<R, true><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
    else k_advection<R, false><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}
template <typename R> int sph_flux_walls(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  if (v.n > 0 && P.heat_flux != 0.0) {
    k_flux_walls<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}
template <typename R> int sph_neumann(hipStream_t s, const DevParams& P, const SphView<R>& v, const CellView& cv) {
  const double m = fmax(fmax(fmax(fabs(P.x_min), fabs(P.x_max)), fmax(fabs(P.y_min), fabs(P.y_max))),
                        fmax(fabs(P.z_min), fabs(P.z_max)));
  if (v.n > 0) {
    if constexpr (std::is_same<R, float>::value) k_pack_neu<<<grid_for(v.n), kBlock, 0, s>>>(v);
    const size_t nt = (size_t)v.n * kNeumannLanes;
    k_neumann<R><<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, s>>>(P, v, cv, m * (1.0 / 2097152.0));
    HIP_TRY(hipGetLastError());
  }
  return 0;
}
template <typename R> int sph_surface(hipStream_t s, const DevParams& P, const SphView<R>& v) { LAUNCH(k_surface, P, v); }
template <typename R> int sph_props(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  if (v.n > 0 && P.prop_table) {
    k_props<R><<<grid_for(v.n), kBlock, 0, s>>>(P, v);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}
template <typename R> int sph_pressure_force(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  if (v.n <= 0) return 0;
  if (P.fp_solve && packed_geometry<R>(P)) k_pack_records<R, false><<<grid_for(v.n), kBlock, 0, s>>>(v, false);
  k_pressure_force<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
  HIP_TRY(hipGetLastError());
  return 0;
}
template <typename R>
int sph_open_boundary(hipStream_t s, const DevParams& P, const SphView<R>& v, const CellView& cv,
                      const OpenBoundary& ob) {
  if (v.n > 0) {
    const size_t nt = (size_t)v.n * kNeumannLanes;
    k_open_boundary<R><<<(unsigned)((nt + kBlock - 1) / kBlock), kBlock, 0, s>>>(P, v, cv, ob);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}
template <typename R> int sph_predictor(hipStream_t s, const DevParams& P, const SphView<R>& v) { LAUNCH(k_predictor, P, v); }
template <typename R> int sph_corrector(hipStream_t s, const DevParams& P, const SphView<R>& v) { LAUNCH(k_corrector, P, v); }
template <typename R> int sph_xsph_only(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  if (v.n > 0) {
    TRY(pack_r32(s, v, false, true));
    k_xsph<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}
template <typename R> int sph_xsph(hipStream_t s, const DevParams& P, const SphView<R>& v) {
  if (v.n > 0) {
    TRY(pack_r32(s, v, false, true));
    k_xsph<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
    HIP_TRY(hipGetLastError());
    k_xsph_calc<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}

#define INST(R)                                                                                              \
  template int sph_mass_init<R>(hipStream_t, const DevParams&, const SphView<R>&);                          \
  template int sph_kgc<R>(hipStream_t, const DevParams&, const SphView<R>&);                                \
  template int sph_porosity<R>(hipStream_t, const DevParams&, const SphView<R>&, const DemView&, const CellView&); \
  template int sph_prep<R>(hipStream_t, const DevParams&, const SphView<R>&);                               \
  template int sph_prep_sums<R>(hipStream_t, const DevParams&, const SphView<R>&);                          \
  template int sph_prep_prep<R>(hipStream_t, const DevParams&, const SphView<R>&);                          \
  template int sph_kgc_porosity<R>(hipStream_t, const DevParams&, const SphView<R>&, const DemView&,         \
                                   const CellView&, bool);                                                  \
  template int sph_advection<R>(hipStream_t, const DevParams&, const SphView<R>&);                          \
  template int sph_flux_walls<R>(hipStream_t, const DevParams&, const SphView<R>&);                         \
  template int sph_neumann<R>(hipStream_t, const DevParams&, const SphView<R>&, const CellView&);           \
  template int sph_surface<R>(hipStream_t, const DevParams&, const SphView<R>&);                            \
  template int sph_props<R>(hipStream_t, const DevParams&, const SphView<R>&);                              \
  template int sph_pressure_force<R>(hipStream_t, const DevParams&, const SphView<R>&);                     \
  template int sph_open_boundary<R>(hipStream_t, const DevParams&, const SphView<R>&, const CellView&,      \
                                    const OpenBoundary&);                                                   \
  template int sph_predictor<R>(hipStream_t, const DevParams&, const SphView<R>&);                          \
  template int sph_corrector<R>(hipStream_t, const DevParams&, const SphView<R>&);                          \
  template int sph_xsph<R>(hipStream_t, const DevParams&, const SphView<R>&);                                \
  template int sph_xsph_only<R>(hipStream_t, const DevParams&, const SphView<R>&);
INST(float)
INST(double)

}  // namespace sophia
