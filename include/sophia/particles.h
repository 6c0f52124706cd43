/*
 * Particle interchange records at the drop-in boundary.
 *
 * These three records are the reference's on-host / on-device particle
 * layout (class_Cuda_Particle_Array2.cuh:11-167): same field names, same
 * order, same types (Real = double, int_t = uint_t = int,
 * Variable_Type.cuh:2-5), so a buffer of them can be handed across the
 * boundary byte for byte.  They are used only for import, export and parity
 * dumps; inside the library every field lives in its own structure-of-arrays
 * stream in device memory.
 *
 * Sizes: part1 1272 B, part2 176 B, part3 496 B (checked by static asserts
 * in the library and by a CPU test).
 */
#ifndef SOPHIA_PARTICLES_H
#define SOPHIA_PARTICLES_H

#ifdef __cplusplus
extern "C" {
#endif

typedef double sophia_real;

/* class_Cuda_Particle_Array2.cuh:11-85 */
typedef struct part1 {
  int i_type, buffer_type, p_type, dem_idx;
  int ct_boundary;
  sophia_real x, y, z;
  sophia_real x_star, y_star, z_star;
  sophia_real ux, uy, uz;
  sophia_real wx, wy, wz;
  sophia_real m;
  sophia_real ri;
  sophia_real rad;
  sophia_real h;
  sophia_real temp, temp1, temp2, temp3;
  sophia_real pres;
  sophia_real rho;
  sophia_real flt_s;
  sophia_real flt_sd;
  sophia_real flt_sd_2;
  sophia_real w_dx;
  sophia_real enthalpy;
  sophia_real concn;
  sophia_real grad_rhox, grad_rhoy, grad_rhoz;
  sophia_real pgf_x, pgf_y, pgf_z;
  sophia_real Fdx_b, Fdy_b, Fdz_b;
  sophia_real Fdx_df, Fdy_df, Fdz_df;
  sophia_real Fdx_da, Fdy_da, Fdz_da;
  sophia_real DEMpor;
  sophia_real dDEMpor;
  sophia_real dDEMpor_prev;
  sophia_real DEMvf;
  sophia_real Q_sd;
  sophia_real Q_sdf;
  sophia_real Q_f;
  sophia_real ct_idx[16];
  sophia_real del_s[16][3];
  sophia_real vol_power;
  sophia_real k_turb, e_turb;
  sophia_real pres_ipp;
  sophia_real pres_ipp_p, pres_ipp_s;
  sophia_real test1, test2, test3, test4, test5, test6, test7, test8;
  sophia_real cond;
  sophia_real vol0, vol;
  sophia_real elix, eliy, eliz;
  sophia_real fbx, fby, fbz;
  sophia_real fcx, fcy, fcz;
  int pos;
  sophia_real PPE1, PPE2, PPE3, PPE4;
  sophia_real XSPH_ux, XSPH_uy, XSPH_uz;
  sophia_real XSPH_temp;
  sophia_real OpenBC_pres, OpenBC_rho, OpenBC_m, OpenBC_temp;
  sophia_real OpenBC_ux, OpenBC_uy, OpenBC_uz;
} part1;

/* class_Cuda_Particle_Array2.cuh:87-106 */
typedef struct part2 {
  sophia_real rho_ref;
  sophia_real SR;
  sophia_real x0, y0, z0;
  sophia_real ux0, uy0, uz0;
  sophia_real wx0, wy0, wz0;
  sophia_real rho0;
  sophia_real drho0;
  sophia_real rad0;
  sophia_real vol0, dvol0;
  sophia_real concn0;
  sophia_real enthalpy0;
  sophia_real temp0, temp10, temp20, temp30;
} part2;

/* class_Cuda_Particle_Array2.cuh:108-167 */
typedef struct part3 {
  sophia_real drho;
  sophia_real dconcn;
  sophia_real denthalpy;
  sophia_real drad;
  sophia_real dtemp, dtemp1, dtemp2, dtemp3;
  sophia_real torqx, torqy, torqz;
  sophia_real ftotalx, ftotaly, ftotalz;
  sophia_real ftotal;
  sophia_real fpx, fpy, fpz;
  sophia_real vis_t;
  sophia_real Sxx, Sxy, Sxz, Syy, Syz, Szz;
  sophia_real dk_turb, de_turb;
  sophia_real lbl_surf;
  sophia_real cc;
  sophia_real Cm[3][3];
  sophia_real nx, ny, nz;
  sophia_real nx_u, ny_u, nz_u;
  sophia_real nx_w, ny_w, nz_w;
  sophia_real nx_t, ny_t, nz_t;
  sophia_real nx_tl, ny_tl, nz_tl;
  sophia_real nx_c, ny_c, nz_c;
  sophia_real nmag;
  sophia_real nmag_c;
  sophia_real curv;
  sophia_real fsx, fsy, fsz;
} part3;

#ifdef __cplusplus
}
#endif

#endif /* SOPHIA_PARTICLES_H */
