/*
 * pg_params.h -- run-time configuration of one disk.
 *
 * The reference configures a run at compile time through the macros and
 * `const` globals of include/parameters.h (reference include/parameters.h:4-207,
 * documented in include/README.md).  This struct carries the same quantities,
 * under the same names, as run-time values, so that one library build can
 * advance disks with different resolutions and physics (a batch of disks
 * holds one pg_params per disk).  pg_params_from_parameters_h() in
 * parameters.h fills it from the compile-time header.
 *
 * Plain C, fixed-width fields, no padding surprises: the struct crosses the
 * C ABI and is mirrored field by field by ctypes on the Python side.
 */
#ifndef PG_PARAMS_H
#define PG_PARAMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pg_params {
    /* grid (parameters.h:47, 67-69, 194-207) */
    int32_t ndim;            /* 2 (cylindrical r, phi) or 3 (spherical r, phi, theta) */
    int32_t xres, yres, zres;
    int32_t geomx, geomy, geomz;   /* 1 cylindrical radius, 3 azimuth (2D);
                                      2 spherical radius, 4 azimuth, 5 polar angle (3D) */
    int32_t gridx, gridy, gridz;   /* 0 = linear spacing, 1 = log spacing */
    /* numerics (parameters.h:43, 115-116) */
    int32_t recon_flag;      /* 6 = PPM4 (4th-order slopes + PPM), 5 = PEM4 */
    int32_t EOS_flag;        /* 0 = locally isothermal, 2 = adiabatic (internal_e_flag 1) */
    /* feature switches (parameters.h:4-13): 0 = off, 1 = on               */
    int32_t visc_flag;
    int32_t OrbAdv_flag;
    int32_t dust_flag;
    int32_t kill_flag;
    int32_t n_planet;        /* 0 or 1 */
    /* boundary types (parameters.h:142-149):
       0 analytic initial state, 1 zero gradient, 2 reflecting, 3 periodic */
    int32_t bound_lft, bound_rgh, bound_bak, bound_frn;
    /* storage precision of the device state: 4 (FP32) or 8 (FP64) */
    int32_t precision;
    /* polar-angle boundary types (ndim 3), same codes */
    int32_t bound_bom, bound_top;
    /* domain (parameters.h:194-199) */
    double xmin, xmax, ymin, ymax, zmin, zmax;
    /* hydro and disk (parameters.h:103, 119-127, 155-171) */
    double gam;
    double CFL;
    double frame_omega;
    double sc_h;
    double p_alpha;
    double p_beta;
    double ss_alpha;
    double kill_width;
    /* dust (parameters.h:133-134) */
    double D_G_ratio;
    double Stokes;
    /* planet (parameters.h:183-188) */
    double planet_mass;
    double planet_radius;
    double planet_ecc;
    double rs_fac;
    double ramp_time;
    /* time control (parameters.h:105-109) */
    double sav_interval;
    double end_time;
    /* seeded initial perturbation (not in the reference, which adds none):
       interior Sigma *= 1 + noise_rho*(2U-1); v_r += noise_vr*c_s(R)*(2U-1),
       U from SplitMix64 keyed by (seed, field, i, j) */
    double noise_rho;
    double noise_vr;
    uint64_t seed;
} pg_params;

#ifdef __cplusplus
}
#endif

#endif /* PG_PARAMS_H */
