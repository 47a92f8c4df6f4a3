/*
 * penguin_oracle.h -- C ABI of the FP64 host oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement, in double
 * precision, of the reference solver's step (PEnGUIn src/solver/solve.cu:212-292
 * and everything it calls).  Only tests/, the smoke check and the CPU-baseline
 * leg of bench.py may load it, and only as the checker: the product library
 * never links, loads or calls it.
 *
 * Layout of state buffers crossing this ABI: interior cells only, x fastest
 * (index i + xres*j), one contiguous block per field, gas fields in the order
 * rho, p, u (radial), v (azimuthal), w and dust fields rho, u, v, w -- the
 * order of the reference's restructure_data/write_check_point
 * (src/output/output.cu:95-180).
 */
#ifndef PENGUIN_ORACLE_H
#define PENGUIN_ORACLE_H

#include "../include/pg_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;

/* grid + initial state (init.cu:472-532) + planet (planet.cu:319) + orbital
   advection rotation (orbit.cu:111) */
orc_grid* orc_create(const pg_params* P);
void orc_destroy(orc_grid* g);
int  orc_num_threads(void);

void orc_get_state(const orc_grid* g, double* gas5, double* dust4);
void orc_set_state(orc_grid* g, const double* gas5, const double* dust4);
/* padded coordinate arrays: xa[xarr+1], ya[yarr+1], xv[xarr], yv[yarr] */
void orc_get_coords(const orc_grid* g, double* xa, double* xv, double* ya, double* yv);
void orc_get_zcoords(const orc_grid* g, double* za, double* zv);
void orc_get_planet(const orc_grid* g, double* body10);
void orc_get_rot(const orc_grid* g, double* rot_xarr);

/* time stepping (timestep.cu:113-140, solve.cu:279-292, main.cu:320-359) */
double orc_global_dt(orc_grid* g, double dt_prev);
void   orc_solve(orc_grid* g, double t, double dt);
/* advance nsteps with the reference loop semantics; dts[n] receives each dt */
void   orc_run(orc_grid* g, int nsteps, double* t_io, double* dt_io, double* dts);

/* module-level entry points (each applies the boundary fills the reference
   applies before it) */
void orc_viscosity(orc_grid* g, int on_T, double* vis_interior2);
void orc_sources(orc_grid* g, double mdt, double dt, int inplace);
void orc_sweepx(orc_grid* g, double dt);
void orc_sweepz(orc_grid* g, double dt);
void orc_sweepy(orc_grid* g, double t, double dt);
void orc_sweepx_dust(orc_grid* g, double dt);
void orc_sweepy_dust(orc_grid* g, double t, double dt);
void orc_orbital_advection(orc_grid* g, double dt);
void orc_get_orbadv(const orc_grid* g, int* inc_xres, double* res_xres);
void orc_killwave(orc_grid* g, double dt);
int  orc_write_snapshot(const orc_grid* g, const char* filename, double t);
/* ghost fills of one direction (dir 0: boundx/boundx_dust, 1: boundy/
   boundy_dust, 2: boundz; boundary.cu:13-374, 462-718) applied to the
   current state, which is returned with the 2*pad ghost cells of that
   direction: field-major as orc_get_state, the padded direction's extent
   n + 2 pad (x fastest).  dust may be NULL; 3D has none. */
void orc_bound(orc_grid* g, int dir, double* gas, double* dust);
void orc_evolve_planet(orc_grid* g, double t);

/* unit-level functions, on caller-supplied row arrays */
void orc_fourth_parameters(int geom, const double* x6, const double* a5, double* par3);
void orc_ppm_adjust(double flat, double* par3);
double orc_ppm_aveL(double q, const double* par3);
double orc_ppm_aveR(double q, const double* par3);
/* PEM sub-cell means (PEM.cu:51-93) */
double orc_pem_aveL(double q, const double* par3);
double orc_pem_aveR(double q, const double* par3);
/* flattening of cell 2 of a 5-cell window (PPM.cu:7-62) */
double orc_flatten(double gam, const double* r5, const double* p5, const double* u5);
/* traced interface state at interface i of a row (states.cu:1-135) with the
   flattening coefficients supplied; S10 = rl,pl,ul,vl,wl,rr,pr,ur,vr,wr */
void orc_set_state_row(int geom, double gam, int n, const double* xa, const double* dv,
                       const double* r, const double* p, const double* u,
                       const double* v, const double* w, const double* flat,
                       int i, double dt, double* S10);
/* wave speeds + HLLC (wave_speeds.cu:319-333, HLLC.cu:10-77):
   out = pm, sl, sm, sr, F.r, F.p, F.u, F.v, F.w, pres */
void orc_riemann_hllc(double gam, int eos, const double* S10, double* out10);
/* whole-row flux difference of the gas sweep (riemann.cu:91-131), Del5[n] */
void orc_riemann_row(int geom, double gam, int n, const double* xa, const double* dv,
                     const double* r, const double* p, const double* u,
                     const double* v, const double* w, double dt, double* Del5);
/* dust interface flux (dust.cu:6-105) with flat = 0; out = F.r,F.u,F.v,F.w */
void orc_dust_flux(int geom, int n, const double* xa, const double* dv,
                   const double* r, const double* u, const double* v,
                   const double* w, int i, double dt, double* F4);
/* analytic Kepler orbit (planet.cu:105-133): body10 = m,a,e,x,y,z,vx,vy,vz,rs */
void orc_orbit_solution(const pg_params* P, double t, double* body10);
/* orc_riemann_hllc plus the interface velocity uprs of the internal-energy
   equation (HLLC.cu:58-68): out11 = the ten values above, uprs */
void orc_riemann_hllc_e(double gam, int eos, const double* S10, double* out11);
/* planet acceleration at (rad, azi, pol) for body10 with dt = 0 (the star's
   own gravity excluded): sph 1 star_planet_grav (forces.cu:14-39, spline
   softening, spherical components r, phi, theta), sph 0
   star_planet_grav_cyl (forces.cu:41-58, pol = z) */
void orc_planet_force(int sph, const double* body10, double rad, double azi, double pol, double* f3);
/* time average of the gas (init_average/averaging, output.cu:250-329; applied
   around each orc_run step as main.cu:344-350 does) */
void orc_average_begin(orc_grid* g, double interval);
void orc_average_end(orc_grid* g);
void orc_average_get(const orc_grid* g, double* gas5);
/* planet torque on the gas (global_sum, output.cu:336-457), 2D cylindrical */
double orc_torque(const orc_grid* g);
/* tangent-linear build only (liboracle_tlm.so): the derivative parts of the
   interior state, in the orc_get_state field order */
void orc_set_tangent(orc_grid* g, const double* gas, const double* dust);
void orc_get_tangent(const orc_grid* g, double* gas, double* dust);

#ifdef __cplusplus
}
#endif

#endif
