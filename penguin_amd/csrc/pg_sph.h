// pg_sph.h -- the three-dimensional disk (ndim 3: spherical radius, azimuth,
// polar angle; adiabatic gas), host entry points called by the C ABI in
// pg_host.cpp when a handle is created with ndim = 3.
#pragma once

#include <cstddef>

#include "../../include/pg_params.h"

namespace pg {

struct Sph;

// error reporting shared with pg_host.cpp (pg_last_error)
int report_error(int code, const char* fmt, ...);
// axis edges with ghosts and cell volumes (make_grid, init.cu:472-500) and the
// 4th-order slope weights of a cell from its five neighbours' edges
void host_make_axis(double* a, double* v, double amin, double amax, int res, int pad, int geom, int grid);
void host_slope_weights(int geom, const double* x6, double wl[3], double wr[3]);

int sph_validate(const pg_params& P);
int sph_create(const pg_params& P, Sph** out);
void sph_destroy(Sph* s);
int sph_get_state(Sph* s, double* gas);
void sph_ghost_x(const Sph* s, int g, int k, double* f5);
void sph_dims(const Sph* s, int* nx, int* ny, int* nz);
int sph_set_state(Sph* s, const double* gas);
int sph_global_dt(Sph* s, double dt_prev, double* dt_out);
int sph_solve(Sph* s, double t, double dt);
int sph_run(Sph* s, int nsteps, double* t_io, double* dt_io, double* dts);
int sph_sweep(Sph* s, int dir, double dt);   // 0 radius, 1 azimuth, 2 polar angle
int sph_orbital_advection(Sph* s, double dt, int* inc, double* res);
int sph_sources(Sph* s, double t_planet, double mdt, double dt, int inplace);
int sph_viscosity(Sph* s, int on_T, double* vis);
int sph_planet(Sph* s, double t, double* body10);
int sph_write_snapshot(Sph* s, const char* filename, double t);
int sph_read_snapshot(Sph* s, const char* filename, double* t_out);
int sph_profile_enable(Sph* s, int on);
int sph_profile_get(Sph* s, const char* kernel, double* avg_us, int* count);
int sph_device_bytes(Sph* s, size_t* bytes);
void* sph_stream(Sph* s);

}  // namespace pg
