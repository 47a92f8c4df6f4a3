/*
 * penguin_amd.h -- C ABI of the MI355X Godunov-sweep library.
 *
 * The reference's boundary for this path is a set of plain C++ host
 * functions over Grid* arrays, all compiled against include/parameters.h
 * (reference include/solver.h:4, timestep.h:4, orbit.h:4-7, planet.h:4-5,
 * viscosity.h:4-5, boundary.h:4-14).  Each entry point below replaces one of
 * them; the comment names the reference function and file:line.
 *
 * Conventions
 *   - every function returns 0 on success or a negative PG_E_* code; the
 *     message of the last failure on the calling thread is pg_last_error();
 *     no C++ exception crosses this boundary;
 *   - a pg_handle owns its device memory and one HIP stream; it is not
 *     thread-safe: one host thread per handle;
 *   - state buffers are host FP64 arrays of interior cells, x (radius)
 *     fastest, one contiguous block per field: gas rho, p, u (radial),
 *     v (azimuthal), w; dust rho, u, v, w -- the order of the reference's
 *     snapshot writer (src/output/output.cu:95-180).
 */
#ifndef PENGUIN_AMD_H
#define PENGUIN_AMD_H

#include <stddef.h>
#include <stdint.h>

#include "pg_params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PG_OK 0
#define PG_E_INVALID (-1)     /* bad argument or unsupported configuration */
#define PG_E_HIP (-2)         /* a HIP runtime call failed                 */
#define PG_E_NOMEM (-3)       /* device or host allocation failed          */
#define PG_E_STATE (-4)       /* call not valid in the handle's state      */

typedef struct pg_handle pg_handle;

/* Last error message of the calling thread ("" if none). */
const char* pg_last_error(void);
/* Library version string. */
const char* pg_version(void);

/* Allocate a disk and fill it with the initial state: grid (init.cu:472-500),
   profiles (init.cu:502-532), planet (planet.cu:319-403) and orbital-advection
   rotation (orbit.cu:111-120); replaces main.cu:157-301.  params->ndim = 3
   creates the spherical adiabatic disk (init_flag 3: geomx/y/z = 2/4/5,
   EOS_flag 2, recon_flag 5 or 6, polar boundaries 2); its state arrays are
   x fastest over (zres, yres, xres).  The dust, wave-killing, averaging,
   torque and adjoint entry points are two-dimensional and return
   PG_E_INVALID for it. */
int pg_create(const pg_params* params, pg_handle** out);
/* Free everything the handle owns (main.cu:433-518). */
int pg_destroy(pg_handle* h);

/* Copy the interior state to/from host FP64 buffers (main.cu:61-104).
   gas: 5*xres*yres*zres doubles (rho,p,u,v,w); dust: 4*xres*yres (rho,u,v,w)
   or NULL.  On set, p is ignored under EOS_flag 0 (it is c_s^2(R)*rho). */
int pg_get_state(pg_handle* h, double* gas, double* dust);
int pg_set_state(pg_handle* h, const double* gas, const double* dust);

/* Ghost fills, one direction each (boundx / boundx_dust, boundy /
   boundy_dust, boundz; include/boundary.h:4-10, boundary.cu:13-374,
   462-718).  No ghost cells are stored on the device: the kernels read the
   radial ghosts from the analytic tables (type 0), the azimuthal ones by
   periodic indexing (type 3) and the polar ones by mirroring with w
   negated (type 2).  These return what those reads see: the current state
   with the 3 ghost cells on each side of the named direction filled as the
   reference fills them, field-major in pg_get_state's field order, that
   direction's extent n + 6 (x fastest; e.g. pg_bound_x: (xres+6)*yres*zres
   per field).  dust may be NULL (PG_E_STATE if given for a disk without
   dust); pg_bound_z needs a 3D disk (PG_E_INVALID otherwise). */
int pg_bound_x(pg_handle* h, double* gas, double* dust);
int pg_bound_y(pg_handle* h, double* gas, double* dust);
int pg_bound_z(pg_handle* h, double* gas, double* dust);

/* CFL step: dt = min(CFL / max signal rate, 1.1*dt_prev)
   (global_dt, timestep.cu:113-140). */
int pg_global_dt(pg_handle* h, double dt_prev, double* dt_out);
/* One full step at time t with step dt (solve, solver.h:4 / solve.cu:279-292:
   DS split order plus wave killing). */
int pg_solve(pg_handle* h, double t, double dt);
/* nsteps of the reference time loop (main.cu:320-359: dt from global_dt with
   the 1.1x cap, clipped to end_time) entirely on the device; t and dt live
   in device memory.  *t_io / *dt_io carry the time and the previous step in
   and out; dts (nsteps doubles, may be NULL) receives every step size.
   use_graph != 0 captures one step in a hipGraph and replays it. */
int pg_run(pg_handle* h, int nsteps, double* t_io, double* dt_io, double* dts, int use_graph);

/* ---- module-level entry points (reference module boundaries) ---------- */
/* viscous acceleration*rho of the current state (or of the predictor buffer
   if on_T): viscosity_tensor_evaluation1/2 (viscosity.cu:489-549).
   vis_out (2*xres*yres doubles, may be NULL) receives f_r, f_phi. */
int pg_viscosity_eval(pg_handle* h, int on_T, double* vis_out);
/* apply_source_terms(_inplace) (sweeps.cu:713-747) with the planet at
   `t_planet`; inplace = 0 writes the predictor buffer. */
int pg_sources(pg_handle* h, double t_planet, double mdt, double dt, int inplace);
/* gas and dust directional sweeps (sweeps.cu:337-456, dust.cu:240-353) */
int pg_sweep_x(pg_handle* h, double dt);
int pg_sweep_y(pg_handle* h, double dt);
int pg_sweep_x_dust(pg_handle* h, double dt);
int pg_sweep_y_dust(pg_handle* h, double dt);
/* polar-angle sweep of a 3D disk (sweepz_inplace, sweeps.cu:458-553, after
   the reflecting boundz of boundary.cu:270-321) */
int pg_sweep_z(pg_handle* h, double dt);
/* set_OrbAdv + shift_OrbAdv + advecty(_dust) (orbit.cu:122-157,
   advection.cu:321-357); inc/res (xres each, xres*zres in 3D, may be NULL)
   receive the integer shift and residual speed per radius (and polar layer) */
int pg_orbital_advection(pg_handle* h, double dt, int32_t* inc, double* res);
/* killwave (killwave.cu:105-117) */
int pg_killwave(pg_handle* h, double dt);
/* planet position and mass at time t (evolve_planet, planet.cu:293-306);
   body10 = m,a,e,x,y,z,vx,vy,vz,rs (may be NULL) */
int pg_planet(pg_handle* h, double t, double* body10);

/* ---- snapshots (the reference's checkpoint files) ------------------------ */
/* Write the current state in the layout of write_check_point
   (output.cu:125-180): FP64 t, xa[xres+1], ya[yres+1], rho, p, u, v
   (xres*yres each, x fastest), then with dust rho_d, u_d, v_d, w_d; 3D:
   t, xa, ya, za[zres+1], rho, p, u, v, w. */
int pg_write_snapshot(pg_handle* h, const char* filename, double t);
/* Restart from such a file (load_grid, output.cu:182-248); unlike the
   reference this also restores the dust fields.  *t_out receives the time. */
int pg_read_snapshot(pg_handle* h, const char* filename, double* t_out);
/* Run label of create_label (output.cu:14-93), e.g. "256x512_h50_1p33E_e0_a-30_OA_PPM4";
   the snapshot of frame n is "binary_" + label + "_" + 5-digit n (main.cu:403). */
int pg_snapshot_label(const pg_params* params, char* buf, size_t len);

/* ---- EOS (get_energy, EOS.h:4 / EOS.cu:4-18) ------------------------------ */
/* Specific energy of a cell under params->EOS_flag: 0 p/rho, 1 p/rho^gam,
   2 p/((gam-1) rho) (internal_e_flag 1, the only form the library runs). */
int pg_get_energy(const pg_params* params, double rho, double p, double u, double v, double w, double* e);

/* ---- diagnostics ------------------------------------------------------- */
/* Start a time average of the gas over a window of `interval`: zero the
   average (init_average, output.cu:316-329) and from now on add
   state * dt/2 / interval before and after every device-loop step
   (averaging, output.cu:299-314, as main.cu:344-350 calls it), in FP64. */
int pg_average_begin(pg_handle* h, double interval);
/* Stop adding to the average (it is kept). */
int pg_average_end(pg_handle* h);
/* The average so far in the pg_get_state gas layout (5 x xres*yres). */
int pg_average_get(pg_handle* h, double* gas);
/* Snapshot of the average in the write_check_point layout, as main.cu:409-414
   writes "binary_" + label + "_ave_" + frame: averaged gas, current dust. */
int pg_write_average_snapshot(pg_handle* h, const char* filename, double t);
/* Torque of the planet on the gas at the current state: the sum over cells
   of R gy rho dV with the planet at the end of the last step (global_sum
   with cal_torque and output_gy, output.cu:336-457, output/forces.cu:41-58,
   106-127), FP64.  main.cu:379-388 prints it divided by q^2/h^2 (2D). */
int pg_torque(pg_handle* h, double* torque);
/* pg_torque for callers that cannot allocate memory (serving loops,
   harnesses that forbid allocation): no counterpart in the reference.  The
   per-tile partial sums live in a fixed device array of the library (16384
   tiles, one per 64 x 16 cells of a large disk, 64 x 4 of a small one; a
   larger disk is refused with PG_E_INVALID), one copy per GPU, shared by
   every call on that GPU: calls that overlap on one GPU race.  Returns the
   same bits as pg_torque. */
int pg_torque_noalloc(pg_handle* h, double* torque);
/* Measurement, no counterpart in the reference: what a persistent step
   kernel would pay per pass boundary.  Runs npass (2 .. 100000) passes of a
   data-carrying pass over `blocks` workgroups of 256 threads (at most the
   co-resident count, else PG_E_INVALID) once as one cooperative launch with
   a grid-wide barrier after each pass and once as npass kernel nodes of a
   captured graph; returns the microseconds per barrier and per node (pass
   included in both) and checks the words the passes leave (PG_E_STATE if
   they are wrong). */
int pg_probe_grid_barrier(int blocks, int npass, double* us_per_barrier, double* us_per_node);

/* ---- introspection ----------------------------------------------------- */
/* HIP stream the handle launches on (as a void*), for external capture. */
int pg_stream(pg_handle* h, void** stream);
/* Average device time (microseconds) per launch of each kernel over the last
   pg_run call with profiling enabled; names are "sweep_x_gas", ... */
int pg_profile_enable(pg_handle* h, int on);
int pg_profile_get(pg_handle* h, const char* kernel, double* avg_us, int* count);
/* Bytes of device memory owned by the handle. */
int pg_device_bytes(pg_handle* h, size_t* bytes);

/* ---- adjoint of the step ---------------------------------------------------
   No counterpart in the reference (which has no derivatives).  Reverse mode
   of the whole DS step of a gas disk without dust or wave killing: source
   terms with the planet, viscosity, x sweep (PPM4 + HLLC), azimuthal sweep,
   orbital shift and residual advection, each transposed at the forward
   point with the step sizes, the integer shifts and every branch of the
   forward arithmetic frozen.  Parameter derivatives are for planet_mass
   (the planet's force is proportional to it) and ss_alpha (the viscous
   forces and the radial velocity -1.5 nu(R)/R of the initial state and of
   the radial ghost cells are proportional to it). */
typedef struct pg_grads {
    double dq;         /* dL/d planet_mass                                     */
    double dalpha;     /* dL/d ss_alpha                                        */
    double* dstate0;   /* 5*xres*yres (pg_get_state layout) or NULL: dL/d of   */
                       /* the start state, rho carrying p = c_s^2 rho          */
} pg_grads;
/* Run nsteps of the device loop from the current state (time t0, previous
   step dt0, as pg_run), keeping the state every ckpt_every steps and the
   step sizes (dts, nsteps doubles, may be NULL).  ckpt_every = 0 keeps every
   step's intermediates instead (the tape: 22 reals per cell and step of
   device memory, e.g. 46 MB per step for a 512x1024 FP32 disk), and the
   backward sweep then recomputes nothing; not with the time average on. */
int pg_adjoint_forward(pg_handle* h, int nsteps, int ckpt_every, double t0, double dt0, double* dts);
/* Given dL/d(final state) in pg_get_state layout (5*xres*yres; the p entry
   is folded in as c_s^2 dL/dp on rho), propagate back through every step of
   the last pg_adjoint_forward, recomputing each segment from its
   checkpoint (or reading the tape); fills *grads.  The handle's state is left at the forward's
   final state. */
int pg_adjoint_backward(pg_handle* h, const double* dL_dstate, pg_grads* grads);
/* Adjoint of one module call at the current state: the transpose of
   pass 0 pg_sweep_x, 1 pg_sweep_y, 2 pg_orbital_advection (gas),
   3 pg_viscosity_eval (lam_out has the 2 force fields), 4/5/6 pg_sources
   kick (mdt = 0) / predictor / corrector with the viscous forces the handle
   holds.  t and dt are the module call's arguments (t: the planet's time
   for the sources).  lam_out, lam_in: FP64, fields rho, u (radial), v
   (azimuthal), xres*yres each (the derivative with respect to v equals the
   one with respect to the stored deviation from the background rotation);
   dparam (2 doubles, may be NULL) receives dL/d planet_mass and
   dL/d ss_alpha of the pass. */
int pg_adjoint_pass(pg_handle* h, int pass, double t, double dt, const double* lam_out, double* lam_in,
                    double* dparam);

/* ---- batches of disks ------------------------------------------------------
   No counterpart in the reference, which runs one Grid per process
   (main.cu:157-359).  A batch advances n independent disks of uneven sizes
   and parameters together: each kernel launch covers every disk through a
   work list of (disk, block) pairs, every disk keeps its own CFL step and
   time in its device control block (the loop of main.cu:320-359 per disk),
   and one hipGraph replays a whole step of the batch.  Per disk the
   arithmetic is that of pg_run, so each disk's state is bitwise what a
   pg_handle holding it alone would produce.  All disks share precision,
   dust_flag, visc_flag and kill_flag; each must pass pg_create's checks. */
typedef struct pg_batch pg_batch;

int pg_batch_create(int n, const pg_params* params, pg_batch** out);
int pg_batch_destroy(pg_batch* b);
int pg_batch_size(pg_batch* b, int* n);
/* state of disk `disk` in pg_get_state's layout */
int pg_batch_get_state(pg_batch* b, int disk, double* gas, double* dust);
int pg_batch_set_state(pg_batch* b, int disk, const double* gas, const double* dust);
/* nsteps of every disk; t_io, dt_io: n entries each (time and previous step
   per disk, in and out); dts (n*nsteps, disk-major) may be NULL */
int pg_batch_run(pg_batch* b, int nsteps, double* t_io, double* dt_io, double* dts, int use_graph);
/* launch schedule: x-sweep tile width (128, 256, 384, 512, 1024, or 122,
   250, 506, whose tiles with their 6 halo cells fill 2, 4, 8 waves); threads
   per azimuthal tile (256: 16 radii x 16 rows, 512: 8 x 128, FP64 8 x 64, 1024: 16 x 64);
   azimuthal tile order (0: each XCD a contiguous band of tile rows of every
   disk, 1: row by row, 2: radius strip by strip); rows of the pointwise and
   stencil tiles (4, 8 or 16; 0 = chosen from the batch size) */
int pg_batch_set_schedule(pg_batch* b, int tile_x, int ring_b, int order, int pt_rows);
int pg_batch_get_schedule(pg_batch* b, int* tile_x, int* ring_b, int* order, int* pt_rows);
/* time `steps` captured steps of each schedule candidate on the batch's
   own state (restored afterwards), time the four fastest again over
   4*steps and keep the fastest of those; *best_ms (may be NULL) receives
   its device time per step */
int pg_batch_autotune(pg_batch* b, int steps, double* best_ms);
int pg_batch_profile_enable(pg_batch* b, int on);
int pg_batch_profile_get(pg_batch* b, const char* kernel, double* avg_us, int* count);
int pg_batch_device_bytes(pg_batch* b, size_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* PENGUIN_AMD_H */
