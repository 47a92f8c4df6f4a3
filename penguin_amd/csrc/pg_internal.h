// pg_internal.h -- data shared between the host module code (pg_host.cpp)
// and the HIP kernels (pg_kernels.hip).
//
// Device data layout (one disk):
//   state     SoA, interior cells only, radius fastest: field f of cell (i,j)
//             at base[f*N + i + nx*j], N = nx*ny.
//             gas  : rho - rho_b(R) (the density's deviation from the
//                    unperturbed initial disk rho_b), u (radial velocity),
//                    vp (azimuthal velocity minus the background rotation
//                    vbg(R) = initial v_phi(R))
//             dust : rho_d, u_d, vp_d (same background subtracted)
//             Pressure is not stored: under the locally isothermal EOS the
//             reference resets p = c_s^2(R) rho after every update
//             (sweeps.cu:255-256, advection.cu:100-102).
//   ghosts    radial ghost cells hold the analytic initial state (boundary
//             type 0, boundary.cu:30-33) and are read from per-radius
//             tables; azimuthal ghosts are periodic (type 3) and come from
//             modular indexing.
//   tables    per padded radius ip = i + NPAD (i interior, ip in
//             [0, nx + 2*NPAD)), precomputed in FP64 on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../../include/pg_params.h"

namespace pg {

constexpr int NPAD = 3;
// Tiles of the pointwise and stencil kernels: PT_X radii x pt_rows rows
// (a launch-set parameter, a multiple of PT_TY up to PT_YMAX), 256 threads
// of PT_X x PT_TY, each thread covering pt_rows / PT_TY rows.  Tall tiles
// keep the one-cell halo of the stencil kernels small; short ones give a
// small disk more workgroups.
constexpr int PT_X = 64, PT_TY = 4, PT_YMAX = 16;
// Radii per tile of the x-sweep adjoint: with its 12 halo slots, 128 slots
// fill two waves, one thread per slot (the host sizes the per-workgroup
// parameter sums of the adjoint from it).
constexpr int ADJ_TX = 116;
constexpr int NGAS = 3;
constexpr int NDUST = 3;
constexpr int RING_THREADS = 512;   // default threads per azimuthal-pass block

struct Body {
    double m, a, e, x, y, z, vx, vy, vz, rs;
};

// A planet as the source kernel uses it: mass, orbital radius, softening,
// m/x^2 (indirect term) and the sine/cosine of its angle and half angle.
struct PlanetK {
    double m, x, rs, m_over_x2, sphi, cphi, shphi, chphi;
};

// What the source kernel needs of a planet, precomputed once per step.
__host__ __device__ inline PlanetK planet_kernel_view(const Body& b) {
    PlanetK k;
    k.m = b.m;
    k.x = b.x;
    k.rs = b.rs;
    k.m_over_x2 = b.m / (b.x * b.x);
    k.sphi = sin(b.y);
    k.cphi = cos(b.y);
    k.shphi = sin(0.5 * b.y);
    k.chphi = cos(0.5 * b.y);
    return k;
}

// Per-step control block, device resident, always FP64.
struct StepCtl {
    double t;            // time at the start of the step
    double dt;           // step size
    double dt_prev;      // previous step size (1.1x cap)
    double end_time;
    unsigned long long rate_bits;  // max CFL rate of the state, bits of a non-negative double
    int nstep;
    int mode;            // 0: dt supplied by host, 1: dt from rate (device loop)
    Body planet0;        // planet at t     (first half kick)
    Body planet1;        // planet at t+dt  (predictor and corrector)
    PlanetK pk0, pk1;    // the same, as the source kernel reads them
    double ave_interval; // time-average window (the averaging weight is dt/2 over it)
    double ave_w;        // sum of the averaging weights so far
    int err;             // device loop: 1 once a step found a CFL rate that is not finite and positive
    int err_step;        // the step at which it did
};

// Per-radius tables (FP64 on the host; cast to the storage type on upload).
enum TabField {
    T_XA = 0,     // left edge of padded cell ip (nx+7 entries used)
    T_XV,         // cell volume per unit angle, (xr^2-xl^2)/2 for geomx 1
    T_XC,         // centre 0.5*(xl+xr)
    T_IDX,        // 1/(xr-xl)
    T_CS2,        // c_s^2(R)
    T_VBG,        // background rotation v_phi,init(R)
    T_EBAL,       // vbg^2/R - 1/R^2 (centrifugal minus stellar gravity at vbg)
    T_NU,         // kinematic viscosity nu(R)
    T_DVBG,       // (vbg[ip+1]-vbg[ip-1])/(xc[ip+1]-xc[ip-1])
    T_D2VBG,      // second difference of vbg on the centre grid (viscosity.cu:12)
    T_WL0, T_WL1, T_WL2,   // 4th-order left-slope weights
    T_WR0, T_WR1, T_WR2,   // 4th-order right-slope weights
    T_G_R, T_G_U, T_G_V,   // gas initial state rho, u, vp (vp = 0)
    T_D_R, T_D_U, T_D_V,   // dust initial state rho_d, u_d, vp_d
    T_KILL,       // wave-killing rate without dt: sin^2(pi/2*d)/R^1.5, 0 outside
    T_C0,         // isothermal sound speed sqrt(c_s^2)
    T_I20,        // 1/(xc[ip+1]-xc[ip-1])
    T_I21,        // 1/(xc[ip+1]-xc[ip])
    T_I10,        // 1/(xc[ip]-xc[ip-1])
    T_IR,         // 1/xc
    T_IXV,        // 1/xv
    T_IC0,        // 1/c_s
    T_TS,         // St * h0 / c_s: dust stopping time times rho (dust.cu:1-4)
    T_ITS,        // 1/T_TS
    T_TBR,        // rho_b[ip] - rho_b[ip-1]: background density difference across edge ip
    T_TBP,        // p_b[ip] - p_b[ip-1], p_b = c_s^2 rho_b
    T_PSRC,       // background pressure force of the x sweep without dt:
                  // -(xa[ip] + xa[ip+1]) (T_TBP[ip] + T_TBP[ip+1]) / 4
    T_NFIELDS
};

template <typename real>
struct DiskArgs {
    int nx, ny;
    int has_dust, visc, orbadv, kill;
    int n_planet;
    real* gas[2];        // 3*N state, ping-pong pair; the launch's parity picks the current one
    real* dust[2];       // 3*N
    real* vis;           // 2*N viscous acceleration*rho of the step start (V0)
    real* vis2;          // 2*N the same of the predicted state (V1)
    const real* tab;     // per-radius tables, field-major: tab[f*tab_stride + ip]
    int tab_stride;      // nx + 2*NPAD + 1 radii
    const real* phitab;  // 4*ny: sin(phi_j), cos(phi_j), sin(phi_j/2), cos(phi_j/2)
    real dy;             // azimuthal cell width (angle), uniform grid
    real idy;            // 1/dy
    real wLy[3], wRy[3]; // 4th-order weights of the uniform azimuthal grid
    real gam, sc_h, Stokes, D_G_ratio, frame_omega;
    double CFL;
    double dy_d;         // dy in FP64 for the orbital shift
    const double* rot;   // 3*nx FP64: [0,nx) vbg(R) - R*frame_omega
                         // (orbit.cu:65), [nx,2nx) R of the interior radii,
                         // [2nx,3nx) the background density rho_b(R)
    int* inc;            // nx, integer azimuthal shift of the step
    real* res;           // nx, residual advection speed of the step
    StepCtl* ctl;
    double ramp_time, planet_mass, planet_ecc, planet_radius, rs;
    double* rate_part;   // per-block maxima of the CFL rate (rate_nblk entries)
    int rate_nblk;
    double* dts;         // step sizes recorded by the device loop (or null)
    int dts_cap;         // entries of dts
    // intermediate states the adjoint's recompute keeps (null otherwise):
    // the gas after the first half kick (x sweep), after the azimuthal
    // sweep before the orbital shift (azimuthal pass), and the predicted
    // state (predictor), 3*N each
    real* out_kick;
    real* out_swept;
    real* out_pred;
    double* ave;         // 3*N FP64 time average of the gas (rho, u, vp), or null
};

// Buffers of one adjoint pass (pg_adjoint.hip): the forward inputs the
// pass recomputes from, the adjoint of its output and where the adjoint of
// its inputs goes.  One disk.
template <typename real>
struct AdjArgs {
    const DiskArgs<real>* d;   // the disk: tables, control block (dt, planets, shift)
    const real* s;             // forward input state of the pass (3N)
    const real* vis;           // viscous forces the pass read (2N)
    const real* s2;            // fused viscosity adjoint: the state the viscosity read (3N)
    const real* vis2;          // fused viscosity adjoint: the forces the predictor read (2N)
    const real* lo;            // adjoint of the pass output (3N)
    real* li;                  // adjoint of the pass input state (3N)
    real* lvis;                // adjoint of vis (2N)
    double* lparam;            // per-workgroup partial sums: [0, nslot) dL/d planet_mass,
                               // [nslot, 2 nslot) dL/d ss_alpha
    int nslot;                 // slots per parameter (>= the largest adjoint grid)
    int accumulate;            // 1: li += adjoint, 0: li = adjoint
    int mode;                  // source-term mode, 0 kick / 1 predictor / 2 corrector
    double q, alpha;           // planet_mass, ss_alpha
};

// A set of disks launched together (one disk, or a batch).  Each kernel's
// grid walks a work list of (disk, local block) pairs, so one launch covers
// every disk of the set whatever their sizes.
template <typename real>
struct LaunchSet {
    const DiskArgs<real>* disks;
    int ndisk;
    const int2* wl_x;    // x sweeps: local block = tile + tiles_per_row * row
    int n_x;
    int tile_x;          // x-sweep tile width (compile-time variants)
    const int2* wl_ring; // rings: local block = radius (-1: padding)
    int n_ring;
    int ring_b;          // ring chunk (threads per block)
    int ny_max;          // largest ring, sizes the LDS
    const int2* wl_pt;   // pointwise and stencil kernels: local block = tile tx + ntx * ty
    int n_pt;
    int pt_rows;         // rows of their tiles (PT_TY .. PT_YMAX)
};

}  // namespace pg
