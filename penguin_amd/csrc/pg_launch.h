// pg_launch.h -- host-side declarations of the kernel launchers defined in
// pg_kernels.hip (explicitly instantiated there for float and double).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "pg_internal.h"

namespace pg {

template <typename real>
struct Launch {
    static int phi_tile_radii(int B);
    static int phi_tile_rows(int B);
    // kick != 0: the first half kick applied to the loaded cells (dust: the
    // pre-kick gas is in the buffer opposite the current one)
    static hipError_t sweep_x_gas(const LaunchSet<real>& ls, int par, int kick, hipStream_t s);
    static hipError_t sweep_x_dust(const LaunchSet<real>& ls, int par, int kick, hipStream_t s);
    static hipError_t ring_gas(const LaunchSet<real>& ls, int par, int flags, hipStream_t s);
    static hipError_t ring_dust(const LaunchSet<real>& ls, int par, int flags, hipStream_t s);
    // gas and dust in one launch: the dust x sweep's drag reads the gas
    // buffer that is current at `par` (the one the gas sweep reads)
    static hipError_t sweep_x_both(const LaunchSet<real>& ls, int par, int kick, hipStream_t s);
    static hipError_t ring_both(const LaunchSet<real>& ls, int par, int flags, hipStream_t s);
    static hipError_t viscosity(const LaunchSet<real>& ls, int par, int which, hipStream_t s);
    static hipError_t sources(const LaunchSet<real>& ls, int par, int mode, int kill, int rate, hipStream_t s);
    // fused: predictor + viscosity of T (d.vis -> d.vis2); corrector + kill +
    // CFL rate + viscosity of the corrected state (d.vis2 -> d.vis)
    static hipError_t predvis(const LaunchSet<real>& ls, int par, hipStream_t s);
    static hipError_t corrvis(const LaunchSet<real>& ls, int par, int kill, hipStream_t s);
    static hipError_t average(const LaunchSet<real>& ls, int par, hipStream_t s);
    static hipError_t torque(const LaunchSet<real>& ls, int par, double* part, hipStream_t s);
    static hipError_t cfl(const LaunchSet<real>& ls, int par, hipStream_t s);
    static hipError_t rate_reduce(const LaunchSet<real>& ls, hipStream_t s);
    static hipError_t prologue(const LaunchSet<real>& ls, int mode, int do_orbadv, hipStream_t s);
    // replay of a recorded step: t = td[0], dt = td[1] (one disk)
    static hipError_t prologue_replay(const LaunchSet<real>& ls, const double* td, hipStream_t s);
    // n replays: descriptor b * stride of ls.disks from t = td[2b], dt = td[2b+1]
    static hipError_t prologue_replay_many(const LaunchSet<real>& ls, int stride, int n, const double* td,
                                           hipStream_t s);
};

// Adjoint passes (pg_adjoint.hip), one disk, grid from nx, ny.
template <typename real>
struct AdjLaunch {
    static hipError_t x(const AdjArgs<real>& A, int nx, int ny, hipStream_t s);
    static hipError_t phi(const AdjArgs<real>& A, int nx, int ny, hipStream_t s);
    static hipError_t adv(const AdjArgs<real>& A, int nx, int ny, hipStream_t s);
    static hipError_t vis(const AdjArgs<real>& A, int mode, int nx, int ny, hipStream_t s);
    static hipError_t src(const AdjArgs<real>& A, int nx, int ny, hipStream_t s);
};

// Grid-barrier probe (pg_probe.hip): microseconds per grid-wide barrier of
// one cooperative launch and per kernel node of a captured graph, for the
// same data-carrying pass over `blocks` workgroups of 256 threads.
int probe_grid_barrier(int blocks, int npass, double* us_per_barrier, double* us_per_node, std::string& msg);

}  // namespace pg
