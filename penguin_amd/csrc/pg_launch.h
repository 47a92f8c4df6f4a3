// pg_launch.h -- host-side declarations of the kernel launchers defined in
// pg_kernels.hip (explicitly instantiated there for float and double).
#pragma once

#include <hip/hip_runtime.h>

#include "pg_internal.h"

namespace pg {

template <typename real>
struct Launch {
    static hipError_t sweep_x_gas(const DiskArgs<real>& d, hipStream_t s);
    static hipError_t sweep_x_dust(const DiskArgs<real>& d, hipStream_t s);
    static size_t ring_lds(int ny, int B);
    static hipError_t ring_gas(const DiskArgs<real>& d, hipStream_t s);
    static hipError_t ring_dust(const DiskArgs<real>& d, hipStream_t s);
    static hipError_t viscosity(const DiskArgs<real>& d, const real* st, hipStream_t s);
    static hipError_t sources(const DiskArgs<real>& d, int mode, int kill, int rate, hipStream_t s);
    static hipError_t cfl(const DiskArgs<real>& d, hipStream_t s);
    static hipError_t rate_reduce(const DiskArgs<real>& d, hipStream_t s);
};

template <typename real>
hipError_t launch_prologue(const DiskArgs<real>& d, int mode, double* dts, hipStream_t s);

}  // namespace pg
