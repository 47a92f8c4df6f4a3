// pg_kernels.hip -- gfx950 kernels of the Godunov step (2D, locally
// isothermal gas + pressureless dust, PPM4 + HLLC).
//
// Every kernel is templated on the storage type `real` (float or double);
// arithmetic runs in that type.  Data layout and tables: pg_internal.h.
//
// Kernels and the reference routines they compute:
//   k_sweep_x_gas   sweepx_inplace  (sweeps.cu:263-358) + riemann()
//   k_sweep_x_dust  sweepx_dust_inplace (dust.cu:178-257)
//   k_ring_gas      sweepy_inplace (sweeps.cu:361-456), then the orbital
//                   shift (orbit.cu:17-49) and the residual advection
//                   (advection.cu:1-173) of one radius, fused: the ring of
//                   one radius stays in LDS between the three stages
//   k_ring_dust     the same for dust (dust.cu:260-353, advection.cu:175-319)
//   k_viscosity     viscous_forces (viscosity.cu:34-197)
//   k_sources       apply_source_terms (sweeps.cu:555-711), optionally
//                   followed by killwave (killwave.cu:17-103) and the CFL
//                   rate of the result (timestep.cu:18-97)
//   k_cfl           get_dt_lv1/lv2 (timestep.cu:18-111)
//   k_prologue      dt selection (timestep.cu:133-139, main.cu:322-329),
//                   planet orbit (planet.cu:105-187), orbital shift and
//                   residual speed per radius (orbit.cu:74-99)
//
// The reference evaluates the block-cooperative phases of riemann()
// (flatten -> traced states -> HLLC -> flux difference) with shared-memory
// barriers between them; here each phase writes LDS arrays over a tile and a
// barrier separates the phases, and each cell's parabola is computed once
// (the reference recomputes it for both of its interfaces).
#include <hip/hip_runtime.h>

#include <cmath>

#include "pg_internal.h"

namespace pg {

// A launch's parity word: bit 0 selects the current gas buffer of every disk
// in the set, bit 1 the current dust buffer; a kernel that writes a fluid
// writes the other buffer of its pair.
__device__ __forceinline__ int gsel(int par) { return par & 1; }
__device__ __forceinline__ int dsel(int par) { return (par >> 1) & 1; }


// per-radius tables, field-major: lanes that differ in radius read
// consecutive addresses
#define TAB(f, ip) d.tab[(size_t)(f) * d.tab_stride + (ip)]

// Division, reciprocal, square root and exponential.  FP32: the hardware
// approximations (v_rcp_f32, v_sqrt_f32, v_exp_f32; about 1 ulp); FP64: IEEE.
template <typename real> struct FM;
template <> struct FM<float> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float exp(float x) { return __expf(x); }
};
template <> struct FM<double> {
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ double div(double a, double b) { return a / b; }
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
};

// ---------------------------------------------------------------- helpers
template <typename real>
__device__ __forceinline__ real lim01(real a) {  // solve.cu:14-17
    return fmin(fmax(a, real(0)), real(1));
}
template <typename real>
__device__ __forceinline__ real exp_lim(real x) {  // solve.cu:29-34
    return fmax(real(1) + x, real(1.0e-6));
}

// 4th-order slopes from precomputed geometric weights (the linear map
// get_4th_parameters applies to the four neighbouring differences,
// recon.cu:216-357), limiter (recon.cu:361-376), flattening and the PPM
// monotonicity correction (PPM.cu:64-91).
template <typename real>
__device__ __forceinline__ void slopes(real am2, real am1, real a0, real ap1, real ap2,
                                       real wl0, real wl1, real wl2, real wr0, real wr1, real wr2,
                                       real flat, real& sL, real& sR) {
    real d0 = am1 - am2, d1 = a0 - am1, d2 = ap1 - a0, d3 = ap2 - ap1;
    real l = wl0 * d0 + wl1 * d1 + wl2 * d2;
    real r = wr0 * d1 + wr1 * d2 + wr2 * d3;
    l = copysign(fmin(fabs(l), fabs(d1)), d1);
    r = copysign(fmin(fabs(r), fabs(d2)), d2);
    if (r * l <= real(0)) {
        l = real(0);
        r = real(0);
    }
    real fk = real(1) - flat;
    r *= fk;
    l *= fk;
    real t3 = r + l;
    real t2 = t3 * t3;
    real t1 = t3 * real(3) * (l - r);
    if (t2 < t1) l = real(2) * r;
    else if (t2 < -t1) r = real(2) * l;
    sL = l;
    sR = r;
}

// Mean of the cell parabola over its first fraction q (get_PPM_aveL,
// PPM.cu:106-113) and over its last fraction x (get_PPM_aveR with x = 1-q,
// PPM.cu:95-104).
template <typename real>
__device__ __forceinline__ real aveL(real q, real sL, real a, real sR) {
    real a6 = real(3) * (sL - sR);
    return (a - sL) + (q * real(0.5)) * ((sR + sL) + (real(1) - q * real(2.0 / 3.0)) * a6);
}
template <typename real>
__device__ __forceinline__ real aveR(real x, real sL, real a, real sR) {
    real a6 = real(3) * (sL - sR);
    return (a + sR) - (x * real(0.5)) * ((sR + sL) - (real(1) - x * real(2.0 / 3.0)) * a6);
}

// Flattening coefficient of the centre cell of a 5-cell window
// (flatten, PPM.cu:7-62).  cs is the cell's sound speed.
template <typename real>
__device__ __forceinline__ real flatten5(real pm2, real pm1, real p0, real pp1, real pp2,
                                         real um1, real u0, real up1, real cs) {
    real f = real(0);
    real dp1f = pp1 - p0, dp2f = pp2 - pm1, dp1b = p0 - pm1, dp2b = pp1 - pm2;
    // Mf > 0 || Mb > 0 with Mf = (u0-up1)/cs, Mb = (um1-u0)/cs and cs > 0
    (void)cs;
    if (u0 - up1 > real(0) || um1 - u0 > real(0)) {
        if (dp2f != real(0) && dp2b != real(0)) {
            real ff = real(10) * fmax(real(0), fabs(FM<real>::div(dp1f, dp2f) - real(0.5)) - real(0.25));
            real fb = real(10) * fmax(real(0), fabs(FM<real>::div(dp1b, dp2b) - real(0.5)) - real(0.25));
            f = fmax(ff, fb);
            f = fmin(real(0.5), FM<real>::div(f * fmax(fabs(dp1f), fabs(dp1b)), p0));
        }
    }
    return f;
}

// Characteristic-traced half-step state on the right of an interface from
// the cell right of it (set_state, states.cu:151-206), locally isothermal
// (EOS_flag 0 fixes gam = 1, parameters.h:118-119, so exp_lim(gam*x) =
// exp_lim(x) and the sound speed of the traced state is sqrt(p0/r0)).
// c: the cell's sound speed; idx = 1/dx.  Returns the state and its sound
// speed cout (= the one HLLC recomputes from pl/rl, wave_speeds.cu:290).
template <typename real>
__device__ __forceinline__ void trace_right(real c, real dt, real idx, real r, real p, real u,
                                            real rsL, real rsR, real psL, real psR, real usL, real usR,
                                            real vsL, real v, real vsR,
                                            real& rr, real& pr, real& ur, real& vr, real& cout) {
    real q = lim01((c - fmin(u, real(0))) * dt * idx);
    real ua = aveL(q, usL, u, usR);
    real dis = fmin(ua, real(0)) * dt;
    q = lim01(-dis * idx);
    real r0 = fmax(aveL(q, rsL, r, rsR), r * real(1.0e-10));
    real p0 = fmax(aveL(q, psL, p, psR), p * real(1.0e-10));
    real u0 = aveL(q, usL, u, usR);
    real ir0 = FM<real>::rcp(r0);
    real cs = FM<real>::sqrt(p0 * ir0);
    q = lim01((cs * dt - dis) * idx);
    real u_ = aveL(q, usL, u, usR);
    real p_ = aveL(q, psL, p, psR);
    real ics = FM<real>::rcp(cs);
    real e = exp_lim((u0 - u_) * ics);
    rr = r0 * e;
    pr = p0 * e;
    ur = u0 + (p0 - p_) * ics * ir0;
    vr = aveL(q, vsL, v, vsR);
    cout = cs;
}

// ... and on the left from the cell left of it (states.cu:208-264).
template <typename real>
__device__ __forceinline__ void trace_left(real c, real dt, real idx, real r, real p, real u,
                                           real rsL, real rsR, real psL, real psR, real usL, real usR,
                                           real vsL, real v, real vsR,
                                           real& rl, real& pl, real& ul, real& vl, real& cout) {
    real x = lim01((fmax(u, real(0)) + c) * dt * idx);
    real ua = aveR(x, usL, u, usR);
    real dis = fmax(ua, real(0)) * dt;
    x = lim01(dis * idx);
    real r0 = fmax(aveR(x, rsL, r, rsR), r * real(1.0e-10));
    real p0 = fmax(aveR(x, psL, p, psR), p * real(1.0e-10));
    real u0 = aveR(x, usL, u, usR);
    real ir0 = FM<real>::rcp(r0);
    real cs = FM<real>::sqrt(p0 * ir0);
    x = lim01((dis + cs * dt) * idx);
    real u_ = aveR(x, usL, u, usR);
    real p_ = aveR(x, psL, p, psR);
    real ics = FM<real>::rcp(cs);
    real e = exp_lim((u_ - u0) * ics);
    rl = r0 * e;
    pl = p0 * e;
    ul = u0 + (p_ - p0) * ics * ir0;
    vl = aveR(x, vsL, v, vsR);
    cout = cs;
}

// The same with one sound speed c for every cell of the row (p = c^2 rho
// with c fixed: the azimuthal sweep of a locally isothermal disk); pressure
// is c^2 times density throughout, so only rho, u, v are traced.
template <typename real>
__device__ __forceinline__ void trace_right_c(real c, real ic, real dt, real idx, real r, real u,
                                              real rsL, real rsR, real usL, real usR,
                                              real vsL, real v, real vsR,
                                              real& rr, real& ur, real& vr) {
    real q = lim01((c - fmin(u, real(0))) * dt * idx);
    real ua = aveL(q, usL, u, usR);
    real dis = fmin(ua, real(0)) * dt;
    q = lim01(-dis * idx);
    real r0 = fmax(aveL(q, rsL, r, rsR), r * real(1.0e-10));
    real u0 = aveL(q, usL, u, usR);
    q = lim01((c * dt - dis) * idx);
    real u_ = aveL(q, usL, u, usR);
    real ra = aveL(q, rsL, r, rsR);
    rr = r0 * exp_lim((u0 - u_) * ic);
    ur = u0 + c * FM<real>::div(r0 - ra, r0);
    vr = aveL(q, vsL, v, vsR);
}

template <typename real>
__device__ __forceinline__ void trace_left_c(real c, real ic, real dt, real idx, real r, real u,
                                             real rsL, real rsR, real usL, real usR,
                                             real vsL, real v, real vsR,
                                             real& rl, real& ul, real& vl) {
    real x = lim01((fmax(u, real(0)) + c) * dt * idx);
    real ua = aveR(x, usL, u, usR);
    real dis = fmax(ua, real(0)) * dt;
    x = lim01(dis * idx);
    real r0 = fmax(aveR(x, rsL, r, rsR), r * real(1.0e-10));
    real u0 = aveR(x, usL, u, usR);
    x = lim01((dis + c * dt) * idx);
    real u_ = aveR(x, usL, u, usR);
    real ra = aveR(x, rsL, r, rsR);
    rl = r0 * exp_lim((u_ - u0) * ic);
    ul = u0 + c * FM<real>::div(ra - r0, r0);
    vl = aveR(x, vsL, v, vsR);
}

// HLLC selection (HLLC.cu:2-77, EOS_flag 0) given the wave speeds.
template <typename real>
__device__ __forceinline__ void hllc_select(real pm, real sl, real sm, real sr,
                                            real rl, real pl, real ul, real vl,
                                            real rr, real pr, real ur, real vr,
                                            real& Fr, real& us, real& vs, real& pres) {
    if (!signbit(sl)) {
        Fr = ul * rl; us = ul; vs = vl; pres = pl;
    } else if (signbit(sr)) {
        Fr = ur * rr; us = ur; vs = vr; pres = pr;
    } else if (!signbit(sm)) {
        real k = FM<real>::rcp(sl - sm);
        Fr = sm * rl * (sl - ul) * k; us = ul; vs = vl; pres = (sl * pm - sm * pl) * k;
    } else {
        real k = FM<real>::rcp(sr - sm);
        Fr = sm * rr * (sr - ur) * k; us = ur; vs = vr; pres = (sr * pm - sm * pr) * k;
    }
    us *= Fr;
    vs *= Fr;
}

// Wave speeds (get_pm_simple, get_lr_speeds_Einfeldt, get_sm;
// wave_speeds.cu:24-35,288-333; gam = 1) with the states' sound speeds cl,
// cr, then the HLLC fluxes: mass flux, upwinded normal and transverse
// momentum flux, interface pressure.
template <typename real>
__device__ __forceinline__ void hllc_iso(real cl, real cr, real rl, real pl, real ul, real vl,
                                         real rr, real pr, real ur, real vr,
                                         real& Fr, real& us, real& vs, real& pres) {
    real csum = cl + cr;
    real ics = FM<real>::rcp(csum);
    real pm = real(0.5) * (pr + pl) * exp_lim((ul - ur) * ics);
    real srl = FM<real>::sqrt(rl), srr = FM<real>::sqrt(rr);
    real iss = FM<real>::rcp(srl + srr);
    real ubar = (srl * ul + srr * ur) * iss;
    real n2 = real(0.5) * srl * srr * iss * iss;
    real du = ur - ul;
    real cbar = FM<real>::sqrt((cl * cl * srl + cr * cr * srr) * iss + n2 * du * du);
    real sl = ubar - cbar, sr = ubar + cbar;
    real sm = FM<real>::div(rl * ul * cl + rr * ur * cr, rl * cl + rr * cr) +
              real(2) * (pl - pr) * ics * FM<real>::rcp(rl + rr);
    hllc_select(pm, sl, sm, sr, rl, pl, ul, vl, rr, pr, ur, vr, Fr, us, vs, pres);
}

// The same with one sound speed c on both sides (azimuthal isothermal sweep):
// cbar^2 = c^2 + n2 du^2 and sm = (rl (ul + c) + rr (ur - c)) / (rl + rr).
template <typename real>
__device__ __forceinline__ void hllc_c(real c, real ic, real c2, real rl, real ul, real vl, real rr, real ur, real vr,
                                       real& Fr, real& us, real& vs, real& pres) {
    real pl = c2 * rl, pr = c2 * rr;
    real pm = real(0.5) * (pr + pl) * exp_lim(real(0.5) * (ul - ur) * ic);
    real srl = FM<real>::sqrt(rl), srr = FM<real>::sqrt(rr);
    real iss = FM<real>::rcp(srl + srr);
    real ubar = (srl * ul + srr * ur) * iss;
    real n2 = real(0.5) * srl * srr * iss * iss;
    real du = ur - ul;
    real cbar = FM<real>::sqrt(c2 + n2 * du * du);
    real sl = ubar - cbar, sr = ubar + cbar;
    real sm = FM<real>::div(rl * (ul + c) + rr * (ur - c), rl + rr);
    hllc_select(pm, sl, sm, sr, rl, pl, ul, vl, rr, pr, ur, vr, Fr, us, vs, pres);
}

// Pressureless dust flux through an interface (dust_flux, dust.cu:6-105)
// from the parabolas of the cells left (L) and right (R) of it.
// idxL, idxR: 1/dx of those cells; dt: the sweep's time step in the
// coordinate's units.
template <typename real>
__device__ __forceinline__ void dust_iface(real dt,
                                           real idxL, real rL, real rLsL, real rLsR, real uL, real uLsL,
                                           real uLsR, real vL, real vLsL, real vLsR,
                                           real idxR, real rR, real rRsL, real rRsR, real uR, real uRsL,
                                           real uRsR, real vR, real vRsL, real vRsR,
                                           real& Fr, real& Fu, real& Fv) {
    // right state from the right cell
    real y = -fmin(uR - uRsL, real(0)) * dt * idxR;
    real q = lim01(y);
    real ur = aveL(q, uRsL, uR, uRsR);
    real du = FM<real>::div(dt * (ur - uR) * idxR, real(1) - y);
    q = lim01(-fmin(ur, real(0)) * dt * idxR);
    real srr = aveL(q, rRsL, rR, rRsR) * exp_lim(du);
    real sur = aveL(q, uRsL, uR, uRsR);
    real svr = aveL(q, vRsL, vR, vRsR);
    // left state from the left cell
    y = fmax(uL + uLsR, real(0)) * dt * idxL;
    real x = lim01(y);
    real ul = aveR(x, uLsL, uL, uLsR);
    du = FM<real>::div(dt * (uL - ul) * idxL, real(1) - y);
    x = lim01(fmax(ul, real(0)) * dt * idxL);
    real srl = aveR(x, rLsL, rL, rLsR) * exp_lim(du);
    real sul = aveR(x, uLsL, uL, uLsR);
    real svl = aveR(x, vLsL, vL, vLsR);
    // donor-cell caps (dust.cu:96-97)
    srl = fmin(rL * dt * idxL, fmax(sul * srl, real(0)));
    srr = fmax(-rR * dt * idxR, fmin(sur * srr, real(0)));
    Fr = srl + srr;
    Fu = srl * sul + srr * sur;
    Fv = srl * svl + srr * svr;
}

// ------------------------------------------------------------ planet (FP64)
__host__ __device__ inline double ramp_function(double t, double tau, double val) {  // planet.cu:7-17
    return t < tau ? val * sin(0.5 * 3.1415926535897932384626434 * t / tau) : val;
}

__host__ __device__ inline double ecc_anomaly(double n, double e, double t) {  // planet.cu:60-103
    const double twopi = 2.0 * 3.1415926535897932384626434;
    const double pi = 3.1415926535897932384626434;
    double l = fmod(n * t, twopi);
    double D = -e * sin(l);
    if (e == 0.0 || l == 0.0 || l == pi || l == twopi) return l;
    double step;
    int count = 0;
    do {
        double es = e * sin(l - D), ec = e * cos(l - D);
        double x = D + es, dx = 1.0 - ec, d2x = -es;
        double eps = 2.0 * x * d2x / dx / dx;
        if (fabs(eps) < 1.0) step = -(x / dx) * (1.0 + eps / 4.0 + eps * eps / 8.0 + 5.0 * eps * eps * eps / 64.0);
        else step = copysign(fmin(fabs(x / dx), 0.25 * pi), -x);
        D += step;
        count++;
    } while ((fabs(step) > 1.0e-12 || count < 10) && count < 200);
    return l - D;
}

// orbit_solution (planet.cu:105-133): analytic Kepler orbit in the frame
// rotating at frame_omega; mass ramp of planet.cu:7-17.
__host__ __device__ inline void orbit_solution(Body& b, double t, double ramp_time, double mass,
                                               double frame_omega) {
    const double pi = 3.1415926535897932384626434;
    double a = b.a, e = b.e;
    double n = sqrt(1.0 / a / a / a);
    double efac = sqrt(1.0 - e * e);
    double B = e / (1.0 + efac);
    double u = ecc_anomaly(n, e, t);
    double f = u + 2.0 * atan(B * sin(u) / (1.0 - B * cos(u)));
    double esinf = e * sin(f), ecosf = e * cos(f);
    b.m = ramp_function(t, ramp_time, mass);
    b.x = a * efac * efac / (1.0 + ecosf);
    b.y = f - fmod(frame_omega * t, 2.0 * pi) + pi;
    b.z = 0.5 * pi;
    b.vx = n * a * esinf / efac;
    b.vy = n * (1.0 + ecosf) * (1.0 + ecosf) / efac / efac / efac;
    b.vz = 0.0;
}

// ------------------------------------------------------------ x sweep, gas
// One block = TX interior cells of one azimuthal row (the whole row when
// nr <= 1024) plus 3 ghost/halo cells each side, staged in LDS; thread t
// owns slot t (interior cell t-3 of the tile).  Phases (barrier between):
// load -> flattening -> parabolas -> interface fluxes -> conservative update.
// TX is a compile-time constant so every LDS array offset is an immediate.
template <typename real, int TX>
__global__ __launch_bounds__(TX) void k_sweep_x_gas(LaunchSet<real> ls, int par) {
    const int2 w = ls.wl_x[blockIdx.x];
    const DiskArgs<real>& d = ls.disks[w.x];
    constexpr int S = TX + 2 * NPAD;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    real* sm = reinterpret_cast<real*>(smem_raw);
    real* s_r = sm;
    real* s_p = sm + S;
    real* s_u = sm + 2 * S;
    real* s_L = sm + 3 * S;
    real* s_vp = sm + 4 * S;
    real* s_f = sm + 5 * S;
    real* s_sl = sm + 6 * S;    // 4 fields: r, p, u, L
    real* s_sr = sm + 10 * S;
    real* s_Fr = sm + 14 * S;
    real* s_Fu = sm + 15 * S;
    real* s_FL = sm + 16 * S;
    real* s_pr = sm + 17 * S;
    const int nx = d.nx;
    const size_t N = (size_t)nx * d.ny;
    const int t = threadIdx.x;
    const int tiles = (nx + TX - 1) / TX;
    const int j = w.y / tiles;
    const int x0 = (w.y - j * tiles) * TX;
    const int nt = min(TX, nx - x0);
    const int ns = nt + 2 * NPAD;
    const real dt = (real)d.ctl->dt;
    const real* __restrict__ in = d.gas[gsel(par)];
    const size_t row = (size_t)j * nx;

    for (int s = t; s < ns; s += TX) {
        const int i = x0 - NPAD + s;
        const int ip = i + NPAD;
        real r, u, vp;
        if (i < 0 || i >= nx) {
            r = TAB(T_G_R, ip); u = TAB(T_G_U, ip); vp = TAB(T_G_V, ip);
        } else {
            r = in[row + i]; u = in[N + row + i]; vp = in[2 * N + row + i];
        }
        s_r[s] = r;
        s_p[s] = TAB(T_CS2, ip) * r;
        s_u[s] = u;
        s_vp[s] = vp;
        s_L[s] = TAB(T_XC, ip) * (TAB(T_VBG, ip) + vp);
    }
    __syncthreads();
    for (int s = t + 2; s < ns - 2; s += TX) {
        s_f[s] = flatten5(s_p[s - 2], s_p[s - 1], s_p[s], s_p[s + 1], s_p[s + 2],
                          s_u[s - 1], s_u[s], s_u[s + 1], real(1));
    }
    __syncthreads();
    for (int s = t + 2; s < ns - 2; s += TX) {
        const int ip = x0 + s;
        const real wl0 = TAB(T_WL0, ip), wl1 = TAB(T_WL1, ip), wl2 = TAB(T_WL2, ip);
        const real wr0 = TAB(T_WR0, ip), wr1 = TAB(T_WR1, ip), wr2 = TAB(T_WR2, ip);
        const real fl = s_f[s];
        real l, r;
        slopes(s_r[s - 2], s_r[s - 1], s_r[s], s_r[s + 1], s_r[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, fl, l, r);
        s_sl[s] = l; s_sr[s] = r;
        slopes(s_p[s - 2], s_p[s - 1], s_p[s], s_p[s + 1], s_p[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, fl, l, r);
        s_sl[S + s] = l; s_sr[S + s] = r;
        slopes(s_u[s - 2], s_u[s - 1], s_u[s], s_u[s + 1], s_u[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, fl, l, r);
        s_sl[2 * S + s] = l; s_sr[2 * S + s] = r;
        slopes(s_L[s - 2], s_L[s - 1], s_L[s], s_L[s + 1], s_L[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, fl, l, r);
        s_sl[3 * S + s] = l; s_sr[3 * S + s] = r;
    }
    __syncthreads();
    for (int k = t + NPAD; k < ns - 2; k += TX) {
        const int ipk = x0 + k;
        real rl, pl, ul, vl, rr, pr, ur, vr, cl, cr;
        trace_left(TAB(T_C0, ipk - 1), dt, TAB(T_IDX, ipk - 1), s_r[k - 1], s_p[k - 1], s_u[k - 1],
                   s_sl[k - 1], s_sr[k - 1], s_sl[S + k - 1], s_sr[S + k - 1],
                   s_sl[2 * S + k - 1], s_sr[2 * S + k - 1],
                   s_sl[3 * S + k - 1], s_L[k - 1], s_sr[3 * S + k - 1], rl, pl, ul, vl, cl);
        trace_right(TAB(T_C0, ipk), dt, TAB(T_IDX, ipk), s_r[k], s_p[k], s_u[k],
                    s_sl[k], s_sr[k], s_sl[S + k], s_sr[S + k], s_sl[2 * S + k], s_sr[2 * S + k],
                    s_sl[3 * S + k], s_L[k], s_sr[3 * S + k], rr, pr, ur, vr, cr);
        real Fr, Fu, FL, pres;
        hllc_iso(cl, cr, rl, pl, ul, vl, rr, pr, ur, vr, Fr, Fu, FL, pres);
        const real g = TAB(T_XA, ipk);  // eval_gfac(1, xa) = edge radius
        s_Fr[k] = Fr * g;
        s_Fu[k] = Fu * g;
        s_FL[k] = FL * g;
        s_pr[k] = pres;
    }
    __syncthreads();
    real* __restrict__ out = d.gas[gsel(par) ^ 1];
    for (int s = t + NPAD; s < ns - NPAD; s += TX) {
        const int ip = x0 + s;
        const int i = ip - NPAD;
        const real g0 = TAB(T_XA, ip), g1 = TAB(T_XA, ip + 1);
        const real f = dt * TAB(T_IXV, ip);
        const real Dr = (s_Fr[s] - s_Fr[s + 1]) * f;
        const real Du = (s_Fu[s] - s_Fu[s + 1] + real(0.5) * (g0 + g1) * (s_pr[s] - s_pr[s + 1])) * f;
        const real DL = (s_FL[s] - s_FL[s + 1]) * f;
        const real r = s_r[s], u = s_u[s], vp = s_vp[s];
        real rn = r + Dr;
        const real rmin = fmax(r * real(1.0e-10), real(1.0e-10));
        real un = u, vn = vp;
        if (rn < rmin) {
            rn = rmin;
        } else {
            const real irn = FM<real>::rcp(rn);
            un = (r * u + Du) * irn;
            vn = (r * vp + (DL * TAB(T_IR, ip) - TAB(T_VBG, ip) * Dr)) * irn;
        }
        out[row + i] = rn;
        out[N + row + i] = un;
        out[2 * N + row + i] = vn;
    }
}

// ----------------------------------------------------------- x sweep, dust
template <typename real, int TX>
__global__ __launch_bounds__(TX) void k_sweep_x_dust(LaunchSet<real> ls, int par) {
    const int2 w = ls.wl_x[blockIdx.x];
    const DiskArgs<real>& d = ls.disks[w.x];
    constexpr int S = TX + 2 * NPAD;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    real* sm = reinterpret_cast<real*>(smem_raw);
    real* s_r = sm;
    real* s_u = sm + S;
    real* s_L = sm + 2 * S;
    real* s_vp = sm + 3 * S;
    real* s_sl = sm + 4 * S;  // 3 fields: r, u, L
    real* s_sr = sm + 7 * S;
    real* s_Fr = sm + 10 * S;
    real* s_Fu = sm + 11 * S;
    real* s_FL = sm + 12 * S;
    const int nx = d.nx;
    const size_t N = (size_t)nx * d.ny;
    const int t = threadIdx.x;
    const int tiles = (nx + TX - 1) / TX;
    const int j = w.y / tiles;
    const int x0 = (w.y - j * tiles) * TX;
    const int nt = min(TX, nx - x0);
    const int ns = nt + 2 * NPAD;
    const real dt = (real)d.ctl->dt;
    const real* __restrict__ in = d.dust[dsel(par)];
    const size_t row = (size_t)j * nx;

    for (int s = t; s < ns; s += TX) {
        const int i = x0 - NPAD + s;
        const int ip = i + NPAD;
        real r, u, vp;
        if (i < 0 || i >= nx) {
            r = TAB(T_D_R, ip); u = TAB(T_D_U, ip); vp = TAB(T_D_V, ip);
        } else {
            r = in[row + i]; u = in[N + row + i]; vp = in[2 * N + row + i];
        }
        s_r[s] = r;
        s_u[s] = u;
        s_vp[s] = vp;
        s_L[s] = TAB(T_XC, ip) * (TAB(T_VBG, ip) + vp);
    }
    __syncthreads();
    for (int s = t + 2; s < ns - 2; s += TX) {
        const int ip = x0 + s;
        const real wl0 = TAB(T_WL0, ip), wl1 = TAB(T_WL1, ip), wl2 = TAB(T_WL2, ip);
        const real wr0 = TAB(T_WR0, ip), wr1 = TAB(T_WR1, ip), wr2 = TAB(T_WR2, ip);
        real l, r;
        slopes(s_r[s - 2], s_r[s - 1], s_r[s], s_r[s + 1], s_r[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, real(0), l, r);
        s_sl[s] = l; s_sr[s] = r;
        slopes(s_u[s - 2], s_u[s - 1], s_u[s], s_u[s + 1], s_u[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, real(0), l, r);
        s_sl[S + s] = l; s_sr[S + s] = r;
        slopes(s_L[s - 2], s_L[s - 1], s_L[s], s_L[s + 1], s_L[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, real(0), l, r);
        s_sl[2 * S + s] = l; s_sr[2 * S + s] = r;
    }
    __syncthreads();
    for (int k = t + NPAD; k < ns - 2; k += TX) {
        const int ipk = x0 + k;
        real Fr, Fu, FL;
        dust_iface(dt, TAB(T_IDX, ipk - 1), s_r[k - 1], s_sl[k - 1], s_sr[k - 1],
                   s_u[k - 1], s_sl[S + k - 1], s_sr[S + k - 1],
                   s_L[k - 1], s_sl[2 * S + k - 1], s_sr[2 * S + k - 1],
                   TAB(T_IDX, ipk), s_r[k], s_sl[k], s_sr[k], s_u[k], s_sl[S + k], s_sr[S + k],
                   s_L[k], s_sl[2 * S + k], s_sr[2 * S + k], Fr, Fu, FL);
        const real g = TAB(T_XA, ipk);
        s_Fr[k] = Fr * g;
        s_Fu[k] = Fu * g;
        s_FL[k] = FL * g;
    }
    __syncthreads();
    real* __restrict__ out = d.dust[dsel(par) ^ 1];
    const real rfloor = d.D_G_ratio * real(1.0e-10);
    for (int s = t + NPAD; s < ns - NPAD; s += TX) {
        const int ip = x0 + s;
        const int i = ip - NPAD;
        const real f = dt * TAB(T_IXV, ip);
        const real Dr = (s_Fr[s] - s_Fr[s + 1]) * f;
        const real Du = (s_Fu[s] - s_Fu[s + 1]) * f;
        const real DL = (s_FL[s] - s_FL[s + 1]) * f;
        const real r = s_r[s], u = s_u[s], vp = s_vp[s];
        real rn = r + Dr;
        const real rmin = fmax(r * real(1.0e-10), rfloor);
        real un = u, vn = vp;
        if (rn < rmin) {
            rn = rmin;
        } else {
            const real irn = FM<real>::rcp(rn);
            un = (r * u + Du) * irn;
            vn = (r * vp + (DL * TAB(T_IR, ip) - TAB(T_VBG, ip) * Dr)) * irn;
        }
        out[row + i] = rn;
        out[N + row + i] = un;
        out[2 * N + row + i] = vn;
    }
}

// ------------------------------------------------------- azimuthal ring
// One block = one radius.  Stage A sweeps the ring in chunks of B cells
// (periodic halo) and writes each updated cell to its orbitally shifted
// position (j + inc) mod ny of an LDS ring; stage B advects that ring by the
// residual speed and writes the result to HBM.
template <typename real, int B>
struct RingLds {
    static constexpr int S = B + 2 * NPAD;
    real* ring;  // 3*ny
    real* b;     // chunk scratch: 14 arrays of S
    __device__ __forceinline__ real* c0() const { return b; }
    __device__ __forceinline__ real* c1() const { return b + S; }
    __device__ __forceinline__ real* c2() const { return b + 2 * S; }
    __device__ __forceinline__ real* fl() const { return b + 3 * S; }
    __device__ __forceinline__ real* sl0() const { return b + 4 * S; }
    __device__ __forceinline__ real* sr0() const { return b + 5 * S; }
    __device__ __forceinline__ real* sl1() const { return b + 6 * S; }
    __device__ __forceinline__ real* sr1() const { return b + 7 * S; }
    __device__ __forceinline__ real* sl2() const { return b + 8 * S; }
    __device__ __forceinline__ real* sr2() const { return b + 9 * S; }
    __device__ __forceinline__ real* F0() const { return b + 10 * S; }
    __device__ __forceinline__ real* F1() const { return b + 11 * S; }
    __device__ __forceinline__ real* F2() const { return b + 12 * S; }
    __device__ __forceinline__ real* F3() const { return b + 13 * S; }
};

__device__ __forceinline__ int wrap(int j, int n) {
    j %= n;
    return j < 0 ? j + n : j;
}

// parabolas of the three ring fields of slot s (flattening coefficient fl)
template <typename real, int B>
__device__ __forceinline__ void ring_slopes(const RingLds<real, B>& L, int s, real fl, const real* w) {
    real l, r;
    const real* c0 = L.c0();
    const real* c1 = L.c1();
    const real* c2 = L.c2();
    slopes(c0[s - 2], c0[s - 1], c0[s], c0[s + 1], c0[s + 2], w[0], w[1], w[2], w[3], w[4], w[5], fl, l, r);
    L.sl0()[s] = l; L.sr0()[s] = r;
    slopes(c1[s - 2], c1[s - 1], c1[s], c1[s + 1], c1[s + 2], w[0], w[1], w[2], w[3], w[4], w[5], fl, l, r);
    L.sl1()[s] = l; L.sr1()[s] = r;
    slopes(c2[s - 2], c2[s - 1], c2[s], c2[s + 1], c2[s + 2], w[0], w[1], w[2], w[3], w[4], w[5], fl, l, r);
    L.sl2()[s] = l; L.sr2()[s] = r;
}

// Stage B (residual advection, advection.cu:1-105) over the shifted ring;
// `dust` selects update_Dust's floor (advection.cu:315) over
// update_Cell_advect (advection.cu:88-105, no floor).
template <typename real, int B>
__device__ __forceinline__ void ring_advect(const DiskArgs<real>& d, const RingLds<real, B>& L, int i, real* __restrict__ out,
                                            real speed, real frac, real fad, bool dust, real rfloor, const real* w) {
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int t = threadIdx.x;
    const real q = lim01(frac);
    for (int c0 = 0; c0 < ny; c0 += B) {
        const int nt = min(B, ny - c0);
        const int ns = nt + 2 * NPAD;
        for (int s = t; s < ns; s += B) {
            const int jj = wrap(c0 - NPAD + s, ny);
            L.c0()[s] = L.ring[jj];
            L.c1()[s] = L.ring[ny + jj];
            L.c2()[s] = L.ring[2 * ny + jj];
        }
        __syncthreads();
        for (int s = t + 2; s < ns - 2; s += B) ring_slopes(L, s, real(0), w);
        __syncthreads();
        for (int k = t + NPAD; k < ns - 2; k += B) {
            real Dr, Dv, Du;
            if (speed > real(0)) {
                const int c = k - 1;
                Dr = aveR(q, L.sl0()[c], L.c0()[c], L.sr0()[c]);
                Dv = aveR(q, L.sl1()[c], L.c1()[c], L.sr1()[c]) * Dr;
                Du = aveR(q, L.sl2()[c], L.c2()[c], L.sr2()[c]) * Dr;
            } else {
                const int c = k;
                Dr = aveL(q, L.sl0()[c], L.c0()[c], L.sr0()[c]);
                Dv = aveL(q, L.sl1()[c], L.c1()[c], L.sr1()[c]) * Dr;
                Du = aveL(q, L.sl2()[c], L.c2()[c], L.sr2()[c]) * Dr;
            }
            L.F0()[k] = Dr * speed; L.F1()[k] = Dv * speed; L.F2()[k] = Du * speed;
        }
        __syncthreads();
        for (int s = t + NPAD; s < ns - NPAD; s += B) {
            const int jj = c0 - NPAD + s;
            const real Dr = (L.F0()[s] - L.F0()[s + 1]) * fad;
            const real Dv = (L.F1()[s] - L.F1()[s + 1]) * fad;
            const real Du = (L.F2()[s] - L.F2()[s + 1]) * fad;
            const real r = L.c0()[s];
            real rn = r + Dr;
            real vn = L.c1()[s], un = L.c2()[s];
            if (dust && rn < fmax(r * real(1.0e-10), rfloor)) {
                rn = fmax(r * real(1.0e-10), rfloor);
            } else {
                const real irn = FM<real>::rcp(rn);
                vn = (r * L.c1()[s] + Dv) * irn;
                un = (r * L.c2()[s] + Du) * irn;
            }
            const size_t o = (size_t)jj * nx + i;
            out[o] = rn;
            out[N + o] = un;
            out[2 * N + o] = vn;
        }
        __syncthreads();
    }
}

template <typename real, int B>
__device__ __forceinline__ void ring_copy_out(const DiskArgs<real>& d, const RingLds<real, B>& L, int i, real* __restrict__ out) {
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    for (int jj = threadIdx.x; jj < ny; jj += B) {
        const size_t o = (size_t)jj * nx + i;
        out[o] = L.ring[jj]; out[N + o] = L.ring[2 * ny + jj]; out[2 * N + o] = L.ring[ny + jj];
    }
}

// flags: bit 0 = sweep stage, bit 1 = orbital shift + residual advection
template <typename real, int B>
__global__ __launch_bounds__(B) void k_ring_gas(LaunchSet<real> ls, int par, int flags) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    real* sm = reinterpret_cast<real*>(smem_raw);
    const int2 w = ls.wl_ring[blockIdx.x];
    const int i = w.y;
    if (i < 0) return;
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int t = threadIdx.x;
    const bool sweep_on = flags & 1;
    const bool orbadv = (flags & 2) && d.orbadv;
    RingLds<real, B> L{sm, sm + 3 * ny};
    const int ip = i + NPAD;
    const real cs2 = TAB(T_CS2, ip);
    const real cs0 = TAB(T_C0, ip);
    const real ic0 = TAB(T_IC0, ip);
    const real iR = TAB(T_IR, ip);
    const real dt = (real)d.ctl->dt;
    const real dts = dt * iR;          // set_state receives dt/rad (riemann.cu:111)
    const real idy = d.idy;
    const real fup = dt * idy * iR;    // Del.multiply(dt/yv/rad_cyl), sweeps.cu:412
    const int inc = orbadv ? d.inc[i] : 0;
    const real* __restrict__ in = d.gas[gsel(par)];
    const real wy[6] = {d.wLy[0], d.wLy[1], d.wLy[2], d.wRy[0], d.wRy[1], d.wRy[2]};

    // ---- stage A: azimuthal sweep (rho, vp normal, u transverse) ----
    for (int c0 = 0; c0 < ny; c0 += B) {
        const int nt = min(B, ny - c0);
        const int ns = nt + 2 * NPAD;
        for (int s = t; s < ns; s += B) {
            const int jj = wrap(c0 - NPAD + s, ny);
            const size_t o = (size_t)jj * nx + i;
            L.c0()[s] = in[o];            // rho
            L.c1()[s] = in[2 * N + o];    // vp: azimuthal velocity in the frame of rot + R*Omega_f
            L.c2()[s] = in[N + o];        // u: radial velocity (transverse here)
        }
        __syncthreads();
        for (int s = t + 2; s < ns - 2; s += B) {
            // under p = c_s^2 rho with c_s^2 fixed on the ring, flatten's
            // pressure ratios equal the density ratios
            const real* c0p = L.c0();
            const real* c1p = L.c1();
            L.fl()[s] = flatten5(c0p[s - 2], c0p[s - 1], c0p[s], c0p[s + 1], c0p[s + 2],
                                 c1p[s - 1], c1p[s], c1p[s + 1], cs0);
        }
        __syncthreads();
        for (int s = t + 2; s < ns - 2; s += B) ring_slopes(L, s, L.fl()[s], wy);
        __syncthreads();
        for (int k = t + NPAD; k < ns - 2; k += B) {
            real rl, ul, vl, rr, ur, vr;
            // p = c_s^2 rho with c_s^2 fixed on the ring: the pressure
            // parabola is c_s^2 times the density parabola
            trace_left_c(cs0, ic0, dts, idy, L.c0()[k - 1], L.c1()[k - 1], L.sl0()[k - 1], L.sr0()[k - 1],
                         L.sl1()[k - 1], L.sr1()[k - 1], L.sl2()[k - 1], L.c2()[k - 1], L.sr2()[k - 1], rl, ul, vl);
            trace_right_c(cs0, ic0, dts, idy, L.c0()[k], L.c1()[k], L.sl0()[k], L.sr0()[k],
                          L.sl1()[k], L.sr1()[k], L.sl2()[k], L.c2()[k], L.sr2()[k], rr, ur, vr);
            real Fr, Fu, Fw, pres;
            hllc_c(cs0, ic0, cs2, rl, ul, vl, rr, ur, vr, Fr, Fu, Fw, pres);
            L.F0()[k] = Fr; L.F1()[k] = Fu; L.F2()[k] = Fw; L.F3()[k] = pres;
        }
        __syncthreads();
        for (int s = t + NPAD; s < ns - NPAD; s += B) {
            const int jj = c0 - NPAD + s;
            const real Dr = (L.F0()[s] - L.F0()[s + 1]) * fup;
            const real Du = (L.F1()[s] - L.F1()[s + 1] + (L.F3()[s] - L.F3()[s + 1])) * fup;
            const real Dw = (L.F2()[s] - L.F2()[s + 1]) * fup;
            const real r = L.c0()[s], vp = L.c1()[s], u = L.c2()[s];
            real rn = r + Dr;
            const real rmin = fmax(r * real(1.0e-10), real(1.0e-10));
            real vn = vp, un = u;
            if (!sweep_on) {
                rn = r;
            } else if (rn < rmin) {
                rn = rmin;
            } else {
                // (rho v + Du + (rot + R Omega_f) Dr) / rho' - vbg == (rho vp + Du) / rho'
                const real irn = FM<real>::rcp(rn);
                vn = (r * vp + Du) * irn;
                un = (r * u + Dw) * irn;
            }
            const int js = wrap(jj + inc, ny);
            L.ring[js] = rn;
            L.ring[ny + js] = vn;
            L.ring[2 * ny + js] = un;
        }
        __syncthreads();
    }
    if (!orbadv) {
        ring_copy_out(d, L, i, d.gas[gsel(par) ^ 1]);
        return;
    }
    // ---- stage B: residual advection at speed res (advection.cu:1-105) ----
    const real speed = d.res[i];
    const real frac = fabs(speed) * dts * idy;   // |speed| dt / (R dy)
    ring_advect(d, L, i, d.gas[gsel(par) ^ 1], speed, frac, fup, false, real(0), wy);
}

// flags: bit 0 = sweep stage, bit 1 = orbital shift + residual advection
template <typename real, int B>
__global__ __launch_bounds__(B) void k_ring_dust(LaunchSet<real> ls, int par, int flags) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    real* sm = reinterpret_cast<real*>(smem_raw);
    const int2 w = ls.wl_ring[blockIdx.x];
    const int i = w.y;
    if (i < 0) return;
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int t = threadIdx.x;
    const bool sweep_on = flags & 1;
    const bool orbadv = (flags & 2) && d.orbadv;
    RingLds<real, B> L{sm, sm + 3 * ny};
    const int ip = i + NPAD;
    const real iR = TAB(T_IR, ip);
    const real dt = (real)d.ctl->dt;
    const real dts = dt * iR;          // dust_flow: dt /= rad for geom > 2 (dust.cu:127)
    const real idy = d.idy;
    const real f12 = dt * idy * iR;    // dust.cu:309-313
    const int inc = orbadv ? d.inc[i] : 0;
    const real* __restrict__ in = d.dust[dsel(par)];
    const real rfloor = d.D_G_ratio * real(1.0e-10);
    const real wy[6] = {d.wLy[0], d.wLy[1], d.wLy[2], d.wRy[0], d.wRy[1], d.wRy[2]};

    for (int c0 = 0; c0 < ny; c0 += B) {
        const int nt = min(B, ny - c0);
        const int ns = nt + 2 * NPAD;
        for (int s = t; s < ns; s += B) {
            const int jj = wrap(c0 - NPAD + s, ny);
            const size_t o = (size_t)jj * nx + i;
            L.c0()[s] = in[o];
            L.c1()[s] = in[2 * N + o];
            L.c2()[s] = in[N + o];
        }
        __syncthreads();
        for (int s = t + 2; s < ns - 2; s += B) ring_slopes(L, s, real(0), wy);
        __syncthreads();
        for (int k = t + NPAD; k < ns - 2; k += B) {
            real Fr, Fu, Fw;
            dust_iface(dts, idy, L.c0()[k - 1], L.sl0()[k - 1], L.sr0()[k - 1], L.c1()[k - 1], L.sl1()[k - 1],
                       L.sr1()[k - 1], L.c2()[k - 1], L.sl2()[k - 1], L.sr2()[k - 1],
                       idy, L.c0()[k], L.sl0()[k], L.sr0()[k], L.c1()[k], L.sl1()[k], L.sr1()[k],
                       L.c2()[k], L.sl2()[k], L.sr2()[k], Fr, Fu, Fw);
            L.F0()[k] = Fr; L.F1()[k] = Fu; L.F2()[k] = Fw;
        }
        __syncthreads();
        for (int s = t + NPAD; s < ns - NPAD; s += B) {
            const int jj = c0 - NPAD + s;
            const real Dr = (L.F0()[s] - L.F0()[s + 1]) * f12;
            const real Du = (L.F1()[s] - L.F1()[s + 1]) * f12;
            const real Dw = (L.F2()[s] - L.F2()[s + 1]) * f12;
            const real r = L.c0()[s], vp = L.c1()[s], u = L.c2()[s];
            real rn = r + Dr;
            const real rmin = fmax(r * real(1.0e-10), rfloor);
            real vn = vp, un = u;
            if (!sweep_on) {
                rn = r;
            } else if (rn < rmin) {
                rn = rmin;
            } else {
                const real irn = FM<real>::rcp(rn);
                vn = (r * vp + Du) * irn;
                un = (r * u + Dw) * irn;
            }
            const int js = wrap(jj + inc, ny);
            L.ring[js] = rn;
            L.ring[ny + js] = vn;
            L.ring[2 * ny + js] = un;
        }
        __syncthreads();
    }
    if (!orbadv) {
        ring_copy_out(d, L, i, d.dust[dsel(par) ^ 1]);
        return;
    }
    const real speed = d.res[i];
    const real frac = fabs(speed) * dts * idy;
    ring_advect(d, L, i, d.dust[dsel(par) ^ 1], speed, frac, f12, true, rfloor, wy);
}

// ----------------------------------------------------------------- viscosity
// 9-point stencil of viscous_forces (viscosity.cu:34-197, ndim 2).  The
// derivatives of v_phi = vbg(R) + vp are split into the tabulated
// background part and the stored deviation (the difference operators are
// linear, so this is the same arithmetic rearranged).
// which = 0: viscosity of the current state; 1: of the other buffer (the
// predictor, viscosity_tensor_evaluation2)
template <typename real>
__global__ __launch_bounds__(256) void k_viscosity(LaunchSet<real> ls, int par, int which) {
    const int2 w = ls.wl_pt[blockIdx.x];
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int segs = (nx + 255) >> 8;
    const int j = w.y / segs;
    const int i = (w.y - j * segs) * 256 + threadIdx.x;
    const real* __restrict__ st = d.gas[gsel(par) ^ which];
    if (i >= nx) return;
    const int ip = i + NPAD;
    real nu[3];
    for (int n = 0; n < 3; ++n) nu[n] = TAB(T_NU, ip - 1 + n);
    real nr[9], u[9], v[9];
    const int jm = wrap(j - 1, ny), jp = wrap(j + 1, ny);
    const int rows[3] = {j, jm, jp};
#pragma unroll
    for (int b = 0; b < 3; ++b) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int n = 3 * b + a;
            int ii = i - 1 + a;
            real rr, uu, vv;
            if (ii < 0 || ii >= nx) {
                rr = TAB(T_G_R, ii + NPAD); uu = TAB(T_G_U, ii + NPAD); vv = TAB(T_G_V, ii + NPAD);
            } else {
                size_t o = (size_t)rows[b] * nx + ii;
                rr = st[o]; uu = st[N + o]; vv = st[2 * N + o];
            }
            nr[n] = rr * nu[a];
            u[n] = uu;
            v[n] = vv;  // deviation vp
        }
    }
    // reciprocal differences of the centre grid (viscosity.cu:9-31)
    const real i20 = TAB(T_I20, ip), i21 = TAB(T_I21, ip), i10 = TAB(T_I10, ip);
    const real ir = TAB(T_IR, ip);
    const real idy = d.idy;
    const real hdy = real(0.5) * idy;            // 1/(t2 - t0)
    // radial derivatives
    real dudx = (u[2] - u[0]) * i20;
    real d2udx2 = ((u[2] - u[1]) * i21 - (u[1] - u[0]) * i10) * (real(2) * i20);
    real dvdx = TAB(T_DVBG, ip) + (v[2] - v[0]) * i20;
    real d2vdx2 = TAB(T_D2VBG, ip) + ((v[2] - v[1]) * i21 - (v[1] - v[0]) * i10) * (real(2) * i20);
    real dndx = (nr[2] - nr[0]) * i20;
    // azimuthal and mixed derivatives
    real dvdy = (v[7] - v[4]) * hdy * ir;
    real d2vdy2 = (v[7] - real(2) * v[1] + v[4]) * (idy * idy) * ir * ir;
    real d2vdxy = (v[8] - v[6] - v[5] + v[3]) * i20 * hdy * ir;
    real dudy = (u[7] - u[4]) * hdy * ir;
    real d2udy2 = (u[7] - real(2) * u[1] + u[4]) * (idy * idy) * ir * ir;
    real d2udxy = (u[8] - u[6] - u[5] + u[3]) * i20 * hdy * ir;
    real dndy = (nr[7] - nr[4]) * hdy * ir;
    const real v1 = TAB(T_VBG, ip) + v[1];
    const real n1 = nr[1];
    real sxx = real(2) * n1 * dudx;
    real fx = real(2) * dndx * dudx + real(2) * n1 * d2udx2 + sxx * ir;
    real syy = real(2) * n1 * (dvdy + u[1] * ir);
    real sxy = n1 * (dudy + dvdx - v1 * ir);
    real fy = real(2) * dndy * dvdy + real(2) * n1 * d2vdy2 + real(2) * n1 * dudy * ir + real(2) * dndy * u[1] * ir;
    fx += dndy * (dudy + dvdx) + n1 * (d2udy2 + d2vdxy) - dndy * v1 * ir - n1 * dvdy * ir - syy * ir;
    fy += dndx * (dudy + dvdx) + n1 * (d2udxy + d2vdx2) - dndx * v1 * ir - n1 * dvdx * ir + n1 * v1 * ir * ir +
          real(2) * sxy * ir;
    size_t o = (size_t)j * nx + i;
    d.vis[o] = fx;
    d.vis[N + o] = fy;
}

// ------------------------------------------------------------- sources
// Planet acceleration (star_planet_grav_cyl, forces.cu:41-58, z = 0) at
// angle phi_j from the tabulated sin/cos of phi_j and phi_j/2 and the
// planet's (sin, cos) of phi_p and phi_p/2.  The squared distance is written
// without cancellation: (R - a)^2 + 4 a R sin^2((phi - phi_p)/2) + rs^2.
template <typename real>
__device__ __forceinline__ void planet_acc(const PlanetK& pk, const real* pt, real R, real& ax, real& ay) {
    const real sp = (real)pk.sphi, cp = (real)pk.cphi, shp = (real)pk.shphi, chp = (real)pk.chphi;
    const real s = pt[0] * cp - pt[1] * sp;      // sin(phi - phi_p)
    const real c = pt[1] * cp + pt[0] * sp;      // cos(phi - phi_p)
    const real h = pt[2] * chp - pt[3] * shp;    // sin((phi - phi_p)/2)
    const real px = (real)pk.x, m = (real)pk.m, rs = (real)pk.rs;
    const real Rp2 = (R - px) * (R - px) + real(4) * px * R * h * h + rs * rs;
    const real iRp = FM<real>::rcp(FM<real>::sqrt(Rp2));
    const real k = iRp * iRp * iRp;
    const real ind = (real)pk.m_over_x2;
    ax = -m * (R - px * c) * k - c * ind;
    ay = -m * px * s * k + s * ind;
}

// Block max of a non-negative value, written to this block's slot of the
// partial-max array (reduced by the prologue or k_rate_reduce): one store per
// block instead of a contended atomic on one address.
template <typename real>
__device__ __forceinline__ void block_max_store(double* part, real v) {
    double x = (double)v;
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    __shared__ double wmax[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wmax[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < (int)((blockDim.x + 63) >> 6); ++k) m = fmax(m, wmax[k]);
        part[0] = m;
    }
}

// Max over the partial array by one block; result in thread 0.  Eight
// independent loads per iteration keep the memory pipeline full.
__device__ __forceinline__ double reduce_partials(const double* part, int n) {
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int stride = blockDim.x * 8;
    int k0 = threadIdx.x;
    for (; k0 + 7 * (int)blockDim.x < n; k0 += stride) {
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = fmax(a[u], part[k0 + u * blockDim.x]);
    }
    for (int k = k0; k < n; k += blockDim.x) a[0] = fmax(a[0], part[k]);
    double x = fmax(fmax(fmax(a[0], a[1]), fmax(a[2], a[3])), fmax(fmax(a[4], a[5]), fmax(a[6], a[7])));
    for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
    __shared__ double wmax[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wmax[w] = x;
    __syncthreads();
    double m = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)((blockDim.x + 63) >> 6); ++k) m = fmax(m, wmax[k]);
    return m;
}

// 1 - exp(-x) (the drag factor dtau, sweeps.cu:673,702)
template <typename real>
__device__ __forceinline__ real one_minus_exp_neg(real x);
template <>
__device__ __forceinline__ float one_minus_exp_neg(float x) { return 1.0f - __expf(-x); }
template <>
__device__ __forceinline__ double one_minus_exp_neg(double x) { return -expm1(-x); }

// Pointwise source terms.  mode: 0 first half kick (planet at t, in place),
// 1 predictor (planet at t+dt, gas written to gasT), 2 corrector (planet at
// t+dt, in place) optionally followed by wave killing and the CFL rate,
// 3 wave killing alone.
template <typename real>
__global__ __launch_bounds__(256) void k_sources(LaunchSet<real> ls, int par, int mode, int do_kill, int do_rate) {
    const int2 w = ls.wl_pt[blockIdx.x];
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int segs = (nx + 255) >> 8;
    const int j = w.y / segs;
    const int i = (w.y - j * segs) * 256 + threadIdx.x;
    real* __restrict__ gas = d.gas[gsel(par)];
    real* __restrict__ gasT = d.gas[gsel(par) ^ 1];
    real* __restrict__ dust = d.dust[dsel(par)];
    real rate = real(0);
    if (i < nx) {
        const int ip = i + NPAD;
        const size_t o = (size_t)j * nx + i;
        const real dt = (real)(0.5 * d.ctl->dt);
        const real mdt = mode == 0 ? real(0) : dt;
        const PlanetK& pk = mode == 0 ? d.ctl->pk0 : d.ctl->pk1;
        const real R = TAB(T_XC, ip);
        const real iR = TAB(T_IR, ip);
        const real vbg = TAB(T_VBG, ip);
        const real ebal = TAB(T_EBAL, ip);
        real px = real(0), py = real(0);
        if (d.n_planet > 0) planet_acc(pk, d.phitab + 4 * j, R, px, py);
        const real rho = gas[o], cu = gas[N + o], cvp = gas[2 * N + o];
        const real irho = FM<real>::rcp(rho);
        real vx = real(0), vy = real(0);
        if (d.visc) { vx = d.vis[o] * irho; vy = d.vis[N + o] * irho; }
        const real ay = vy + py;               // fy + gy, independent of the state
        real Tu = cu, Tvp = cvp;
        real Qu = 0, Qvp = 0, Du = 0, Dvp = 0, ts = 0, its = 0;
        if (d.has_dust) {
            Qu = dust[N + o]; Qvp = dust[2 * N + o];
            Du = Qu; Dvp = Qvp;
            ts = TAB(T_TS, ip) * irho;            // St / (c_s/h0) / rho (dust.cu:1-4)
            its = rho * TAB(T_ITS, ip);
        }
        // fx + gx at azimuthal velocity vbg + vp: centrifugal - stellar
        // gravity (tabulated at vbg) + deviation part + planet (+ viscosity)
        auto axd = [&](real vp) { return ebal + vp * (real(2) * vbg + vp) * iR + px; };
        auto axg = [&](real vp) { return axd(vp) + vx; };
        if (mode == 1 || mode == 2) {
            Tu = cu + axg(cvp) * mdt;
            Tvp = cvp + ay * mdt;
            if (d.has_dust) {
                const real dtau = one_minus_exp_neg(mdt * its);
                Du = Qu + (axd(Dvp) * ts + cu - Qu) * dtau;
                Dvp = Qvp + ((cvp - Qvp) + py * ts) * dtau;
            }
        }
        real nu_ = cu + axg(Tvp) * dt;
        real nvp = cvp + ay * dt;
        real nr_ = rho;
        if (mode == 3) { nu_ = cu; nvp = cvp; }   // wave killing alone
        if (mode == 1) {
            gasT[o] = nr_; gasT[N + o] = nu_; gasT[2 * N + o] = nvp;
        } else {
            real ndu = Qu, ndvp = Qvp;
            if (d.has_dust && mode != 3) {
                const real dtau = one_minus_exp_neg(dt * its);
                ndu = Qu + (axd(Dvp) * ts + Tu - Qu) * dtau;
                ndvp = Qvp + ((Tvp - Qvp) + py * ts) * dtau;
            }
            if (do_kill) {   // killwave on the corrected state (solve.cu:287-289)
                const real kk = TAB(T_KILL, ip) * (real)d.ctl->dt;
                nr_ += kk * (TAB(T_G_R, ip) - nr_);
                nu_ += kk * (TAB(T_G_U, ip) - nu_);
                nvp += kk * (real(0) - nvp);
            }
            if (do_kill) gas[o] = nr_;   // rho changes only by wave killing
            gas[N + o] = nu_; gas[2 * N + o] = nvp;
            if (d.has_dust) { dust[N + o] = ndu; dust[2 * N + o] = ndvp; }
            if (do_rate) {
                const real cs = TAB(T_C0, ip);
                real sx = fabs(nu_) + cs, sy = fabs(nvp) + cs;
                if (d.has_dust) { sx = fmax(fabs(ndu), sx); sy = fmax(fabs(ndvp), sy); }
                rate = fmax(sx * TAB(T_IDX, ip), sy * d.idy * iR);
            }
        }
    }
    if (do_rate) block_max_store(d.rate_part + w.y, rate);
}

// CFL rate of the current state (get_dt_lv1, timestep.cu:18-97).
template <typename real>
__global__ __launch_bounds__(256) void k_cfl(LaunchSet<real> ls, int par) {
    const int2 w = ls.wl_pt[blockIdx.x];
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int segs = (nx + 255) >> 8;
    const int j = w.y / segs;
    const int i = (w.y - j * segs) * 256 + threadIdx.x;
    const real* __restrict__ gas = d.gas[gsel(par)];
    const real* __restrict__ dust = d.dust[dsel(par)];
    real rate = real(0);
    if (i < nx) {
        const int ip = i + NPAD;
        const size_t o = (size_t)j * nx + i;
        const real cs = TAB(T_C0, ip);
        real sx = fabs(gas[N + o]) + cs, sy = fabs(gas[2 * N + o]) + cs;
        if (d.has_dust) { sx = fmax(fabs(dust[N + o]), sx); sy = fmax(fabs(dust[2 * N + o]), sy); }
        rate = fmax(sx * TAB(T_IDX, ip), sy * d.idy * TAB(T_IR, ip));
    }
    block_max_store(d.rate_part + w.y, rate);
}

// Max of the partials into the control block (host-driven global_dt).
template <typename real>
__global__ __launch_bounds__(1024) void k_rate_reduce(LaunchSet<real> ls) {
    const DiskArgs<real>& d = ls.disks[blockIdx.x];
    double m = reduce_partials(d.rate_part, d.rate_nblk);
    if (threadIdx.x == 0) d.ctl->rate_bits = (unsigned long long)__double_as_longlong(m);
}

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

// Step prologue, one block.  mode 0: t and dt were set by the host.
// mode 1 (device loop): finish the previous step (t += dt), take
// dt = min(CFL/rate, 1.1 dt_prev) clipped to end_time, record it.  Then the
// planet at t and t+dt and the orbital shift of every radius.
template <typename real>
__global__ __launch_bounds__(1024) void k_prologue(LaunchSet<real> ls, int mode, int do_orbadv) {
    const DiskArgs<real>& d = ls.disks[blockIdx.x];
    double* dts = d.dts;
    StepCtl* ctl = d.ctl;
    __shared__ double s_dt;
    double rate = mode == 1 ? reduce_partials(d.rate_part, d.rate_nblk) : 0.0;
    if (threadIdx.x == 0) {
        double t = ctl->t, dt = ctl->dt;
        if (mode == 1) {
            if (ctl->nstep > 0) t += dt;
            ctl->rate_bits = (unsigned long long)__double_as_longlong(rate);
            double dtn = d.CFL / rate;
            double cap = dt * 1.1;
            dtn = dtn < cap ? dtn : cap;
            if (t + dtn > ctl->end_time) dtn = ctl->end_time - t;
            dt = dtn;
            ctl->t = t;
            ctl->dt = dt;
            ctl->dt_prev = dt;
            if (dts && ctl->nstep < d.dts_cap) dts[ctl->nstep] = dt;
            ctl->nstep += 1;
        }
        Body b0 = ctl->planet0;
        orbit_solution(b0, t, d.ramp_time, d.planet_mass, (double)d.frame_omega);
        ctl->planet0 = b0;
        ctl->pk0 = planet_kernel_view(b0);
        Body b1 = ctl->planet1;
        orbit_solution(b1, t + dt, d.ramp_time, d.planet_mass, (double)d.frame_omega);
        ctl->planet1 = b1;
        ctl->pk1 = planet_kernel_view(b1);
        s_dt = dt;
    }
    __syncthreads();
    if (d.orbadv && do_orbadv) {
        const double dt = s_dt;
        for (int i = threadIdx.x; i < d.nx; i += blockDim.x) {
            // set_orbital_advection (orbit.cu:74-99), FP64
            double dy = d.dy_d * d.rot[d.nx + i];
            double rt = d.rot[i];
            int shift = (int)floor(rt * dt / dy + 0.5);
            double speed = shift * dy / dt;
            d.inc[i] = shift;
            d.res[i] = (real)(rt - speed);
        }
    }
}

// ------------------------------------------------------------- launchers
template <typename Kern>
static hipError_t set_lds(Kern k, size_t lds) {
    if (lds <= 65536) return hipSuccess;
    return hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <typename real, int TX>
static hipError_t launch_sweep_x(const LaunchSet<real>& ls, int par, bool dust, hipStream_t s) {
    const size_t S = TX + 2 * NPAD;
    if (dust) {
        size_t lds = 13 * S * sizeof(real);
        hipError_t e = set_lds(k_sweep_x_dust<real, TX>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_sweep_x_dust<real, TX>), dim3(ls.n_x), dim3(TX), lds, s, ls, par);
    } else {
        size_t lds = 18 * S * sizeof(real);
        hipError_t e = set_lds(k_sweep_x_gas<real, TX>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_sweep_x_gas<real, TX>), dim3(ls.n_x), dim3(TX), lds, s, ls, par);
    }
    return hipGetLastError();
}

template <typename real>
static hipError_t sweep_x_any(const LaunchSet<real>& ls, int par, bool dust, hipStream_t s) {
    switch (ls.tile_x) {
        case 128: return launch_sweep_x<real, 128>(ls, par, dust, s);
        case 384: return launch_sweep_x<real, 384>(ls, par, dust, s);
        case 512: return launch_sweep_x<real, 512>(ls, par, dust, s);
        case 1024: return launch_sweep_x<real, 1024>(ls, par, dust, s);
        default: return launch_sweep_x<real, 256>(ls, par, dust, s);
    }
}

template <typename real, int B>
static hipError_t launch_ring(const LaunchSet<real>& ls, int par, int flags, bool dust, hipStream_t s) {
    size_t lds = ((size_t)3 * ls.ny_max + (size_t)14 * (B + 2 * NPAD)) * sizeof(real);
    if (dust) {
        hipError_t e = set_lds(k_ring_dust<real, B>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_ring_dust<real, B>), dim3(ls.n_ring), dim3(B), lds, s, ls, par, flags);
    } else {
        hipError_t e = set_lds(k_ring_gas<real, B>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_ring_gas<real, B>), dim3(ls.n_ring), dim3(B), lds, s, ls, par, flags);
    }
    return hipGetLastError();
}

template <typename real>
static hipError_t ring_any(const LaunchSet<real>& ls, int par, int flags, bool dust, hipStream_t s) {
    switch (ls.ring_b) {
        case 256: return launch_ring<real, 256>(ls, par, flags, dust, s);
        case 1024: return launch_ring<real, 1024>(ls, par, flags, dust, s);
        default: return launch_ring<real, 512>(ls, par, flags, dust, s);
    }
}

template <typename real>
struct Launch {
    static size_t ring_lds(int ny, int B) { return ((size_t)3 * ny + (size_t)14 * (B + 2 * NPAD)) * sizeof(real); }
    static hipError_t sweep_x_gas(const LaunchSet<real>& ls, int par, hipStream_t s) { return sweep_x_any(ls, par, false, s); }
    static hipError_t sweep_x_dust(const LaunchSet<real>& ls, int par, hipStream_t s) { return sweep_x_any(ls, par, true, s); }
    static hipError_t ring_gas(const LaunchSet<real>& ls, int par, int flags, hipStream_t s) { return ring_any(ls, par, flags, false, s); }
    static hipError_t ring_dust(const LaunchSet<real>& ls, int par, int flags, hipStream_t s) { return ring_any(ls, par, flags, true, s); }
    static hipError_t viscosity(const LaunchSet<real>& ls, int par, int which, hipStream_t s) {
        hipLaunchKernelGGL(k_viscosity<real>, dim3(ls.n_pt), dim3(256), 0, s, ls, par, which);
        return hipGetLastError();
    }
    static hipError_t sources(const LaunchSet<real>& ls, int par, int mode, int kill, int rate, hipStream_t s) {
        hipLaunchKernelGGL(k_sources<real>, dim3(ls.n_pt), dim3(256), 0, s, ls, par, mode, kill, rate);
        return hipGetLastError();
    }
    static hipError_t cfl(const LaunchSet<real>& ls, int par, hipStream_t s) {
        hipLaunchKernelGGL(k_cfl<real>, dim3(ls.n_pt), dim3(256), 0, s, ls, par);
        return hipGetLastError();
    }
    static hipError_t rate_reduce(const LaunchSet<real>& ls, hipStream_t s) {
        hipLaunchKernelGGL(k_rate_reduce<real>, dim3(ls.ndisk), dim3(1024), 0, s, ls);
        return hipGetLastError();
    }
    static hipError_t prologue(const LaunchSet<real>& ls, int mode, int do_orbadv, hipStream_t s) {
        hipLaunchKernelGGL(k_prologue<real>, dim3(ls.ndisk), dim3(1024), 0, s, ls, mode, do_orbadv);
        return hipGetLastError();
    }
};

template struct Launch<float>;
template struct Launch<double>;

}  // namespace pg
