// pg_device.h -- per-cell arithmetic of the Godunov step shared by the
// forward kernels (pg_kernels.hip) and their adjoints (pg_adjoint.hip).
//
// The pure functions are __host__ __device__ so that the host can evaluate
// them in FP64 (local derivative checks); FP32 uses the hardware
// approximations of FM<float> and exists on the device only.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>

#include "pg_internal.h"

#define PG_HD __host__ __device__ __forceinline__

namespace pg {

// A launch's parity word: bit 0 selects the current gas buffer of every disk
// in the set, bit 1 the current dust buffer; a kernel that writes a fluid
// writes the other buffer of its pair.
__device__ __forceinline__ int gsel(int par) { return par & 1; }
__device__ __forceinline__ int dsel(int par) { return (par >> 1) & 1; }


// Device-memory (global address space) view of a pointer held in a
// descriptor: loads through it compile to global loads instead of flat
// loads, which would also count against the LDS counter the sweeps wait on.
#define PG_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const PG_GLOBAL T* gmem(const T* p) { return (const PG_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ PG_GLOBAL T* gmem(T* p) { return (PG_GLOBAL T*)p; }

// Load at a uniform base plus an unsigned 32-bit byte offset: the form the
// scalar-base global load takes, so no 64-bit address arithmetic per load.
template <typename T>
__device__ __forceinline__ T ld_off(const PG_GLOBAL T* base, unsigned byte_off) {
    return *(const PG_GLOBAL T*)((const PG_GLOBAL char*)base + byte_off);
}

// per-radius tables, field-major: lanes that differ in radius read
// consecutive addresses
#define TAB(f, ip) ld_off(gmem(d.tab) + (size_t)(f) * (size_t)d.tab_stride, (unsigned)(ip) * (unsigned)sizeof(real))

// Division, reciprocal, square root, reciprocal square root and exponential.
// FP32: the hardware approximations (v_rcp_f32, v_sqrt_f32, v_rsq_f32,
// v_exp_f32; about 1 ulp); FP64: IEEE.
template <typename real> struct FM;
template <> struct FM<float> {
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    static __device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
    static __device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float exp(float x) { return __expf(x); }
};
template <> struct FM<double> {
    static PG_HD double rcp(double x) { return 1.0 / x; }
    static PG_HD double div(double a, double b) { return a / b; }
    static PG_HD double sqrt(double x) { return ::sqrt(x); }
    static PG_HD double rsqrt(double x) { return 1.0 / ::sqrt(x); }
    static PG_HD double exp(double x) { return ::exp(x); }
};

// ---------------------------------------------------------------- helpers
template <typename real>
PG_HD real lim01(real a) {  // solve.cu:14-17
    return fmin(fmax(a, real(0)), real(1));
}
template <typename real>
PG_HD real exp_lim(real x) {  // solve.cu:29-34
    return fmax(real(1) + x, real(1.0e-6));
}

// 4th-order slopes from precomputed geometric weights (the linear map
// get_4th_parameters applies to the four neighbouring differences,
// recon.cu:216-357), limiter (recon.cu:361-376), flattening and the PPM
// monotonicity correction (PPM.cu:64-91).
// slopes_d takes the four differences of neighbouring cell means around the
// cell (d1 = a0 - am1 across its left edge, d2 = ap1 - a0 across its right
// edge); slopes forms them from the five means.
template <typename real>
PG_HD void slopes_d(real d0, real d1, real d2, real d3,
                    real wl0, real wl1, real wl2, real wr0, real wr1, real wr2,
                    real flat, real& sL, real& sR) {
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
template <typename real>
PG_HD void slopes(real am2, real am1, real a0, real ap1, real ap2,
                                       real wl0, real wl1, real wl2, real wr0, real wr1, real wr2,
                                       real flat, real& sL, real& sR) {
    slopes_d(am1 - am2, a0 - am1, ap1 - a0, ap2 - ap1, wl0, wl1, wl2, wr0, wr1, wr2, flat, sL, sR);
}

// Mean of the cell parabola over its first fraction q (get_PPM_aveL,
// PPM.cu:106-113) and over its last fraction x (get_PPM_aveR with x = 1-q,
// PPM.cu:95-104).  The reference's form a - sL + q/2 (sR + sL + (1 - 2q/3) a6)
// with a6 = 3 (sL - sR) is the quadratic (a - sL) + q ((2 sL - sR) - q (sL - sR));
// written so, each evaluation is two FMAs on coefficients that the callers'
// repeated evaluations of one parabola share.
template <typename real>
PG_HD real aveL(real q, real sL, real a, real sR) {
    return fma(q, fma(-q, sL - sR, real(2) * sL - sR), a - sL);
}
template <typename real>
PG_HD real aveR(real x, real sL, real a, real sR) {
    return fma(x, fma(-x, sL - sR, sL - real(2) * sR), a + sR);
}

// Flattening coefficient of the centre cell of a 5-cell window
// (flatten, PPM.cu:7-62).  cs is the cell's sound speed.
// flatten_d takes the pressure differences dp1f = p[i+1] - p[i],
// dp2f = p[i+2] - p[i-1], dp1b = p[i] - p[i-1], dp2b = p[i+1] - p[i-2]
// and the cell's pressure p0.
template <typename real>
PG_HD real flatten_d(real dp1f, real dp2f, real dp1b, real dp2b, real p0, real um1, real u0, real up1) {
    real f = real(0);
    // Mf > 0 || Mb > 0 with Mf = (u0-up1)/cs, Mb = (um1-u0)/cs and cs > 0
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
template <typename real>
PG_HD real flatten5(real pm2, real pm1, real p0, real pp1, real pp2,
                                         real um1, real u0, real up1, real cs) {
    (void)cs;
    return flatten_d(pp1 - p0, pp2 - pm1, p0 - pm1, pp1 - pm2, p0, um1, u0, up1);
}

// Characteristic-traced half-step state on the right of an interface from
// the cell right of it (set_state, states.cu:17-75), locally isothermal
// (EOS_flag 0 fixes gam = 1, parameters.h:118-119, so exp_lim(gam*x) =
// exp_lim(x) and the sound speed of the traced state is sqrt(p0/r0)).
// c: the cell's sound speed; idx = 1/dx.  Returns the state and its sound
// speed cout (= the one HLLC recomputes from pl/rl, wave_speeds.cu:290).
template <typename real>
PG_HD void trace_right(real c, real dt, real idx, real r, real p, real u,
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
    real w = p0 * ir0;
    real ics = FM<real>::rsqrt(w);
    real cs = w * ics;   // sqrt(p0/r0) and its reciprocal from one rsqrt
    q = lim01((cs * dt - dis) * idx);
    real u_ = aveL(q, usL, u, usR);
    real p_ = aveL(q, psL, p, psR);
    real e = exp_lim((u0 - u_) * ics);
    rr = r0 * e;
    pr = p0 * e;
    ur = u0 + (p0 - p_) * ics * ir0;
    vr = aveL(q, vsL, v, vsR);
    cout = cs;
}

// ... and on the left from the cell left of it (states.cu:76-132).
template <typename real>
PG_HD void trace_left(real c, real dt, real idx, real r, real p, real u,
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
    real w = p0 * ir0;
    real ics = FM<real>::rsqrt(w);
    real cs = w * ics;
    x = lim01((dis + cs * dt) * idx);
    real u_ = aveR(x, usL, u, usR);
    real p_ = aveR(x, psL, p, psR);
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
PG_HD void trace_right_c(real c, real ic, real dt, real idx, real r, real u,
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
PG_HD void trace_left_c(real c, real ic, real dt, real idx, real r, real u,
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

// HLLC selection (HLLC.cu:2-77, EOS_flag 0) given the wave speeds: left
// supersonic (sl >= 0) -> F(U_L); right supersonic (sr < 0) -> F(U_R);
// otherwise the star state on the side of the contact (sm >= 0: left).
// Written without branches: the side is selected first, then one
// reciprocal serves the star state of that side.
template <typename real>
PG_HD void hllc_select(real pm, real sl, real sm, real sr,
                       real rl, real pl, real ul, real vl,
                       real rr, real pr, real ur, real vr,
                       real& Fr, real& us, real& vs, real& pres) {
    const bool sup_l = !__builtin_signbit(sl);
    const bool sup_r = !sup_l && __builtin_signbit(sr);
    const bool left = sup_l || (!sup_r && !__builtin_signbit(sm));
    const real r_ = left ? rl : rr, p_ = left ? pl : pr, u_ = left ? ul : ur, v_ = left ? vl : vr;
    const real s_ = left ? sl : sr;
    const real k = FM<real>::rcp(s_ - sm);
    const real Fs = sm * r_ * (s_ - u_) * k;
    const real ps = (s_ * pm - sm * p_) * k;
    const bool sup = sup_l || sup_r;
    Fr = sup ? u_ * r_ : Fs;
    pres = sup ? p_ : ps;
    us = u_ * Fr;
    vs = v_ * Fr;
}

// Wave speeds (get_pm_simple, get_lr_speeds_Einfeldt, get_sm;
// wave_speeds.cu:24-35,288-333; gam = 1) with the states' sound speeds cl,
// cr, then the HLLC fluxes: mass flux, upwinded normal and transverse
// momentum flux, interface pressure.
template <typename real>
PG_HD void hllc_iso(real cl, real cr, real rl, real pl, real ul, real vl,
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
    // sm = num/den + 2 (pl - pr) ics / (rl + rr), both divisions from one
    // reciprocal of den (rl + rr)
    const real den = rl * cl + rr * cr, rs = rl + rr;
    const real X = FM<real>::rcp(den * rs);
    real sm = (rl * ul * cl + rr * ur * cr) * rs * X + real(2) * (pl - pr) * ics * den * X;
    hllc_select(pm, sl, sm, sr, rl, pl, ul, vl, rr, pr, ur, vr, Fr, us, vs, pres);
}

// The same with one sound speed c on both sides (azimuthal isothermal sweep):
// cbar^2 = c^2 + n2 du^2 and sm = (rl (ul + c) + rr (ur - c)) / (rl + rr).
template <typename real>
PG_HD void hllc_c(real c, real ic, real c2, real rl, real ul, real vl, real rr, real ur, real vr,
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

// ------------------------------------------------- deviation form (gas)
// The gas density is held as its deviation from the time-independent radial
// background rho_b(R) (the unperturbed initial disk) and the pressure as
// p_b(R) + c_s^2(R) times that deviation, p_b = c_s^2 rho_b.  Differences of
// neighbouring values (the slopes, the traced states' averages, the
// interface pressures whose differences are the pressure force) are then
// formed from deviations, whose perturbation part keeps the full precision
// of the storage type, plus background differences tabulated in FP64.  In
// real arithmetic each function equals its counterpart above with
// value = background + deviation; only the rounding differs.
//
// exp_lim(x) - 1 without forming 1 + x (exp_lim, solve.cu:29-34)
template <typename real>
PG_HD real exp_lim_m1(real x) {
    return fmax(x, real(1.0e-6) - real(1));
}

// The stored deviation of a cell set to the density floor rmin over the
// background rb: rmin - rb, but never so close to -rb that rb + the
// deviation rounds to zero in the storage precision (FP32 resolves a
// deviation to ~6e-8 rb, so there a floor below ~5e-7 rb is held at
// ~5e-7 rb; FP64 keeps the reference's floor).
template <typename real>
PG_HD real floor_dev(real rmin, real rb) {
    return fmax(rmin - rb, -rb * (real(1) - real(4) * std::numeric_limits<real>::epsilon()));
}

// trace_right with the cell's background rb, pb and the deviations rd, pd
// of its means; the slopes are of the full fields.  Returns the traced
// density rr itself (positive: it is r0 >= the floor times exp_lim > 0) and
// the traced state's deviations rrd, prd from the same background.
template <typename real>
PG_HD void trace_right_d(real c, real dt, real idx, real rb, real pb, real rd, real pd, real u,
                         real rsL, real rsR, real psL, real psR, real usL, real usR,
                         real vsL, real v, real vsR,
                         real& rr, real& rrd, real& prd, real& ur, real& vr, real& cout) {
    real q = lim01((c - fmin(u, real(0))) * dt * idx);
    real ua = aveL(q, usL, u, usR);
    real dis = fmin(ua, real(0)) * dt;
    q = lim01(-dis * idx);
    real r0d = aveL(q, rsL, rd, rsR), r0 = rb + r0d;
    real p0d = aveL(q, psL, pd, psR), p0 = pb + p0d;
    const real rmin = (rb + rd) * real(1.0e-10), pmin = (pb + pd) * real(1.0e-10);
    if (r0 < rmin) { r0 = rmin; r0d = rmin - rb; }
    if (p0 < pmin) { p0 = pmin; p0d = pmin - pb; }
    real u0 = aveL(q, usL, u, usR);
    // y = 1/sqrt(p0 r0): the sound speed sqrt(p0/r0) = p0 y, its reciprocal
    // r0 y and 1/(r0 cs) = y, from one reciprocal square root
    const real y = FM<real>::rsqrt(p0 * r0);
    const real ics = r0 * y;
    const real cs = p0 * y;
    q = lim01((cs * dt - dis) * idx);
    real u_ = aveL(q, usL, u, usR);
    real p_d = aveL(q, psL, pd, psR);
    real xe = exp_lim_m1((u0 - u_) * ics);
    rr = fma(r0, xe, r0);
    rrd = fma(r0, xe, r0d);
    prd = fma(p0, xe, p0d);
    ur = u0 + (p0d - p_d) * y;
    vr = aveL(q, vsL, v, vsR);
    cout = cs;
}

template <typename real>
PG_HD void trace_left_d(real c, real dt, real idx, real rb, real pb, real rd, real pd, real u,
                        real rsL, real rsR, real psL, real psR, real usL, real usR,
                        real vsL, real v, real vsR,
                        real& rl, real& rld, real& pld, real& ul, real& vl, real& cout) {
    real x = lim01((fmax(u, real(0)) + c) * dt * idx);
    real ua = aveR(x, usL, u, usR);
    real dis = fmax(ua, real(0)) * dt;
    x = lim01(dis * idx);
    real r0d = aveR(x, rsL, rd, rsR), r0 = rb + r0d;
    real p0d = aveR(x, psL, pd, psR), p0 = pb + p0d;
    const real rmin = (rb + rd) * real(1.0e-10), pmin = (pb + pd) * real(1.0e-10);
    if (r0 < rmin) { r0 = rmin; r0d = rmin - rb; }
    if (p0 < pmin) { p0 = pmin; p0d = pmin - pb; }
    real u0 = aveR(x, usL, u, usR);
    const real y = FM<real>::rsqrt(p0 * r0);
    const real ics = r0 * y;
    const real cs = p0 * y;
    x = lim01((dis + cs * dt) * idx);
    real u_ = aveR(x, usL, u, usR);
    real p_d = aveR(x, psL, pd, psR);
    real xe = exp_lim_m1((u_ - u0) * ics);
    rl = fma(r0, xe, r0);
    rld = fma(r0, xe, r0d);
    pld = fma(p0, xe, p0d);
    ul = u0 + (p_d - p0d) * y;
    vl = aveR(x, vsL, v, vsR);
    cout = cs;
}

// hllc_iso for the traced states of the cells left (pressure background pbl)
// and right of an interface: densities rl, rr, pressures as deviations pld,
// prd from their cells' backgrounds.  tbp is the right background pressure
// minus pbl.  The interface pressure is returned as its deviation presd
// from the interface reference pbl + tbp/2 (the mean of the two
// backgrounds).
template <typename real>
PG_HD void hllc_iso_d(real cl, real cr, real pbl, real tbp, real rl, real pld, real ul, real vl,
                      real rr, real prd, real ur, real vr,
                      real& Fr, real& us, real& vs, real& presd) {
    const real h = real(0.5) * tbp;
    const real pref = pbl + h;
    const real pl = pld - h, pr = prd + h;    // deviations from pref
    // Roe averages through a = sqrt(rr/rl) (one reciprocal square root):
    // weights 1/(1 + a) and a/(1 + a), n2 = a/(2 (1 + a)^2).  The three
    // reciprocals 1/(cl + cr), 1/(1 + a) and 1/(den rs) come from the
    // reciprocal of their product (all three factors are positive).
    const real a = rr * FM<real>::rsqrt(rl * rr);
    const real csum = cl + cr, opa = real(1) + a;
    const real den = rl * cl + rr * cr, rs = rl + rr;
    const real dr = den * rs, co = csum * opa;
    const real Z = FM<real>::rcp(co * dr);
    const real ics = opa * dr * Z, iss = csum * dr * Z, X = co * Z;
    const real xm = exp_lim_m1((ul - ur) * ics);
    const real pm = fma(real(0.5) * (pr + pl), real(1) + xm, pref * xm);
    real ubar = (ul + a * ur) * iss;
    real n2 = real(0.5) * a * iss * iss;
    real du = ur - ul;
    real cbar = FM<real>::sqrt((cl * cl + a * (cr * cr)) * iss + n2 * du * du);
    real sl = ubar - cbar, sr = ubar + cbar;
    real sm = (rl * ul * cl + rr * ur * cr) * rs * X + real(2) * (pl - pr) * ics * den * X;
    // the star pressure (s pm - sm p)/(s - sm) is affine in (pm, p) with
    // weights summing to one, so it carries the common reference unchanged
    hllc_select(pm, sl, sm, sr, rl, pl, ul, vl, rr, pr, ur, vr, Fr, us, vs, presd);
}

// trace_right_c / trace_left_c with the ring's background density rb and
// the deviation rd of the cell's mean; returns the traced density (positive)
// and its deviation.  dti = dt idx, constant along the ring: the fractions
// (c - min(u, 0)) dt idx, -dis idx and (c dt - dis) idx with dis = w dt,
// w = min(ua, 0), are (c - min(u, 0)) dti, -w dti and (c - w) dti.
template <typename real>
PG_HD void trace_right_cd(real c, real ic, real dti, real rb, real rd, real u,
                          real rsL, real rsR, real usL, real usR, real vsL, real v, real vsR,
                          real& rr, real& rrd, real& ur, real& vr) {
    real q = lim01((c - fmin(u, real(0))) * dti);
    real ua = aveL(q, usL, u, usR);
    const real w = fmin(ua, real(0));
    q = lim01(-w * dti);
    real r0d = aveL(q, rsL, rd, rsR), r0 = rb + r0d;
    const real rmin = (rb + rd) * real(1.0e-10);
    if (r0 < rmin) { r0 = rmin; r0d = rmin - rb; }
    real u0 = aveL(q, usL, u, usR);
    q = lim01((c - w) * dti);
    real u_ = aveL(q, usL, u, usR);
    real rad = aveL(q, rsL, rd, rsR);
    const real xe = exp_lim_m1((u0 - u_) * ic);
    rr = fma(r0, xe, r0);
    rrd = fma(r0, xe, r0d);
    ur = u0 + c * FM<real>::div(r0d - rad, r0);
    vr = aveL(q, vsL, v, vsR);
}

template <typename real>
PG_HD void trace_left_cd(real c, real ic, real dti, real rb, real rd, real u,
                         real rsL, real rsR, real usL, real usR, real vsL, real v, real vsR,
                         real& rl, real& rld, real& ul, real& vl) {
    real x = lim01((fmax(u, real(0)) + c) * dti);
    real ua = aveR(x, usL, u, usR);
    const real w = fmax(ua, real(0));   // dis = w dt
    x = lim01(w * dti);
    real r0d = aveR(x, rsL, rd, rsR), r0 = rb + r0d;
    const real rmin = (rb + rd) * real(1.0e-10);
    if (r0 < rmin) { r0 = rmin; r0d = rmin - rb; }
    real u0 = aveR(x, usL, u, usR);
    x = lim01((w + c) * dti);
    real u_ = aveR(x, usL, u, usR);
    real rad = aveR(x, rsL, rd, rsR);
    const real xe = exp_lim_m1((u_ - u0) * ic);
    rl = fma(r0, xe, r0);
    rld = fma(r0, xe, r0d);
    ul = u0 + c * FM<real>::div(rad - r0d, r0);
    vl = aveR(x, vsL, v, vsR);
}

// hllc_c with the ring's background rb and the traced densities rl, rr
// and their deviations; the interface pressure is returned as its
// deviation from c^2 rb.
template <typename real>
PG_HD void hllc_cd(real c, real ic, real c2, real rb, real rl, real rld, real ul, real vl, real rr, real rrd,
                   real ur, real vr, real& Fr, real& us, real& vs, real& presd) {
    const real pl = c2 * rld, pr = c2 * rrd;
    const real xm = exp_lim_m1(real(0.5) * (ul - ur) * ic);
    const real pm = c2 * fma(real(0.5) * (rld + rrd), real(1) + xm, rb * xm);
    // Roe averages through a = sqrt(rr/rl); 1/(1 + a) and 1/(rl + rr) from
    // the reciprocal of their product
    const real a = rr * FM<real>::rsqrt(rl * rr);
    const real opa = real(1) + a, rs = rl + rr;
    const real Z = FM<real>::rcp(opa * rs);
    const real iss = rs * Z;
    real ubar = (ul + a * ur) * iss;
    real n2 = real(0.5) * a * iss * iss;
    real du = ur - ul;
    real cbar = FM<real>::sqrt(c2 + n2 * du * du);
    real sl = ubar - cbar, sr = ubar + cbar;
    // sm = (rl (ul + c) + rr (ur - c)) / (rl + rr) with c (rl - rr) formed
    // from the deviations: the terms of order c rho cancel to order u rho,
    // and in FP32 their rounding would be ~1e-3 of the contact speed
    real sm = (rl * ul + rr * ur + c * (rld - rrd)) * (opa * Z);
    hllc_select(pm, sl, sm, sr, rl, pl, ul, vl, rr, pr, ur, vr, Fr, us, vs, presd);
}

// Pressureless dust flux through an interface (dust_flux, dust.cu:6-105)
// from the parabolas of the cells left (L) and right (R) of it.
// idxL, idxR: 1/dx of those cells; dt: the sweep's time step in the
// coordinate's units.
template <typename real>
PG_HD void dust_iface(real dt,
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

__device__ __forceinline__ int wrap(int j, int n) {
    j %= n;
    return j < 0 ? j + n : j;
}


// ------------------------------------------------------- planet force
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
    const real iRp = FM<real>::rsqrt(Rp2);
    const real k = iRp * iRp * iRp;
    const real ind = (real)pk.m_over_x2;
    ax = -m * (R - px * c) * k - c * ind;
    ay = -m * px * s * k + s * ind;
}

}  // namespace pg
