// pg_adjoint_fn.h -- reverse-mode derivatives of the per-cell arithmetic of
// pg_device.h.
//
// Each *_adj function takes the forward inputs of its forward counterpart
// (it recomputes the forward values it needs) and the adjoints of the
// forward outputs, and ADDS the adjoints of the inputs into its reference
// arguments.  Branches (limiters, floors, min/max selections, the HLLC
// upwind choice) are taken as the forward pass takes them, so the result is
// the exact derivative of the forward arithmetic wherever it is
// differentiable.  Sound speeds, 1/dx, dt and geometry are constants.
#pragma once

#include "pg_device.h"

namespace pg {

template <typename real>
PG_HD real lim01_d(real a) { return (a > real(0) && a < real(1)) ? real(1) : real(0); }
template <typename real>
PG_HD real exp_lim_d(real x) { return (real(1) + x > real(1.0e-6)) ? real(1) : real(0); }
// +1 / -1 from the sign bit (copysign and fabs follow the sign bit)
template <typename real>
PG_HD real sbit(real x) { return __builtin_signbit(x) ? real(-1) : real(1); }

// aveL (PPM.cu:106-113): y = (a - sL) + q ((2 sL - sR) - q (sL - sR))
template <typename real>
PG_HD void aveL_adj(real q, real sL, real a, real sR, real ly, real& lq, real& lsL, real& la, real& lsR) {
    la += ly;
    lsL += ly * (real(-1) + q * (real(2) - q));
    lsR += ly * (q * (q - real(1)));
    lq += ly * ((real(2) * sL - sR) - real(2) * q * (sL - sR));
}
// aveR (PPM.cu:95-104): y = (a + sR) + x ((sL - 2 sR) - x (sL - sR))
template <typename real>
PG_HD void aveR_adj(real x, real sL, real a, real sR, real ly, real& lx, real& lsL, real& la, real& lsR) {
    la += ly;
    lsL += ly * (x * (real(1) - x));
    lsR += ly * (real(1) + x * (x - real(2)));
    lx += ly * ((sL - real(2) * sR) - real(2) * x * (sL - sR));
}

// slopes(): adjoints of the four neighbour differences d0..d3
// (d0 = am1-am2, d1 = a0-am1, d2 = ap1-a0, d3 = ap2-ap1) and of flat.
template <typename real>
PG_HD void slopes_adj(real am2, real am1, real a0, real ap1, real ap2,
                      real wl0, real wl1, real wl2, real wr0, real wr1, real wr2, real flat,
                      real lsL, real lsR, real ld[4], real& lflat) {
    const real d0 = am1 - am2, d1 = a0 - am1, d2 = ap1 - a0, d3 = ap2 - ap1;
    const real l0 = wl0 * d0 + wl1 * d1 + wl2 * d2;
    const real r0 = wr0 * d1 + wr1 * d2 + wr2 * d3;
    const bool usel = fabs(l0) < fabs(d1);
    const bool user = fabs(r0) < fabs(d2);
    const real l1 = copysign(fmin(fabs(l0), fabs(d1)), d1);
    const real r1 = copysign(fmin(fabs(r0), fabs(d2)), d2);
    const bool zero = r1 * l1 <= real(0);
    const real l2 = zero ? real(0) : l1, r2 = zero ? real(0) : r1;
    const real fk = real(1) - flat;
    const real l3 = l2 * fk, r3 = r2 * fk;
    const real t3 = r3 + l3;
    const real t2 = t3 * t3;
    const real t1 = t3 * real(3) * (l3 - r3);
    real ll3, lr3;
    if (t2 < t1) {          // sL = 2 r3, sR = r3
        ll3 = real(0);
        lr3 = real(2) * lsL + lsR;
    } else if (t2 < -t1) {  // sL = l3, sR = 2 l3
        ll3 = lsL + real(2) * lsR;
        lr3 = real(0);
    } else {
        ll3 = lsL;
        lr3 = lsR;
    }
    lflat -= ll3 * l2 + lr3 * r2;
    real ll1 = zero ? real(0) : ll3 * fk;
    real lr1 = zero ? real(0) : lr3 * fk;
    real ll0 = real(0), lr0 = real(0);
    real ld0 = real(0), ld1 = real(0), ld2 = real(0), ld3 = real(0);
    if (usel) ll0 = ll1 * sbit(l0) * sbit(d1);
    else ld1 += ll1;
    if (user) lr0 = lr1 * sbit(r0) * sbit(d2);
    else ld2 += lr1;
    ld0 += wl0 * ll0;
    ld1 += wl1 * ll0 + wr0 * lr0;
    ld2 += wl2 * ll0 + wr1 * lr0;
    ld3 += wr2 * lr0;
    ld[0] += ld0; ld[1] += ld1; ld[2] += ld2; ld[3] += ld3;
}

// slopes_adj with the differences pushed to the five cells
template <typename real>
PG_HD void slopes_adj5(const real a[5], real wl0, real wl1, real wl2, real wr0, real wr1, real wr2, real flat,
                       real lsL, real lsR, real la[5], real& lflat) {
    real ld[4] = {0, 0, 0, 0};
    slopes_adj(a[0], a[1], a[2], a[3], a[4], wl0, wl1, wl2, wr0, wr1, wr2, flat, lsL, lsR, ld, lflat);
    la[0] -= ld[0];
    la[1] += ld[0] - ld[1];
    la[2] += ld[1] - ld[2];
    la[3] += ld[2] - ld[3];
    la[4] += ld[3];
}

// flatten5 (PPM.cu:7-62): adjoints of the five pressures (the velocities
// only select the branch)
template <typename real>
PG_HD void flatten5_adj(real pm2, real pm1, real p0, real pp1, real pp2, real um1, real u0, real up1,
                        real lf, real lp[5]) {
    const real dp1f = pp1 - p0, dp2f = pp2 - pm1, dp1b = p0 - pm1, dp2b = pp1 - pm2;
    if (!(u0 - up1 > real(0) || um1 - u0 > real(0))) return;
    if (!(dp2f != real(0) && dp2b != real(0))) return;
    const real Rf = FM<real>::div(dp1f, dp2f), Rb = FM<real>::div(dp1b, dp2b);
    const real ef = fabs(Rf - real(0.5)) - real(0.25), eb = fabs(Rb - real(0.5)) - real(0.25);
    const real ff = real(10) * fmax(real(0), ef), fb = real(10) * fmax(real(0), eb);
    const real g = fmax(ff, fb);
    const real M = fmax(fabs(dp1f), fabs(dp1b));
    const real ip0 = FM<real>::rcp(p0);
    const real h = g * M * ip0;
    if (!(h < real(0.5))) return;   // f = 0.5
    const real lh = lf;
    const real lg = lh * M * ip0;
    const real lM = lh * g * ip0;
    real lp0 = -lh * h * ip0;
    real ldp1f = 0, ldp2f = 0, ldp1b = 0, ldp2b = 0;
    if (fabs(dp1f) >= fabs(dp1b)) ldp1f += lM * sbit(dp1f);
    else ldp1b += lM * sbit(dp1b);
    real lff = 0, lfb = 0;
    if (ff >= fb) lff = lg;
    else lfb = lg;
    if (ef > real(0)) {
        const real lR = real(10) * lff * sbit(Rf - real(0.5));
        const real i2 = FM<real>::rcp(dp2f);
        ldp1f += lR * i2;
        ldp2f -= lR * Rf * i2;
    }
    if (eb > real(0)) {
        const real lR = real(10) * lfb * sbit(Rb - real(0.5));
        const real i2 = FM<real>::rcp(dp2b);
        ldp1b += lR * i2;
        ldp2b -= lR * Rb * i2;
    }
    lp[0] += -ldp2b;
    lp[1] += -ldp2f - ldp1b;
    lp[2] += lp0 - ldp1f + ldp1b;
    lp[3] += ldp1f + ldp2b;
    lp[4] += ldp2f;
}

// Adjoint of one traced state.  Inputs (index): 0 r, 1 p, 2 u, 3 rsL, 4 rsR,
// 5 psL, 6 psR, 7 usL, 8 usR, 9 vsL, 10 v, 11 vsR.  Outputs: rr, pr, ur,
// vr, cout.
enum { TI_R, TI_P, TI_U, TI_RSL, TI_RSR, TI_PSL, TI_PSR, TI_USL, TI_USR, TI_VSL, TI_V, TI_VSR, TI_N };

// trace_right (states.cu:17-75)
template <typename real>
PG_HD void trace_right_adj(real c, real dt, real idx, const real* in, real lrr, real lpr, real lur, real lvr,
                           real lcout, real* lin) {
    const real r = in[TI_R], p = in[TI_P], u = in[TI_U];
    const real rsL = in[TI_RSL], rsR = in[TI_RSR], psL = in[TI_PSL], psR = in[TI_PSR];
    const real usL = in[TI_USL], usR = in[TI_USR], vsL = in[TI_VSL], v = in[TI_V], vsR = in[TI_VSR];
    // forward
    const real z1 = (c - fmin(u, real(0))) * dt * idx;
    const real q1 = lim01(z1);
    const real ua = aveL(q1, usL, u, usR);
    const real dis = fmin(ua, real(0)) * dt;
    const real z2 = -dis * idx;
    const real q2 = lim01(z2);
    const real r0a = aveL(q2, rsL, r, rsR), p0a = aveL(q2, psL, p, psR);
    const bool r0sel = !(r0a < r * real(1.0e-10));
    const bool p0sel = !(p0a < p * real(1.0e-10));
    const real r0 = fmax(r0a, r * real(1.0e-10));
    const real p0 = fmax(p0a, p * real(1.0e-10));
    const real u0 = aveL(q2, usL, u, usR);
    const real ir0 = FM<real>::rcp(r0);
    const real w0 = p0 * ir0;
    const real ics = FM<real>::rsqrt(w0);
    const real cs = w0 * ics;
    const real z3 = (cs * dt - dis) * idx;
    const real q3 = lim01(z3);
    const real u_ = aveL(q3, usL, u, usR);
    const real p_ = aveL(q3, psL, p, psR);
    const real earg = (u0 - u_) * ics;
    const real e = exp_lim(earg);
    // reverse
    real lq3 = 0, lq2 = 0, lq1 = 0;
    real lcs = lcout;
    aveL_adj(q3, vsL, v, vsR, lvr, lq3, lin[TI_VSL], lin[TI_V], lin[TI_VSR]);
    real lu0 = lur, lp0 = lur * ics * ir0, lp_ = -lur * ics * ir0;
    real lics = lur * (p0 - p_) * ir0, lir0 = lur * (p0 - p_) * ics;
    lp0 += lpr * e;
    real le = lpr * p0 + lrr * r0;
    real lr0 = lrr * e;
    const real learg = le * exp_lim_d(earg);
    lu0 += learg * ics;
    real lu_ = -learg * ics;
    lics += learg * (u0 - u_);
    lcs -= lics * ics * ics;
    aveL_adj(q3, psL, p, psR, lp_, lq3, lin[TI_PSL], lin[TI_P], lin[TI_PSR]);
    aveL_adj(q3, usL, u, usR, lu_, lq3, lin[TI_USL], lin[TI_U], lin[TI_USR]);
    const real lz3 = lq3 * lim01_d(z3);
    lcs += lz3 * dt * idx;
    real ldis = -lz3 * idx;
    const real lw = lcs * real(0.5) * ics;   // d sqrt(w) = dw / (2 cs)
    lp0 += lw * ir0;
    lir0 += lw * p0;
    lr0 -= lir0 * ir0 * ir0;
    aveL_adj(q2, usL, u, usR, lu0, lq2, lin[TI_USL], lin[TI_U], lin[TI_USR]);
    if (p0sel) aveL_adj(q2, psL, p, psR, lp0, lq2, lin[TI_PSL], lin[TI_P], lin[TI_PSR]);
    else lin[TI_P] += lp0 * real(1.0e-10);
    if (r0sel) aveL_adj(q2, rsL, r, rsR, lr0, lq2, lin[TI_RSL], lin[TI_R], lin[TI_RSR]);
    else lin[TI_R] += lr0 * real(1.0e-10);
    ldis -= lq2 * lim01_d(z2) * idx;
    const real lua = ua < real(0) ? ldis * dt : real(0);
    aveL_adj(q1, usL, u, usR, lua, lq1, lin[TI_USL], lin[TI_U], lin[TI_USR]);
    const real lz1 = lq1 * lim01_d(z1);
    if (u < real(0)) lin[TI_U] -= lz1 * dt * idx;
}

// trace_left (states.cu:76-132)
template <typename real>
PG_HD void trace_left_adj(real c, real dt, real idx, const real* in, real lrl, real lpl, real lul, real lvl,
                          real lcout, real* lin) {
    const real r = in[TI_R], p = in[TI_P], u = in[TI_U];
    const real rsL = in[TI_RSL], rsR = in[TI_RSR], psL = in[TI_PSL], psR = in[TI_PSR];
    const real usL = in[TI_USL], usR = in[TI_USR], vsL = in[TI_VSL], v = in[TI_V], vsR = in[TI_VSR];
    const real z1 = (fmax(u, real(0)) + c) * dt * idx;
    const real x1 = lim01(z1);
    const real ua = aveR(x1, usL, u, usR);
    const real dis = fmax(ua, real(0)) * dt;
    const real z2 = dis * idx;
    const real x2 = lim01(z2);
    const real r0a = aveR(x2, rsL, r, rsR), p0a = aveR(x2, psL, p, psR);
    const bool r0sel = !(r0a < r * real(1.0e-10));
    const bool p0sel = !(p0a < p * real(1.0e-10));
    const real r0 = fmax(r0a, r * real(1.0e-10));
    const real p0 = fmax(p0a, p * real(1.0e-10));
    const real u0 = aveR(x2, usL, u, usR);
    const real ir0 = FM<real>::rcp(r0);
    const real w0 = p0 * ir0;
    const real ics = FM<real>::rsqrt(w0);
    const real cs = w0 * ics;
    const real z3 = (dis + cs * dt) * idx;
    const real x3 = lim01(z3);
    const real u_ = aveR(x3, usL, u, usR);
    const real p_ = aveR(x3, psL, p, psR);
    const real earg = (u_ - u0) * ics;
    const real e = exp_lim(earg);
    real lx3 = 0, lx2 = 0, lx1 = 0;
    real lcs = lcout;
    aveR_adj(x3, vsL, v, vsR, lvl, lx3, lin[TI_VSL], lin[TI_V], lin[TI_VSR]);
    // ul = u0 + (p_ - p0) ics ir0
    real lu0 = lul, lp0 = -lul * ics * ir0, lp_ = lul * ics * ir0;
    real lics = lul * (p_ - p0) * ir0, lir0 = lul * (p_ - p0) * ics;
    lp0 += lpl * e;
    real le = lpl * p0 + lrl * r0;
    real lr0 = lrl * e;
    const real learg = le * exp_lim_d(earg);
    real lu_ = learg * ics;
    lu0 -= learg * ics;
    lics += learg * (u_ - u0);
    lcs -= lics * ics * ics;
    aveR_adj(x3, psL, p, psR, lp_, lx3, lin[TI_PSL], lin[TI_P], lin[TI_PSR]);
    aveR_adj(x3, usL, u, usR, lu_, lx3, lin[TI_USL], lin[TI_U], lin[TI_USR]);
    const real lz3 = lx3 * lim01_d(z3);
    real ldis = lz3 * idx;
    lcs += lz3 * dt * idx;
    const real lw = lcs * real(0.5) * ics;
    lp0 += lw * ir0;
    lir0 += lw * p0;
    lr0 -= lir0 * ir0 * ir0;
    aveR_adj(x2, usL, u, usR, lu0, lx2, lin[TI_USL], lin[TI_U], lin[TI_USR]);
    if (p0sel) aveR_adj(x2, psL, p, psR, lp0, lx2, lin[TI_PSL], lin[TI_P], lin[TI_PSR]);
    else lin[TI_P] += lp0 * real(1.0e-10);
    if (r0sel) aveR_adj(x2, rsL, r, rsR, lr0, lx2, lin[TI_RSL], lin[TI_R], lin[TI_RSR]);
    else lin[TI_R] += lr0 * real(1.0e-10);
    ldis += lx2 * lim01_d(z2) * idx;
    const real lua = ua > real(0) ? ldis * dt : real(0);
    aveR_adj(x1, usL, u, usR, lua, lx1, lin[TI_USL], lin[TI_U], lin[TI_USR]);
    const real lz1 = lx1 * lim01_d(z1);
    if (u > real(0)) lin[TI_U] += lz1 * dt * idx;
}

// Constant-sound-speed traces.  Inputs: 0 r, 1 u, 2 rsL, 3 rsR, 4 usL,
// 5 usR, 6 vsL, 7 v, 8 vsR.  Outputs: r, u, v of the state.
enum { TC_R, TC_U, TC_RSL, TC_RSR, TC_USL, TC_USR, TC_VSL, TC_V, TC_VSR, TC_N };

template <typename real>
PG_HD void trace_right_c_adj(real c, real ic, real dt, real idx, const real* in, real lrr, real lur, real lvr,
                             real* lin) {
    const real r = in[TC_R], u = in[TC_U], rsL = in[TC_RSL], rsR = in[TC_RSR];
    const real usL = in[TC_USL], usR = in[TC_USR], vsL = in[TC_VSL], v = in[TC_V], vsR = in[TC_VSR];
    const real z1 = (c - fmin(u, real(0))) * dt * idx;
    const real q1 = lim01(z1);
    const real ua = aveL(q1, usL, u, usR);
    const real dis = fmin(ua, real(0)) * dt;
    const real z2 = -dis * idx;
    const real q2 = lim01(z2);
    const real r0a = aveL(q2, rsL, r, rsR);
    const bool r0sel = !(r0a < r * real(1.0e-10));
    const real r0 = fmax(r0a, r * real(1.0e-10));
    const real u0 = aveL(q2, usL, u, usR);
    const real z3 = (c * dt - dis) * idx;
    const real q3 = lim01(z3);
    const real u_ = aveL(q3, usL, u, usR);
    const real ra = aveL(q3, rsL, r, rsR);
    const real earg = (u0 - u_) * ic;
    const real e = exp_lim(earg);
    const real ir0 = FM<real>::rcp(r0);
    real lq1 = 0, lq2 = 0, lq3 = 0;
    aveL_adj(q3, vsL, v, vsR, lvr, lq3, lin[TC_VSL], lin[TC_V], lin[TC_VSR]);
    // ur = u0 + c (r0 - ra) / r0 = u0 + c - c ra / r0
    real lu0 = lur;
    real lra = -lur * c * ir0;
    real lr0 = lur * c * ra * ir0 * ir0;
    // rr = r0 e
    lr0 += lrr * e;
    const real learg = lrr * r0 * exp_lim_d(earg);
    lu0 += learg * ic;
    const real lu_ = -learg * ic;
    aveL_adj(q3, rsL, r, rsR, lra, lq3, lin[TC_RSL], lin[TC_R], lin[TC_RSR]);
    aveL_adj(q3, usL, u, usR, lu_, lq3, lin[TC_USL], lin[TC_U], lin[TC_USR]);
    real ldis = -lq3 * lim01_d(z3) * idx;
    aveL_adj(q2, usL, u, usR, lu0, lq2, lin[TC_USL], lin[TC_U], lin[TC_USR]);
    if (r0sel) aveL_adj(q2, rsL, r, rsR, lr0, lq2, lin[TC_RSL], lin[TC_R], lin[TC_RSR]);
    else lin[TC_R] += lr0 * real(1.0e-10);
    ldis -= lq2 * lim01_d(z2) * idx;
    const real lua = ua < real(0) ? ldis * dt : real(0);
    aveL_adj(q1, usL, u, usR, lua, lq1, lin[TC_USL], lin[TC_U], lin[TC_USR]);
    const real lz1 = lq1 * lim01_d(z1);
    if (u < real(0)) lin[TC_U] -= lz1 * dt * idx;
}

template <typename real>
PG_HD void trace_left_c_adj(real c, real ic, real dt, real idx, const real* in, real lrl, real lul, real lvl,
                            real* lin) {
    const real r = in[TC_R], u = in[TC_U], rsL = in[TC_RSL], rsR = in[TC_RSR];
    const real usL = in[TC_USL], usR = in[TC_USR], vsL = in[TC_VSL], v = in[TC_V], vsR = in[TC_VSR];
    const real z1 = (fmax(u, real(0)) + c) * dt * idx;
    const real x1 = lim01(z1);
    const real ua = aveR(x1, usL, u, usR);
    const real dis = fmax(ua, real(0)) * dt;
    const real z2 = dis * idx;
    const real x2 = lim01(z2);
    const real r0a = aveR(x2, rsL, r, rsR);
    const bool r0sel = !(r0a < r * real(1.0e-10));
    const real r0 = fmax(r0a, r * real(1.0e-10));
    const real u0 = aveR(x2, usL, u, usR);
    const real z3 = (dis + c * dt) * idx;
    const real x3 = lim01(z3);
    const real u_ = aveR(x3, usL, u, usR);
    const real ra = aveR(x3, rsL, r, rsR);
    const real earg = (u_ - u0) * ic;
    const real e = exp_lim(earg);
    const real ir0 = FM<real>::rcp(r0);
    real lx1 = 0, lx2 = 0, lx3 = 0;
    aveR_adj(x3, vsL, v, vsR, lvl, lx3, lin[TC_VSL], lin[TC_V], lin[TC_VSR]);
    // ul = u0 + c (ra - r0) / r0 = u0 + c ra / r0 - c
    real lu0 = lul;
    real lra = lul * c * ir0;
    real lr0 = -lul * c * ra * ir0 * ir0;
    lr0 += lrl * e;
    const real learg = lrl * r0 * exp_lim_d(earg);
    const real lu_ = learg * ic;
    lu0 -= learg * ic;
    aveR_adj(x3, rsL, r, rsR, lra, lx3, lin[TC_RSL], lin[TC_R], lin[TC_RSR]);
    aveR_adj(x3, usL, u, usR, lu_, lx3, lin[TC_USL], lin[TC_U], lin[TC_USR]);
    real ldis = lx3 * lim01_d(z3) * idx;
    aveR_adj(x2, usL, u, usR, lu0, lx2, lin[TC_USL], lin[TC_U], lin[TC_USR]);
    if (r0sel) aveR_adj(x2, rsL, r, rsR, lr0, lx2, lin[TC_RSL], lin[TC_R], lin[TC_RSR]);
    else lin[TC_R] += lr0 * real(1.0e-10);
    ldis += lx2 * lim01_d(z2) * idx;
    const real lua = ua > real(0) ? ldis * dt : real(0);
    aveR_adj(x1, usL, u, usR, lua, lx1, lin[TC_USL], lin[TC_U], lin[TC_USR]);
    const real lz1 = lx1 * lim01_d(z1);
    if (u > real(0)) lin[TC_U] += lz1 * dt * idx;
}

// HLLC selection adjoint.  S = rl, pl, ul, vl, rr, pr, ur, vr; adds into
// lS[8], lpm, lsl, lsm, lsr.
template <typename real>
PG_HD void hllc_select_adj(real pm, real sl, real sm, real sr, const real* S, real lFr, real lus, real lvs,
                           real lpres, real* lS, real& lpm, real& lsl, real& lsm, real& lsr) {
    const real rl = S[0], pl = S[1], ul = S[2], vl = S[3], rr = S[4], pr = S[5], ur = S[6], vr = S[7];
    if (!__builtin_signbit(sl)) {
        const real Fr = ul * rl;
        const real lF = lFr + lus * ul + lvs * vl;
        lS[2] += lus * Fr + lF * rl;
        lS[3] += lvs * Fr;
        lS[0] += lF * ul;
        lS[1] += lpres;
    } else if (__builtin_signbit(sr)) {
        const real Fr = ur * rr;
        const real lF = lFr + lus * ur + lvs * vr;
        lS[6] += lus * Fr + lF * rr;
        lS[7] += lvs * Fr;
        lS[4] += lF * ur;
        lS[5] += lpres;
    } else if (!__builtin_signbit(sm)) {
        const real k = FM<real>::rcp(sl - sm);
        const real A = sm * rl * (sl - ul);
        const real Fr = A * k;
        const real lF = lFr + lus * ul + lvs * vl;
        lS[2] += lus * Fr;
        lS[3] += lvs * Fr;
        const real lA = lF * k;
        const real P = sl * pm - sm * pl;
        const real lk = lF * A + lpres * P;
        lsl += lpres * pm * k;
        lpm += lpres * sl * k;
        lsm -= lpres * pl * k;
        lS[1] -= lpres * sm * k;
        lsm += lA * rl * (sl - ul);
        lS[0] += lA * sm * (sl - ul);
        lsl += lA * sm * rl;
        lS[2] -= lA * sm * rl;
        lsl -= lk * k * k;
        lsm += lk * k * k;
    } else {
        const real k = FM<real>::rcp(sr - sm);
        const real A = sm * rr * (sr - ur);
        const real Fr = A * k;
        const real lF = lFr + lus * ur + lvs * vr;
        lS[6] += lus * Fr;
        lS[7] += lvs * Fr;
        const real lA = lF * k;
        const real P = sr * pm - sm * pr;
        const real lk = lF * A + lpres * P;
        lsr += lpres * pm * k;
        lpm += lpres * sr * k;
        lsm -= lpres * pr * k;
        lS[5] -= lpres * sm * k;
        lsm += lA * rr * (sr - ur);
        lS[4] += lA * sm * (sr - ur);
        lsr += lA * sm * rr;
        lS[6] -= lA * sm * rr;
        lsr -= lk * k * k;
        lsm += lk * k * k;
    }
}

// hllc_iso adjoint: S as above, lS[8] and lc[2] (cl, cr) accumulate.
template <typename real>
PG_HD void hllc_iso_adj(real cl, real cr, const real* S, real lFr, real lus, real lvs, real lpres, real* lS,
                        real& lcl, real& lcr) {
    const real rl = S[0], pl = S[1], ul = S[2], rr = S[4], pr = S[5], ur = S[6];
    const real csum = cl + cr;
    const real ics = FM<real>::rcp(csum);
    const real pmarg = (ul - ur) * ics;
    const real pme = exp_lim(pmarg);
    const real pm = real(0.5) * (pr + pl) * pme;
    const real srl = FM<real>::sqrt(rl), srr = FM<real>::sqrt(rr);
    const real iss = FM<real>::rcp(srl + srr);
    const real ubar = (srl * ul + srr * ur) * iss;
    const real n2 = real(0.5) * srl * srr * iss * iss;
    const real du = ur - ul;
    const real cq = cl * cl * srl + cr * cr * srr;
    const real cbar = FM<real>::sqrt(cq * iss + n2 * du * du);
    const real sl = ubar - cbar, sr = ubar + cbar;
    const real num = rl * ul * cl + rr * ur * cr, den = rl * cl + rr * cr;
    const real X = FM<real>::rcp(den * (rl + rr));
    const real irs = den * X;
    const real smA = num * (rl + rr) * X;
    const real sm = smA + real(2) * (pl - pr) * ics * den * X;
    real lpm = 0, lsl = 0, lsm = 0, lsr = 0;
    hllc_select_adj(pm, sl, sm, sr, S, lFr, lus, lvs, lpres, lS, lpm, lsl, lsm, lsr);
    const real lubar = lsl + lsr;
    const real lcbar = lsr - lsl;
    // sm
    const real iden = (rl + rr) * X;
    const real lnum = lsm * iden, lden = -lsm * smA * iden;
    lS[0] += lnum * ul * cl + lden * cl;
    lS[2] += lnum * rl * cl;
    lcl += lnum * rl * ul + lden * rl;
    lS[4] += lnum * ur * cr + lden * cr;
    lS[6] += lnum * rr * cr;
    lcr += lnum * rr * ur + lden * rr;
    lS[1] += lsm * real(2) * ics * irs;
    lS[5] -= lsm * real(2) * ics * irs;
    real lics = lsm * real(2) * (pl - pr) * irs;
    const real lirs = lsm * real(2) * (pl - pr) * ics;
    lS[0] -= lirs * irs * irs;
    lS[4] -= lirs * irs * irs;
    // cbar
    const real lcb2 = lcbar * real(0.5) * FM<real>::rcp(cbar);
    lcl += lcb2 * real(2) * cl * srl * iss;
    lcr += lcb2 * real(2) * cr * srr * iss;
    real lsrl = lcb2 * cl * cl * iss, lsrr = lcb2 * cr * cr * iss;
    real liss = lcb2 * cq;
    const real ln2 = lcb2 * du * du;
    const real ldu = lcb2 * real(2) * n2 * du;
    lS[6] += ldu;
    lS[2] -= ldu;
    lsrl += ln2 * real(0.5) * srr * iss * iss;
    lsrr += ln2 * real(0.5) * srl * iss * iss;
    liss += ln2 * srl * srr * iss;
    // ubar
    lsrl += lubar * ul * iss;
    lS[2] += lubar * srl * iss;
    lsrr += lubar * ur * iss;
    lS[6] += lubar * srr * iss;
    liss += lubar * (srl * ul + srr * ur);
    // pm
    lS[5] += lpm * real(0.5) * pme;
    lS[1] += lpm * real(0.5) * pme;
    const real lpmarg = lpm * real(0.5) * (pr + pl) * exp_lim_d(pmarg);
    lS[2] += lpmarg * ics;
    lS[6] -= lpmarg * ics;
    lics += lpmarg * (ul - ur);
    // iss, sqrt, ics
    lsrl -= liss * iss * iss;
    lsrr -= liss * iss * iss;
    lS[0] += lsrl * real(0.5) * FM<real>::rcp(srl);
    lS[4] += lsrr * real(0.5) * FM<real>::rcp(srr);
    const real lcsum = -lics * ics * ics;
    lcl += lcsum;
    lcr += lcsum;
}

// hllc_c adjoint: Sc = rl, ul, vl, rr, ur, vr; lSc[6] accumulates.
template <typename real>
PG_HD void hllc_c_adj(real c, real ic, real c2, const real* Sc, real lFr, real lus, real lvs, real lpres,
                      real* lSc) {
    const real rl = Sc[0], ul = Sc[1], vl = Sc[2], rr = Sc[3], ur = Sc[4], vr = Sc[5];
    const real pl = c2 * rl, pr = c2 * rr;
    const real pmarg = real(0.5) * (ul - ur) * ic;
    const real pme = exp_lim(pmarg);
    const real pm = real(0.5) * (pr + pl) * pme;
    const real srl = FM<real>::sqrt(rl), srr = FM<real>::sqrt(rr);
    const real iss = FM<real>::rcp(srl + srr);
    const real ubar = (srl * ul + srr * ur) * iss;
    const real n2 = real(0.5) * srl * srr * iss * iss;
    const real du = ur - ul;
    const real cbar = FM<real>::sqrt(c2 + n2 * du * du);
    const real sl = ubar - cbar, sr = ubar + cbar;
    const real num = rl * (ul + c) + rr * (ur - c), den = rl + rr;
    const real sm = FM<real>::div(num, den);
    const real S[8] = {rl, pl, ul, vl, rr, pr, ur, vr};
    real lS[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    real lpm = 0, lsl = 0, lsm = 0, lsr = 0;
    hllc_select_adj(pm, sl, sm, sr, S, lFr, lus, lvs, lpres, lS, lpm, lsl, lsm, lsr);
    real lrl = lS[0] + c2 * lS[1], lul = lS[2], lvl = lS[3];
    real lrr = lS[4] + c2 * lS[5], lur = lS[6], lvr = lS[7];
    const real lubar = lsl + lsr;
    const real lcbar = lsr - lsl;
    // sm = num / den
    const real iden = FM<real>::rcp(den);
    const real lnum = lsm * iden, lden = -lsm * sm * iden;
    lrl += lnum * (ul + c) + lden;
    lul += lnum * rl;
    lrr += lnum * (ur - c) + lden;
    lur += lnum * rr;
    // cbar = sqrt(c2 + n2 du^2)
    const real lcb2 = lcbar * real(0.5) * FM<real>::rcp(cbar);
    const real ln2 = lcb2 * du * du;
    const real ldu = lcb2 * real(2) * n2 * du;
    lur += ldu;
    lul -= ldu;
    real lsrl = ln2 * real(0.5) * srr * iss * iss;
    real lsrr = ln2 * real(0.5) * srl * iss * iss;
    real liss = ln2 * srl * srr * iss;
    lsrl += lubar * ul * iss;
    lul += lubar * srl * iss;
    lsrr += lubar * ur * iss;
    lur += lubar * srr * iss;
    liss += lubar * (srl * ul + srr * ur);
    // pm = (pr + pl)/2 exp_lim((ul - ur) ic / 2)
    lrr += lpm * real(0.5) * pme * c2;
    lrl += lpm * real(0.5) * pme * c2;
    const real lpmarg = lpm * real(0.5) * (pr + pl) * exp_lim_d(pmarg);
    lul += lpmarg * real(0.5) * ic;
    lur -= lpmarg * real(0.5) * ic;
    lsrl -= liss * iss * iss;
    lsrr -= liss * iss * iss;
    lrl += lsrl * real(0.5) * FM<real>::rcp(srl);
    lrr += lsrr * real(0.5) * FM<real>::rcp(srr);
    lSc[0] += lrl; lSc[1] += lul; lSc[2] += lvl;
    lSc[3] += lrr; lSc[4] += lur; lSc[5] += lvr;
}

}  // namespace pg
