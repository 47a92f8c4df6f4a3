// penguin_oracle.cpp -- FP64 host restatement of the reference solver's step.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Nothing in the product library
// links, loads or calls this file; tests/, the smoke check and bench.py's
// CPU-baseline leg load it through ctypes.
//
// What it restates, and from where (reference = fakesuji/PEnGUIn):
//   * grid and profiles        src/geom/geom.cu:3-33, src/init/init.cu:7-532
//   * reconstruction           src/solver/recon/recon.cu:216-501, PPM.cu:7-113
//   * traced states + Riemann  src/solver/riemann/states.cu:1-135,
//                              wave_speeds.cu:24-35,288-333, HLLC.cu:2-77,
//                              riemann.cu:5-131
//   * sweeps + update          src/solver/sweeps.cu:218-456
//   * dust                     src/solver/dust/dust.cu:1-353
//   * orbital advection        src/orbit/orbit.cu:17-157,
//                              src/solver/advection/advection.cu:1-357
//   * viscosity                src/viscosity/viscosity.cu:9-197,489-549
//   * source terms + forces    src/solver/sweeps.cu:555-747,
//                              src/solver/force/forces.cu:41-58,110-327
//   * CFL                      src/timestep/timestep.cu:18-140
//   * ghost fills              src/boundary/boundary.cu:13-268,462-848
//   * planet                   src/planet/planet.cu:7-187,293-317
//   * wave killing             src/killwave/killwave.cu:8-117
//   * split order              src/solver/solve.cu:212-292, main.cu:320-359
//
// The reference's block-cooperative device routines (flatten, riemann,
// net_source, net_flux: shared memory + __syncthreads) are executed here as
// ordered phases over a whole row, which is what the barriers enforce.  Rows
// are independent and run under OpenMP.
//
// Scope: ndim = 2 (geomx = 1 cylindrical radius, geomy = 3 azimuth),
// init_flag 2 disk, recon_flag 6 (PPM4), EOS_flag 0 (locally isothermal),
// twobd_flag 0, cool_flag 0, gas and dust; and ndim = 3 (geomx 2 spherical
// radius, geomy 4 azimuth, geomz 5 polar angle), init_flag 3 disk,
// recon_flag 5 (PEM4) or 6, EOS_flag 2 with internal_e_flag 1 (adiabatic,
// internal energy), gas only, no wave killing: the polar sweep
// (sweeps.cu:458-553), polar ghost fills (boundary.cu:270-321), spherical
// forces with the spline-softened planet (forces.cu:14-39, 110-327) and the
// per-(radius, polar layer) orbital advection (orbit.cu:51-99).
// Resolutions of the reference's ambiguities:
//   * dust sweeps and dust advection read the flattening coefficient from
//     uninitialised shared memory under recon_flag 6 (dust.cu:116-122,
//     advection.cu:186-192 vs PPM.cu:69-78); here it is 0, the convention the
//     gas advection uses (advection.cu:12-18);
//   * killwave's __sinf (killwave.cu:11) is sinf on a float argument, PEM's
//     __logf (PEM.cu:13,29) the correctly rounded float logarithm of a float
//     argument (CUDA's approximation is not reproduced).

#include "penguin_oracle.h"
#include "oracle_real.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const real_t kPi = 3.1415926535897932384626434;
const real_t kHalfPi = 0.5 * kPi;
const real_t kTwoPi = 2.0 * kPi;
const real_t kSmallR = 1.0e-10;
const real_t kSmallP = 1.0e-10;
const int kPad = 3;  // npad for recon_flag > 4 (parameters.h:49-53)

struct Cell { real_t r, p, u, v, w; };
struct Dust { real_t r, u, v, w; };
struct Body { real_t m, a, e, x, y, z, vx, vy, vz, rs; };

// ---------------------------------------------------------------- profiles
// init.cu:7-50, 72-349 (init_flag 2, geomx 1, ndim 2), 351-454.

real_t cs2_at(const pg_params& P, real_t x) {
    real_t coeff = P.sc_h * P.sc_h * (1.0 / P.planet_radius);
    return coeff * mpow(x / P.planet_radius, -P.p_beta);
}
real_t h_at(const pg_params& P, real_t x) { return msqrt(cs2_at(P, x) * mpow(x, 3)); }
real_t nu_at(const pg_params& P, real_t x) {
    if (!P.visc_flag) return 0.0;
    real_t rp = P.planet_radius;
    real_t coeff = P.ss_alpha * msqrt(cs2_at(P, rp)) * h_at(P, rp);
    return coeff * mpow(x / rp, P.p_alpha);
}
real_t rho_init(const pg_params& P, real_t x) { return mpow(x / P.planet_radius, -P.p_alpha); }
real_t p_init(const pg_params& P, real_t x) { return cs2_at(P, x) * rho_init(P, x); }
real_t u_init(const pg_params& P, real_t x) { return -1.5 * nu_at(P, x) / x; }
real_t v_init(const pg_params& P, real_t x) {
    real_t rho = rho_init(P, x);
    real_t dr = 1.1514178e-7;
    real_t dPdr = (p_init(P, x + dr) - p_init(P, x - dr)) / (2.0 * dr);
    return msqrt(1.0 / x + (x / rho) * dPdr);
}
Cell cell_init(const pg_params& P, real_t x) {
    return Cell{rho_init(P, x), p_init(P, x), u_init(P, x), v_init(P, x), 0.0};
}

// init_flag 3 (ndim 3, geomx 2): x is the spherical radius, z the polar
// angle, x sin z the cylindrical radius (init.cu:7-70, 72-349)
real_t cs2_3(const pg_params& P, real_t x, real_t z) {
    real_t coeff = P.sc_h * P.sc_h * (1.0 / P.planet_radius);
    return coeff * mpow(x * msin(z) / P.planet_radius, -P.p_beta);
}
real_t h_3(const pg_params& P, real_t x, real_t z) { return msqrt(cs2_3(P, x, z) * mpow(x * msin(z), 3)); }
real_t nu_3(const pg_params& P, real_t x, real_t z) {
    if (!P.visc_flag) return 0.0;
    real_t rp = P.planet_radius;
    real_t coeff = P.ss_alpha * msqrt(cs2_3(P, rp, z)) * h_3(P, rp, z);
    return coeff * mpow(x * msin(z) / rp, P.p_alpha);
}
real_t vertical_3(const pg_params& P, real_t x, real_t z) {  // get_vertical, geomx 2
    return mexp((1.0 / x - 1.0 / (x * msin(z))) / cs2_3(P, x, z));
}
real_t rho_3(const pg_params& P, real_t x, real_t z) {
    return mpow(x * msin(z) / P.planet_radius, -P.p_alpha) * vertical_3(P, x, z);
}
real_t p_3(const pg_params& P, real_t x, real_t z) { return cs2_3(P, x, z) * rho_3(P, x, z); }
real_t v_3(const pg_params& P, real_t x, real_t z) {
    real_t rho = rho_3(P, x, z);
    real_t dr = 1.1514178e-7;
    real_t dPdr = (p_3(P, x + dr, z) - p_3(P, x - dr, z)) / (2.0 * dr);
    return msqrt(1.0 / x + (x / rho) * dPdr);
}
Cell cell_init3(const pg_params& P, real_t x, real_t z) {
    return Cell{rho_3(P, x, z), p_3(P, x, z), -1.5 * nu_3(P, x, z) / (x * msin(z)), v_3(P, x, z), 0.0};
}
Dust dust_init(const pg_params& P, real_t x) {
    real_t St = P.Stokes;
    real_t f_gas = v_init(P, x) / msqrt(1.0 / x);
    real_t tmp = St * St * f_gas;
    real_t L = ((1.0 + tmp) / tmp) *
               (1.0 - msqrt(1.0 - 2.0 * tmp * (1.0 - f_gas) / (1.0 + tmp) / (1.0 + tmp)));
    real_t u0 = St * (L * L - 2.0 * L);
    real_t u = (1.0 - msqrt(1.0 - 2.0 * St * u0)) / St;
    L = (1.0 - f_gas + St * u / 2.0) / (1.0 + St * u / 2.0);
    u0 = St * (L * L - 2.0 * L);
    u = (1.0 - msqrt(1.0 - 2.0 * St * u0)) / St;
    Dust d;
    d.r = P.D_G_ratio * mpow(x / P.planet_radius, -P.p_alpha);
    d.u = u * msqrt(1.0 / x);
    d.v = msqrt(1.0 / x);
    d.w = 0.0;
    return d;
}

// ---------------------------------------------------------------- geometry
// geom.cu:3-33, init.cu:472-500

real_t dv_dr(int geom, real_t ra, real_t dr) {
    if (geom == 5) return dr < 1.0e-4 ? msin(ra + 0.5 * dr) : (mcos(ra) - mcos(ra + dr)) / dr;
    if (geom == 2) return ra * ra + ra * dr + dr * dr / 3.0;
    if (geom == 1) return ra + dr / 2.0;
    return 1.0;
}
real_t cell_volume(int geom, real_t ra, real_t dr) { return dr * dv_dr(geom, ra, dr); }

void build_axis(std::vector<real_t>& a, std::vector<real_t>& v, real_t amin, real_t amax,
                int res, int pad, int geom, int grid) {
    a.assign(res + 2 * pad + 1, 0.0);
    v.assign(res + 2 * pad, 0.0);
    if (grid == 1) {
        real_t t = mlog(amax / amin);
        for (int i = 0; i <= res; ++i) a[pad + i] = amin * mexp(t * (real_t)i / (real_t)res);
    } else {
        for (int i = 0; i <= res; ++i) a[pad + i] = amin + (amax - amin) * (real_t)i / (real_t)res;
    }
    for (int i = 0; i < pad; ++i) {
        if (grid == 1) {
            a[i] = amin * mexp(mlog(amax / amin) * (real_t)(i - pad) / (real_t)res);
            a[res + pad + 1 + i] = amin * mexp(mlog(amax / amin) * (real_t)(res + 1 + i) / (real_t)res);
        } else {
            a[i] = a[pad] - (a[2 * pad - i] - a[pad]);
            a[res + pad + 1 + i] = a[res + pad] + (a[res + pad] - a[res + pad - 1 - i]);
        }
    }
    for (int i = 0; i < res + 2 * pad; ++i) v[i] = cell_volume(geom, a[i], a[i + 1] - a[i]);
}

// ---------------------------------------------------------- reconstruction
inline real_t lim01(real_t a) { return mfmin(mfmax(a, 0.0), 1.0); }  // solve.cu:14-17
inline real_t exp_lim(real_t x) { return mfmax(1.0 + x, 1.0e-6); }       // solve.cu:29-34

// Cell means of x, x^2, x^3 under the geometric weight (1, x, x^2 for
// geom 0/3/4, 1, 2), recon.cu:242-301.
inline void cell_moments(int geom, real_t xl, real_t xr, real_t m[3]) {
    if (geom == 2) {
        real_t d = xr * xr * xr - xl * xl * xl;
        m[0] = (xr * xr * xr * xr - xl * xl * xl * xl) / d / (4.0 / 3.0);
        m[1] = (xr * xr * xr * xr * xr - xl * xl * xl * xl * xl) / d / (5.0 / 3.0);
        m[2] = (xr * xr * xr * xr * xr * xr - xl * xl * xl * xl * xl * xl) / d / 2.0;
    } else if (geom == 1) {
        real_t d = xr * xr - xl * xl;
        m[0] = (xr * xr * xr - xl * xl * xl) / d / 1.5;
        m[1] = (xr * xr * xr * xr - xl * xl * xl * xl) / d / 2.0;
        m[2] = (xr * xr * xr * xr * xr - xl * xl * xl * xl * xl) / d / 2.5;
    } else {
        real_t d = xr - xl;
        m[0] = (xr * xr - xl * xl) / d / 2.0;
        m[1] = (xr * xr * xr - xl * xl * xl) / d / 3.0;
        m[2] = (xr * xr * xr * xr - xl * xl * xl * xl) / d / 4.0;
    }
}

// Coefficients c1..c3 of the cubic whose cell means over four consecutive
// cells (moments M[0..3]) differ by d[0..2] (recon.cu:305-328).
inline void cubic_from_means(const real_t M[4][3], const real_t d[3], real_t c[3]) {
    real_t D0[3], D1[3], D2[3];
    for (int k = 0; k < 3; ++k) {
        D0[k] = M[1][k] - M[0][k];
        D1[k] = M[2][k] - M[1][k];
        D2[k] = M[3][k] - M[2][k];
    }
    // D?2 -> index 0, D?3 -> index 1, D?4 -> index 2
    real_t den_a = D2[1] * D0[0] - D0[1] * D2[0];
    real_t den_b = D1[1] * D0[0] - D0[1] * D1[0];
    real_t c3 = (D0[0] * d[2] - D2[0] * d[0]) / den_a - (d[1] * D0[0] - d[0] * D1[0]) / den_b;
    real_t t = (D0[0] * D2[2] - D2[0] * D0[2]) / den_a - (D1[2] * D0[0] - D0[2] * D1[0]) / den_b;
    c3 /= t;
    t = (d[1] * D0[0] - d[0] * D1[0]) / den_b;
    real_t c2 = t - c3 * (D1[2] * D0[0] - D0[2] * D1[0]) / den_b;
    t = d[0] / D0[0];
    real_t c1 = t - c2 * D0[1] / D0[0] - c3 * D0[2] / D0[0];
    c[0] = c1; c[1] = c2; c[2] = c3;
}

// get_4th_parameters (recon.cu:216-378): par = {sL, <a>, sR} of cell i, with
// x[i-2..i+3] the edges and a[i-2..i+2] the cell means.
void fourth_parameters(int i, int geom, const real_t* x, const real_t* a, real_t par[3]) {
    real_t M[5][3];
    for (int c = 0; c < 5; ++c) cell_moments(geom, x[i - 2 + c], x[i - 1 + c], M[c]);
    real_t a0 = a[i - 2], a1 = a[i - 1], a2 = a[i], a3 = a[i + 1], a4 = a[i + 2];
    real_t x2 = x[i], x3 = x[i + 1];
    real_t c[3];
    {
        real_t Mw[4][3];
        for (int m = 0; m < 4; ++m) for (int k = 0; k < 3; ++k) Mw[m][k] = M[m][k];
        real_t d[3] = {a1 - a0, a2 - a1, a3 - a2};
        cubic_from_means(Mw, d, c);
    }
    real_t sL = c[0] * (M[2][0] - x2) + c[1] * (M[2][1] - x2 * x2) + c[2] * (M[2][2] - x2 * x2 * x2);
    {
        real_t Mw[4][3];
        for (int m = 0; m < 4; ++m) for (int k = 0; k < 3; ++k) Mw[m][k] = M[m + 1][k];
        real_t d[3] = {a2 - a1, a3 - a2, a4 - a3};
        cubic_from_means(Mw, d, c);
    }
    real_t sR = c[0] * (x3 - M[2][0]) + c[1] * (x3 * x3 - M[2][1]) + c[2] * (x3 * x3 * x3 - M[2][2]);
    sL = mcopysign(mfmin(mfabs(sL), mfabs(a2 - a1)), a2 - a1);
    sR = mcopysign(mfmin(mfabs(sR), mfabs(a3 - a2)), a3 - a2);
    if (sR * sL <= 0.0) { par[0] = 0.0; par[1] = a2; par[2] = 0.0; }
    else                { par[0] = sL;  par[1] = a2; par[2] = sR; }
}

// adjust_PPM_parameters (PPM.cu:64-91)
void ppm_adjust(real_t flat, real_t par[3]) {
    real_t sL = par[0], sR = par[2];
    sR *= (1.0 - flat);
    sL *= (1.0 - flat);
    real_t t3 = sR + sL;
    real_t t2 = t3 * t3;
    real_t t1 = t3 * 3.0 * (sL - sR);
    if (t2 < t1) sL = 2.0 * sR;
    else if (t2 < -t1) sR = 2.0 * sL;
    par[0] = sL;
    par[2] = sR;
}

// get_PPM_aveR / aveL (PPM.cu:95-113)
inline real_t ppm_aveR(real_t x, const real_t* par) {
    real_t a6 = 3.0 * (par[0] - par[2]);
    real_t aR = par[1] + par[2];
    real_t S = par[2] + par[0];
    x = 1.0 - x;
    return aR - (x / 2.0) * (S - (1.0 - x / 1.5) * a6);
}
inline real_t ppm_aveL(real_t x, const real_t* par) {
    real_t a6 = 3.0 * (par[0] - par[2]);
    real_t aL = par[1] - par[0];
    real_t S = par[2] + par[0];
    return aL + (x / 2.0) * (S + (1.0 - x / 1.5) * a6);
}

// recon_flag 6: 4th-order parameters, then flattening + PPM monotonicity
inline void recon_par(int i, int geom, const real_t* x, const real_t* a, const real_t* flat, real_t par[3]) {
    fourth_parameters(i, geom, x, a, par);
    ppm_adjust(flat[i], par);
}

// get_PEM_0 / get_PEM_1 (PEM.cu:3-48): power-law sub-cell profiles; the
// logarithm is the single-precision one of a float argument (__logf,
// PEM.cu:13,29), here correctly rounded.  get_PEM_1's x >= 1 assignment is overwritten by the
// branch that follows it (PEM.cu:23-27) and so has no effect here either.
inline real_t pem0(real_t x, real_t S, real_t aB, real_t eta) {
    real_t v = x == 0.0 ? real_t(0.0) : mexp(eta * mlogf(x));
    return aB + S * v;
}
inline real_t pem1(real_t x, real_t S, real_t aB, real_t eta) {
    real_t v;
    if (eta * x > 0.001) {
        v = (1.0 - x) * (mexp(eta * mlogf(1.0 - x)) - 1.0) / (eta * x) + 1.0;
    } else {
        real_t t = (eta + 1.0) * x / 2.0;
        v = t;
        t *= -(eta - 1.0) * x / 3.0;
        v += t;
        t *= -(eta - 2.0) * x / 4.0;
        v += t;
        t *= -(eta - 3.0) * x / 5.0;
        v += t;
    }
    return aB + S * v;
}
// get_PEM_aveL / aveR (PEM.cu:51-93): the parabola while the slope ratio is
// at most 2, the power law beyond
inline real_t pem_aveL(real_t x, const real_t* par) {
    real_t sL = par[0], aM = par[1], sR = par[2];
    if (sL * sR == 0.0) return aM;
    if (sR / sL >= 1.0) return sR / sL <= 2.0 ? ppm_aveL(x, par) : pem0(x, sL, aM - sL, sR / sL);
    return sL / sR <= 2.0 ? ppm_aveL(x, par) : pem1(x, sL, aM - sL, sL / sR);
}
inline real_t pem_aveR(real_t x, const real_t* par) {
    real_t sL = par[0], aM = par[1], sR = par[2];
    if (sR * sL == 0.0) return aM;
    if (sR / sL > 1.0) return sR / sL <= 2.0 ? ppm_aveR(x, par) : pem1(1.0 - x, -sR, aM + sR, sR / sL);
    return sL / sR <= 2.0 ? ppm_aveR(x, par) : pem0(1.0 - x, -sR, aM + sR, sL / sR);
}
// get_CON/PRM_parameters and _aveL/_aveR dispatch (recon.cu:383-501) for
// recon_flag 6 (PPM4: 4th-order slopes, flattening, PPM) and 5 (PEM4:
// 4th-order slopes, PEM averages, no flattening)
inline void con_par(int recon, int i, int geom, const real_t* x, const real_t* a, const real_t* flat, real_t par[3]) {
    fourth_parameters(i, geom, x, a, par);
    if (recon == 6) ppm_adjust(flat[i], par);
}
inline real_t con_aveL(int recon, real_t q, const real_t* par) { return recon == 5 ? pem_aveL(q, par) : ppm_aveL(q, par); }
inline real_t con_aveR(int recon, real_t q, const real_t* par) { return recon == 5 ? pem_aveR(q, par) : ppm_aveR(q, par); }

// flatten (PPM.cu:7-62) for cell i of a row of n cells
real_t flatten_cell(int i, int n, real_t gam, const real_t* r, const real_t* p, const real_t* u) {
    real_t f = 0.0;
    if (i >= 2 && i < n - 2) {
        real_t pb = p[i - 1], p0 = p[i], pf = p[i + 1];
        real_t cs = msqrt(gam * p0 / r[i]);
        real_t dp1f = pf - p0;
        real_t dp2f = p[i + 2] - pb;
        real_t dp1b = p0 - pb;
        real_t dp2b = pf - p[i - 2];
        real_t Mf = (u[i] - u[i + 1]) / cs;
        real_t Mb = (u[i - 1] - u[i]) / cs;
        if (Mf > 0.0 || Mb > 0.0) {
            if (dp2f != 0.0 && dp2b != 0.0) {
                real_t ff = 10.0 * mfmax(0.0, mfabs(dp1f / dp2f - 0.5) - 0.25);
                real_t fb = 10.0 * mfmax(0.0, mfabs(dp1b / dp2b - 0.5) - 0.25);
                f = mfmax(ff, fb);
                f = mfmin(0.5, f * mfmax(mfabs(dp1f), mfabs(dp1b)) / p0);
            }
        }
    }
    return f;
}

// ------------------------------------------------------------ Riemann step
struct State { real_t rl, pl, ul, vl, wl, rr, pr, ur, vr, wr; };

// set_state (states.cu:1-135) for interface i (between cells i-1 and i)
void trace_state(int i, int geom, int recon, real_t gam, const real_t* xa, const real_t* r, const real_t* p,
                 const real_t* u, const real_t* v, const real_t* w, const real_t* flat,
                 real_t dt, real_t us, State& S) {
    real_t rp[3], pp[3], up[3], vp[3], wp[3];
    real_t hdt = 0.5 * dt;
    // right state: cell i, swept region adjacent to its left edge
    con_par(recon, i, geom, xa, r, flat, rp);
    con_par(recon, i, geom, xa, p, flat, pp);
    con_par(recon, i, geom, xa, u, flat, up);
    real_t xl = xa[i], xr = xa[i + 1];
    real_t cs = msqrt(gam * p[i] / r[i]);
    real_t tmp = xl - mfmin(u[i], 0.0) * dt + cs * dt;
    real_t q = lim01((tmp - xl) / (xr - xl));
    real_t rmin = r[i] * 1.0e-10, pmin = p[i] * 1.0e-10;
    real_t ur = con_aveL(recon, q, up);
    real_t dis = mfmin(ur, 0.0) * dt;
    tmp = xl - dis;
    q = lim01((tmp - xl) / (xr - xl));
    real_t r0 = con_aveL(recon, q, rp), p0 = con_aveL(recon, q, pp), u0 = con_aveL(recon, q, up);
    r0 = mfmax(r0, rmin);
    p0 = mfmax(p0, pmin);
    cs = msqrt(gam * p0 / r0);
    tmp = xl - dis + cs * dt;
    q = lim01((tmp - xl) / (xr - xl));
    real_t u_ = con_aveL(recon, q, up), p_ = con_aveL(recon, q, pp);
    u_ = 2.0 * (u_ - u0) / (cs * dt);
    p_ = 2.0 * (p_ - p0) / (cs * dt);
    u_ *= -hdt;
    p_ *= -hdt;
    S.rr = r0 * exp_lim(u_);
    S.pr = p0 * exp_lim(gam * u_);
    S.ur = u0 + p_ / r0 + us;
    con_par(recon, i, geom, xa, v, flat, vp);
    con_par(recon, i, geom, xa, w, flat, wp);
    S.vr = con_aveL(recon, q, vp);
    S.wr = con_aveL(recon, q, wp);
    // left state: cell i-1, swept region adjacent to its right edge
    con_par(recon, i - 1, geom, xa, r, flat, rp);
    con_par(recon, i - 1, geom, xa, p, flat, pp);
    con_par(recon, i - 1, geom, xa, u, flat, up);
    xl = xa[i - 1];
    xr = xa[i];
    cs = msqrt(gam * p[i - 1] / r[i - 1]);
    tmp = xr - mfmax(u[i - 1], 0.0) * dt - cs * dt;
    q = lim01((tmp - xl) / (xr - xl));
    rmin = r[i - 1] * 1.0e-10;
    pmin = p[i - 1] * 1.0e-10;
    real_t ul = con_aveR(recon, q, up);
    dis = mfmax(ul, 0.0) * dt;
    tmp = xr - dis;
    q = lim01((tmp - xl) / (xr - xl));
    r0 = con_aveR(recon, q, rp);
    p0 = con_aveR(recon, q, pp);
    u0 = con_aveR(recon, q, up);
    r0 = mfmax(r0, rmin);
    p0 = mfmax(p0, pmin);
    cs = msqrt(gam * p0 / r0);
    tmp = xr - dis - cs * dt;
    q = lim01((tmp - xl) / (xr - xl));
    u_ = con_aveR(recon, q, up);
    p_ = con_aveR(recon, q, pp);
    u_ = 2.0 * (u0 - u_) / (cs * dt);
    p_ = 2.0 * (p0 - p_) / (cs * dt);
    u_ *= -hdt;
    p_ *= -hdt;
    S.rl = r0 * exp_lim(u_);
    S.pl = p0 * exp_lim(gam * u_);
    S.ul = u0 + p_ / r0 + us;
    con_par(recon, i - 1, geom, xa, v, flat, vp);
    con_par(recon, i - 1, geom, xa, w, flat, wp);
    S.vl = con_aveR(recon, q, vp);
    S.wl = con_aveR(recon, q, wp);
}

// get_pm_simple, get_lr_speeds_Einfeldt, get_sm (wave_speeds.cu:24-35,288-315)
void wave_speeds(const State& S, real_t gam, real_t us, real_t& pm, real_t& sl, real_t& sm, real_t& sr) {
    real_t cl = msqrt(gam * S.pl / S.rl);
    real_t cr = msqrt(gam * S.pr / S.rr);
    real_t c_ = 0.5 * (cr + cl);
    real_t p_ = 0.5 * (S.pr + S.pl);
    pm = p_ * exp_lim(0.5 * gam * (S.ul - S.ur) / c_);
    real_t srl = msqrt(S.rl), srr = msqrt(S.rr);
    real_t ubar = (srl * S.ul + srr * S.ur) / (srl + srr);
    real_t n2 = 0.5 * srl * srr / ((srl + srr) * (srl + srr));
    real_t cbar = msqrt((cl * cl * srl + cr * cr * srr) / (srl + srr) + n2 * (S.ur - S.ul) * (S.ur - S.ul));
    sl = ubar - cbar;
    sr = ubar + cbar;
    sm = (S.rl * S.ul * cl + S.rr * S.ur * cr) / (S.rl * cl + S.rr * cr);
    sm += (S.pl - S.pr) / (S.rl + S.rr) / (0.5 * (cl + cr));
    sm += us * (mfmax(S.ul, 0.0) + mfmax(-S.ur, 0.0)) / (cl + cr);
}

// upwind selection (HLLC.cu:2-8)
inline real_t pick(real_t FL, real_t F_L, real_t F_R, real_t FR, real_t sl, real_t sm, real_t sr) {
    if (!msignbit(sl)) return FL;
    if (msignbit(sr)) return FR;
    if (!msignbit(sm)) return F_L;
    return F_R;
}

real_t energy(int eos, real_t gam, real_t r, real_t p, real_t u, real_t v, real_t w) {  // EOS.cu:4-18
    (void)u; (void)v; (void)w;
    if (eos == 2) return p / ((gam - 1.0) * r);  // internal_e_flag 1
    if (eos == 1) return p / mpow(r, gam);
    return p / r;
}

// HLLC_fluxes (HLLC.cu:10-77), internal_e_flag 1 for EOS_flag 2
void hllc(const State& S, int eos, real_t gam, real_t pm, real_t sl, real_t sm, real_t sr,
          Cell& F, real_t& pres, real_t& uprs) {
    real_t fl = S.ul * S.rl;
    real_t f_l = sm * S.rl * (sl - S.ul) / (sl - sm);
    real_t f_r = sm * S.rr * (sr - S.ur) / (sr - sm);
    real_t fr = S.ur * S.rr;
    F.r = pick(fl, f_l, f_r, fr, sl, sm, sr);
    F.u = pick(S.ul, S.ul, S.ur, S.ur, sl, sm, sr) * F.r;
    F.v = pick(S.vl, S.vl, S.vr, S.vr, sl, sm, sr) * F.r;
    F.w = pick(S.wl, S.wl, S.wr, S.wr, sl, sm, sr) * F.r;
    if (eos > 0) {
        real_t el = energy(eos, gam, S.rl, S.pl, S.ul, 0.0, 0.0);
        real_t er = energy(eos, gam, S.rr, S.pr, S.ur, 0.0, 0.0);
        F.p = pick(el, el, er, er, sl, sm, sr) * F.r;
    } else {
        F.p = 0.0;
    }
    pres = pick(S.pl, (sl * pm - sm * S.pl) / (sl - sm), (sr * pm - sm * S.pr) / (sr - sm), S.pr, sl, sm, sr);
    if (eos == 2) {
        uprs = pick(S.ul, sm * (sl * pm - S.ul * S.pl) / (sl * pm - sm * S.pl),
                    sm * (sr * pm - S.ur * S.pr) / (sr * pm - sm * S.pr), S.ur, sl, sm, sr);
    } else {
        uprs = 0.0;
    }
}

inline real_t geom_factor(int geom, real_t x) {  // eval_gfac (riemann.cu:5-11)
    if (geom == 1) return x;
    if (geom == 2) return x * x;
    if (geom == 5) return msin(x);
    return 1.0;
}

struct RowScratch {
    std::vector<real_t> r, p, u, v, w, flat, pres, uprs, gfac;
    std::vector<Cell> F, Del;
    std::vector<Dust> FD, DelD;
    void resize(int n) {
        r.resize(n); p.resize(n); u.resize(n); v.resize(n); w.resize(n);
        flat.resize(n); pres.resize(n + 1); uprs.resize(n + 1); gfac.resize(n + 1);
        F.resize(n + 1); Del.resize(n); FD.resize(n + 1); DelD.resize(n);
    }
};

// riemann() + net_source() over a whole row of n cells (riemann.cu:40-131).
// Del[i] is valid for cells i in [kPad, n-kPad).
void riemann_row(int geom, int eos, real_t gam, int n, const real_t* xa, RowScratch& s,
                 real_t dt_state, int recon = 6) {
    if (recon == 6)   // flattening only under recon_flag 6 (riemann.cu:99-102)
        for (int i = 0; i < n; ++i) s.flat[i] = flatten_cell(i, n, gam, s.r.data(), s.p.data(), s.u.data());
    for (int i = kPad; i < n + 1 - kPad; ++i) {
        State S;
        trace_state(i, geom, recon, gam, xa, s.r.data(), s.p.data(), s.u.data(), s.v.data(), s.w.data(),
                    s.flat.data(), dt_state, 0.0, S);
        real_t pm, sl, sm, sr;
        wave_speeds(S, gam, 0.0, pm, sl, sm, sr);
        hllc(S, eos, gam, pm, sl, sm, sr, s.F[i], s.pres[i], s.uprs[i]);
        s.gfac[i] = geom_factor(geom, xa[i]);
    }
    for (int i = kPad; i < n - kPad; ++i) {
        real_t g0 = s.gfac[i], g1 = s.gfac[i + 1];
        real_t du = 0.5 * (g0 + g1) * (s.pres[i] - s.pres[i + 1]);
        real_t de = 0.0;
        if (eos == 2) de = 0.5 * (s.pres[i] + s.pres[i + 1]) * (s.uprs[i] * g0 - s.uprs[i + 1] * g1);
        Cell& D = s.Del[i];
        D.r = s.F[i].r * g0 - s.F[i + 1].r * g1;
        D.p = s.F[i].p * g0 - s.F[i + 1].p * g1 + de;
        D.u = s.F[i].u * g0 - s.F[i + 1].u * g1 + du;
        D.v = s.F[i].v * g0 - s.F[i + 1].v * g1;
        D.w = s.F[i].w * g0 - s.F[i + 1].w * g1;
    }
}

// ------------------------------------------------------------------ dust
// dust_flux (dust.cu:6-105), flattening coefficient 0
Dust dust_interface(int i, int geom, const real_t* xa, const real_t* r, const real_t* u,
                    const real_t* v, const real_t* w, const real_t* flat, real_t dt, real_t us) {
    State S;
    real_t rp[3], upar[3];
    real_t xl = xa[i], xr = xa[i + 1];
    recon_par(i, geom, xa, u, flat, upar);
    real_t tmp = xl - mfmin(upar[1] - upar[0], 0.0) * dt;
    real_t q = lim01((tmp - xl) / (xr - xl));
    real_t ur = ppm_aveL(q, upar);
    real_t u_ = 0.5 * dt * (ur - u[i]) / (0.5 * (xr - tmp));
    tmp = xl - mfmin(ur, 0.0) * dt;
    q = lim01((tmp - xl) / (xr - xl));
    recon_par(i, geom, xa, r, flat, rp);
    real_t r0 = ppm_aveL(q, rp);
    real_t u0 = ppm_aveL(q, upar);
    S.rr = r0 * exp_lim(u_);
    S.ur = u0 + us;
    recon_par(i, geom, xa, v, flat, rp);
    recon_par(i, geom, xa, w, flat, upar);
    S.vr = ppm_aveL(q, rp);
    S.wr = ppm_aveL(q, upar);

    xl = xa[i - 1];
    xr = xa[i];
    recon_par(i - 1, geom, xa, u, flat, upar);
    tmp = xr - mfmax(upar[1] + upar[2], 0.0) * dt;
    q = lim01((tmp - xl) / (xr - xl));
    real_t ul = ppm_aveR(q, upar);
    u_ = 0.5 * dt * (u[i - 1] - ul) / (0.5 * (tmp - xl));
    tmp = xr - mfmax(ul, 0.0) * dt;
    q = lim01((tmp - xl) / (xr - xl));
    recon_par(i - 1, geom, xa, r, flat, rp);
    r0 = ppm_aveR(q, rp);
    u0 = ppm_aveR(q, upar);
    S.rl = r0 * exp_lim(u_);
    S.ul = u0 + us;
    recon_par(i - 1, geom, xa, v, flat, rp);
    recon_par(i - 1, geom, xa, w, flat, upar);
    S.vl = ppm_aveR(q, rp);
    S.wl = ppm_aveR(q, upar);

    S.rl = mfmin(r[i - 1] * dt / (xa[i] - xa[i - 1]), mfmax(S.ul * S.rl, 0.0));
    S.rr = mfmax(r[i] * dt / (xa[i] - xa[i + 1]), mfmin(S.ur * S.rr, 0.0));
    Dust F;
    F.r = S.rl + S.rr;
    F.u = S.rl * S.ul + S.rr * S.ur;
    F.v = S.rl * S.vl + S.rr * S.vr;
    F.w = S.rl * S.wl + S.rr * S.wr;
    return F;
}

// dust_flow (dust.cu:107-147) over a row; rows hold r,u,v,w in s.r,s.u,s.v,s.w
void dust_row(int geom, int n, const real_t* xa, RowScratch& s, real_t dt_flux) {
    std::fill(s.flat.begin(), s.flat.begin() + n, 0.0);
    for (int i = kPad; i < n + 1 - kPad; ++i) {
        s.FD[i] = dust_interface(i, geom, xa, s.r.data(), s.u.data(), s.v.data(), s.w.data(),
                                 s.flat.data(), dt_flux, 0.0);
        s.gfac[i] = geom_factor(geom, xa[i]);
    }
    for (int i = kPad; i < n - kPad; ++i) {
        real_t g0 = s.gfac[i], g1 = s.gfac[i + 1];
        Dust& D = s.DelD[i];
        D.r = s.FD[i].r * g0 - s.FD[i + 1].r * g1;
        D.u = s.FD[i].u * g0 - s.FD[i + 1].u * g1;
        D.v = s.FD[i].v * g0 - s.FD[i + 1].v * g1;
        D.w = s.FD[i].w * g0 - s.FD[i + 1].w * g1;
    }
}

// ---------------------------------------------------------- cell updates
Cell update_cell(const pg_params& P, real_t x, Cell Q, const Cell& D) {  // sweeps.cu:218-261
    real_t r_min = mfmax(Q.r * 1.0e-10, kSmallR);
    real_t p_min = mfmax(Q.p * 1.0e-10, kSmallP);
    real_t uo = Q.u, vo = Q.v, wo = Q.w;
    Q.p = energy(P.EOS_flag, P.gam, Q.r, Q.p, Q.u, Q.v, Q.w);
    Q.p *= Q.r;
    Q.u *= Q.r; Q.v *= Q.r; Q.w *= Q.r;
    Q.r += D.r; Q.p += D.p; Q.u += D.u; Q.v += D.v; Q.w += D.w;
    Q.u /= Q.r; Q.v /= Q.r; Q.w /= Q.r;
    if (Q.r < r_min) {
        Q.r = r_min; Q.p = p_min; Q.u = uo; Q.v = vo; Q.w = wo;
    } else if (P.EOS_flag == 2) {
        Q.p = mfmax(Q.p * (P.gam - 1.0), p_min);
    } else if (P.EOS_flag == 0) {
        Q.p = cs2_at(P, x) * Q.r;
    }
    return Q;
}

Cell update_cell_advect(const pg_params& P, real_t x, Cell Q, const Cell& D) {  // advection.cu:88-105
    Q.u *= Q.r; Q.v *= Q.r; Q.w *= Q.r;
    Q.r += D.r; Q.p += D.p; Q.u += D.u; Q.v += D.v; Q.w += D.w;
    Q.u /= Q.r; Q.v /= Q.r; Q.w /= Q.r;
    if (P.EOS_flag == 0) Q.p = cs2_at(P, x) * Q.r;
    return Q;
}

Dust update_dust(const pg_params& P, Dust Q, const Dust& D) {  // dust.cu:149-176
    real_t r_min = mfmax(Q.r * 1.0e-10, P.D_G_ratio * kSmallR);
    real_t uo = Q.u, vo = Q.v, wo = Q.w;
    Q.u *= Q.r; Q.v *= Q.r; Q.w *= Q.r;
    Q.r += D.r; Q.u += D.u; Q.v += D.v; Q.w += D.w;
    Q.u /= Q.r; Q.v /= Q.r; Q.w /= Q.r;
    if (Q.r < r_min) { Q.r = r_min; Q.u = uo; Q.v = vo; Q.w = wo; }
    return Q;
}

// ----------------------------------------------------------------- planet
real_t ramp(real_t t, real_t tau, real_t val) { return t < tau ? val * msin(kHalfPi * t / tau) : val; }

real_t ecc_anomaly(real_t n, real_t e, real_t t) {  // planet.cu:60-103
    real_t l = mfmod(n * t, kTwoPi);
    real_t D = -e * msin(l);
    if (e == 0.0 || l == 0.0 || l == kPi || l == kTwoPi) return l;
    real_t tor = 1.0e-12, step;
    int count = 0;
    do {
        real_t es = msin(l - D), ec = mcos(l - D);
        es *= e; ec *= e;
        real_t x = D + es, dx = 1.0 - ec, d2x = -es;
        real_t eps = 2.0 * x * d2x / dx / dx;
        if (mfabs(eps) < 1.0) step = -(x / dx) * (1.0 + eps / 4.0 + eps * eps / 8.0 + 5.0 * eps * eps * eps / 64.0);
        else step = mcopysign(mfmin(mfabs(x / dx), kHalfPi / 2.0), -x);
        D += step;
        count++;
    } while (mfabs(step) > tor || count < 10);
    return l - D;
}

void orbit_solution(const pg_params& P, Body& b, real_t t) {  // planet.cu:105-133
    real_t a = b.a, e = b.e;
    real_t n = msqrt(1.0 / a / a / a);
    real_t efac = msqrt(1.0 - e * e);
    real_t B = e / (1.0 + efac);
    real_t u = ecc_anomaly(n, e, t);
    real_t sinu = msin(u), cosu = mcos(u);
    real_t f = u + 2.0 * matan(B * sinu / (1.0 - B * cosu));
    real_t esinf = msin(f) * e, ecosf = mcos(f) * e;
    b.m = ramp(t, P.ramp_time, P.planet_mass);
    b.x = a * efac * efac / (1.0 + ecosf);
    b.y = f - mfmod(P.frame_omega * t, kTwoPi) + kPi;
    b.z = kHalfPi;
    b.vx = n * a * esinf / efac;
    b.vy = n * (1.0 + ecosf) * (1.0 + ecosf) / efac / efac / efac;
    b.vz = 0.0;
}

// star_planet_grav_cyl (forces.cu:41-58), evaluated with dt = 0
void planet_force_cyl(real_t rad, real_t azi, real_t z, const Body& b, real_t& fx, real_t& fy, real_t& fz) {
    real_t s = msin(azi - (b.y + 0.0 * b.vy)), c = mcos(azi - (b.y + 0.0 * b.vy));
    real_t Rp = msqrt(rad * rad + b.x * b.x - 2.0 * b.x * rad * c + z * z + b.rs * b.rs);
    real_t k = 1.0 / Rp / Rp / Rp;
    fx = -b.m * (rad - b.x * c) * k - b.m * c / b.x / b.x;
    fy = -b.m * b.x * s * k + b.m * s / b.x / b.x;
    fz = -b.m * z * k;
}

// star_planet_grav (forces.cu:14-39), evaluated with dt = 0: the
// spline-softened planet in spherical coordinates (rad, azimuth, polar
// angle) plus the indirect term
void planet_force_sph(real_t rad, real_t azi, real_t pol, const Body& b, real_t& fx, real_t& fy, real_t& fz) {
    real_t s = msin(azi - (b.y + 0.0 * b.vy)), c = mcos(azi - (b.y + 0.0 * b.vy));
    real_t sp = msin(pol), cp = mcos(pol);
    real_t Rp = msqrt(rad * rad + b.x * b.x - 2.0 * b.x * rad * c * sp);
    real_t k = mfmax(4.0 - 3.0 * Rp / b.rs, 1.0) / mfmax(b.rs * b.rs * b.rs, Rp * Rp * Rp);
    fx = -b.m * (rad - b.x * c * sp) * k - b.m * sp * c / b.x / b.x;
    fy = -b.m * b.x * s * k + b.m * sp * s / b.x / b.x;
    fz = b.m * b.x * c * cp * k - b.m * cp * c / (sp * b.x * b.x);
}

}  // namespace

// ================================================================== grid
struct orc_grid {
    pg_params P;
    int xres, yres, xpad, ypad, xarr, yarr;
    int zres = 1, zpad = 0, zarr = 1;   // ndim 3: the polar-angle axis
    std::vector<real_t> xa, xv, ya, yv, za, zv;
    std::vector<Cell> C, T;
    std::vector<Dust> CD, TD;
    std::vector<real_t> vis;
    std::vector<real_t> rot, res;   // per (i, k): ik(i, k)
    std::vector<int> inc;
    Body planet;
    std::vector<Cell> A;        // time average of the gas (Grid::A)
    real_t ave_interval = 0.0;  // 0: not averaging
    size_t ind(int i, int j, int k = 0) const { return (size_t)i + (size_t)xarr * ((size_t)j + (size_t)yarr * k); }
    size_t ik(int i, int k) const { return (size_t)i + (size_t)xarr * k; }
    real_t xc(int i) const { return 0.5 * (xa[i + 1] + xa[i]); }
    real_t yc(int j) const { return 0.5 * (ya[j + 1] + ya[j]); }
    real_t zc(int k) const { return 0.5 * (za[k + 1] + za[k]); }
    bool sph() const { return P.ndim == 3; }
    // cylindrical radius of cell (i, k) (sweeps.cu:288-291, 375-381)
    real_t rcyl(int i, int k) const { return sph() ? xc(i) * msin(zc(k)) : xc(i); }
    // the analytic state at the centre of cell (i, k) (init_C, init.cu:456-468)
    Cell init_at(int i, int k) const { return sph() ? cell_init3(P, xc(i), zc(k)) : cell_init(P, xc(i)); }
    // interior k range (a single layer in 2D)
    int k0() const { return zpad; }
    int k1() const { return zpad + zres; }
};

namespace {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// uniform in [0,1) keyed by (seed, field, interior i, interior j)
real_t keyed_uniform(uint64_t seed, int field, int i, int j) {
    uint64_t key = splitmix64(seed) ^ splitmix64(((uint64_t)field << 48) ^ ((uint64_t)j << 24) ^ (uint64_t)i);
    return (real_t)(splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

// boundx / boundx_dust / boundx2 for one device with types 0 or 1
// (boundary.cu:13-113, 462-560)
void fill_x_ghosts(orc_grid& g, std::vector<Cell>& X) {
    const pg_params& P = g.P;
    for (int k = 0; k < g.zarr; ++k)
    for (int j = 0; j < g.yarr; ++j) {
        for (int n = 0; n < g.xpad; ++n) {
            size_t l = g.ind(n, j, k), r = g.ind(g.xres + g.xpad + n, j, k);
            if (P.bound_lft == 0) X[l] = g.init_at(n, k);
            else if (P.bound_lft == 1) X[l] = X[g.ind(g.xpad, j, k)];
            if (P.bound_rgh == 0) X[r] = g.init_at(g.xres + g.xpad + n, k);
            else if (P.bound_rgh == 1) X[r] = X[g.ind(g.xres + g.xpad - 1, j, k)];
        }
    }
}
void fill_x_ghosts_dust(orc_grid& g, std::vector<Dust>& X) {
    const pg_params& P = g.P;
    for (int j = 0; j < g.yarr; ++j) {
        for (int n = 0; n < g.xpad; ++n) {
            size_t l = g.ind(n, j), r = g.ind(g.xres + g.xpad + n, j);
            if (P.bound_lft == 0) X[l] = dust_init(P, g.xc(n));
            else if (P.bound_lft == 1) X[l] = X[g.ind(g.xpad, j)];
            if (P.bound_rgh == 0) X[r] = dust_init(P, g.xc(g.xres + g.xpad + n));
            else if (P.bound_rgh == 1) X[r] = X[g.ind(g.xres + g.xpad - 1, j)];
        }
    }
}
// boundy / boundy_dust, periodic (boundary.cu:141-268), all i
template <class T>
void fill_y_ghosts(orc_grid& g, std::vector<T>& X) {
    for (int k = 0; k < g.zarr; ++k)
    for (int i = 0; i < g.xarr; ++i)
        for (int n = 0; n < g.ypad; ++n) {
            X[g.ind(i, n, k)] = X[g.ind(i, g.yres + n, k)];
            X[g.ind(i, g.yres + g.ypad + n, k)] = X[g.ind(i, g.ypad + n, k)];
        }
}
// boundz (boundary.cu:270-321), all i and j: 3 periodic, 2 reflecting (w
// negated), 1 zero gradient, 0 the analytic state
void fill_z_ghosts(orc_grid& g, std::vector<Cell>& X) {
    const pg_params& P = g.P;
    if (g.zpad == 0) return;
    for (int j = 0; j < g.yarr; ++j)
        for (int i = 0; i < g.xarr; ++i)
            for (int n = 0; n < g.zpad; ++n) {
                Cell T;
                if (P.bound_bom == 3) T = X[g.ind(i, j, g.zres + n)];
                else if (P.bound_bom == 2) { T = X[g.ind(i, j, 2 * g.zpad - 1 - n)]; T.w *= -1.0; }
                else if (P.bound_bom == 1) T = X[g.ind(i, j, g.zpad)];
                else T = g.init_at(i, n);
                X[g.ind(i, j, n)] = T;
                if (P.bound_top == 3) T = X[g.ind(i, j, g.zpad + n)];
                else if (P.bound_top == 2) { T = X[g.ind(i, j, g.zres + g.zpad - 1 - n)]; T.w *= -1.0; }
                else if (P.bound_top == 1) T = X[g.ind(i, j, g.zres + g.zpad - 1)];
                else T = g.init_at(i, g.zres + g.zpad + n);
                X[g.ind(i, j, g.zres + g.zpad + n)] = T;
            }
}

// ------------------------------------------------------------------ sweeps
void sweep_x_gas(orc_grid& g, real_t dt) {  // sweeps.cu:263-358
    const pg_params& P = g.P;
    const int n = g.xarr;
    const int nrow = g.yres * g.zres;
#pragma omp parallel
    {
        RowScratch s;
        s.resize(n);
#pragma omp for schedule(static)
        for (int row = 0; row < nrow; ++row) {
            const int j = g.ypad + row % g.yres, k = g.zpad + row / g.yres;
            for (int i = 0; i < n; ++i) {
                const Cell& c = g.C[g.ind(i, j, k)];
                real_t rad = 0.5 * (g.xa[i + 1] + g.xa[i]);
                s.r[i] = c.r; s.p[i] = c.p; s.u[i] = c.u;
                if (g.sph()) {   // angular momenta: v R_cyl, w r
                    s.v[i] = c.v * (rad * msin(g.zc(k)));
                    s.w[i] = c.w * rad;
                } else {
                    s.v[i] = c.v * rad; s.w[i] = c.w;
                }
            }
            riemann_row(P.geomx, P.EOS_flag, P.gam, n, g.xa.data(), s, dt / 1.0, P.recon_flag);
            for (int i = g.xpad; i < n - g.xpad; ++i) {
                Cell D = s.Del[i];
                real_t f = dt / g.xv[i];
                D.r *= f; D.p *= f; D.u *= f; D.v *= f; D.w *= f;
                real_t rad = 0.5 * (g.xa[i + 1] + g.xa[i]);
                if (g.sph()) {
                    D.v /= rad * msin(g.zc(k));
                    D.w /= rad;
                } else {
                    D.v /= rad;
                }
                g.T[g.ind(i, j, k)] = update_cell(P, g.xc(i), g.C[g.ind(i, j, k)], D);
            }
        }
    }
    std::swap(g.C, g.T);
}

void sweep_y_gas(orc_grid& g, real_t dt) {  // sweeps.cu:361-456
    const pg_params& P = g.P;
    const int n = g.yarr;
    const int nrow = g.xres * g.zres;
#pragma omp parallel
    {
        RowScratch s;
        s.resize(n);
#pragma omp for schedule(static)
        for (int row = 0; row < nrow; ++row) {
            const int i = g.xpad + row % g.xres, k = g.zpad + row / g.xres;
            real_t R = g.rcyl(i, k);
            real_t rot = g.rot[g.ik(i, k)];
            for (int j = 0; j < n; ++j) {
                const Cell& c = g.C[g.ind(i, j, k)];
                s.r[j] = c.r; s.p[j] = c.p;
                s.u[j] = c.v - rot;
                s.u[j] -= R * P.frame_omega;
                s.v[j] = c.w; s.w[j] = c.u;
            }
            riemann_row(P.geomy, P.EOS_flag, P.gam, n, g.ya.data(), s, dt / R, P.recon_flag);
            for (int j = g.ypad; j < n - g.ypad; ++j) {
                Cell D = s.Del[j];
                real_t f = dt / g.yv[j] / R;
                D.r *= f; D.p *= f; D.u *= f; D.v *= f; D.w *= f;
                D.u += (rot + R * P.frame_omega) * D.r;
                real_t t = D.v;
                D.v = D.u; D.u = D.w; D.w = t;
                g.T[g.ind(i, j, k)] = update_cell(P, g.xc(i), g.C[g.ind(i, j, k)], D);
            }
        }
    }
    std::swap(g.C, g.T);
}

void sweep_z_gas(orc_grid& g, real_t dt) {  // sweepz_inplace (sweeps.cu:458-553), geomz 5
    const pg_params& P = g.P;
    const int n = g.zarr;
    const int nrow = g.xres * g.yres;
#pragma omp parallel
    {
        RowScratch s;
        s.resize(n);
#pragma omp for schedule(static)
        for (int row = 0; row < nrow; ++row) {
            const int i = g.xpad + row % g.xres, j = g.ypad + row / g.xres;
            const real_t rad = g.xc(i);
            for (int k = 0; k < n; ++k) {
                const Cell& c = g.C[g.ind(i, j, k)];
                // normal w, transverse u and the angular momentum v R_cyl
                s.r[k] = c.r; s.p[k] = c.p; s.u[k] = c.w; s.v[k] = c.u;
                s.w[k] = c.v * (rad * msin(0.5 * (g.za[k + 1] + g.za[k])));
            }
            riemann_row(P.geomz, P.EOS_flag, P.gam, n, g.za.data(), s, dt / rad, P.recon_flag);
            for (int k = g.zpad; k < n - g.zpad; ++k) {
                Cell D = s.Del[k];
                real_t f = dt / g.zv[k];
                D.r *= f; D.p *= f; D.u *= f; D.v *= f; D.w *= f;
                f = 1.0 / rad;
                D.r *= f; D.p *= f; D.u *= f; D.v *= f; D.w *= f;
                D.w /= rad * msin(0.5 * (g.za[k + 1] + g.za[k]));
                real_t t = D.w;
                D.w = D.u; D.u = D.v; D.v = t;
                g.T[g.ind(i, j, k)] = update_cell(P, g.xc(i), g.C[g.ind(i, j, k)], D);
            }
        }
    }
    std::swap(g.C, g.T);
}

void sweep_x_dust(orc_grid& g, real_t dt) {  // dust.cu:178-257
    const pg_params& P = g.P;
    const int n = g.xarr;
#pragma omp parallel
    {
        RowScratch s;
        s.resize(n);
#pragma omp for schedule(static)
        for (int j = g.ypad; j < g.ypad + g.yres; ++j) {
            for (int i = 0; i < n; ++i) {
                const Dust& c = g.CD[g.ind(i, j)];
                real_t rad = 0.5 * (g.xa[i + 1] + g.xa[i]);
                s.r[i] = c.r; s.u[i] = c.u; s.v[i] = c.v * rad; s.w[i] = c.w;
            }
            dust_row(P.geomx, n, g.xa.data(), s, dt);
            for (int i = g.xpad; i < n - g.xpad; ++i) {
                Dust D = s.DelD[i];
                real_t f = dt / g.xv[i];
                D.r *= f; D.u *= f; D.v *= f; D.w *= f;
                real_t rad = 0.5 * (g.xa[i + 1] + g.xa[i]);
                D.v /= rad;
                g.TD[g.ind(i, j)] = update_dust(P, g.CD[g.ind(i, j)], D);
            }
        }
    }
    std::swap(g.CD, g.TD);
}

void sweep_y_dust(orc_grid& g, real_t dt) {  // dust.cu:260-353
    const pg_params& P = g.P;
    const int n = g.yarr;
#pragma omp parallel
    {
        RowScratch s;
        s.resize(n);
#pragma omp for schedule(static)
        for (int i = g.xpad; i < g.xpad + g.xres; ++i) {
            real_t R = g.xc(i);
            real_t rot = g.rot[i];
            for (int j = 0; j < n; ++j) {
                const Dust& c = g.CD[g.ind(i, j)];
                s.r[j] = c.r;
                s.u[j] = c.v - rot;
                s.u[j] -= R * P.frame_omega;
                s.v[j] = c.w; s.w[j] = c.u;
            }
            dust_row(P.geomy, n, g.ya.data(), s, dt / R);
            for (int j = g.ypad; j < n - g.ypad; ++j) {
                Dust D = s.DelD[j];
                real_t f = dt / g.yv[j];
                D.r *= f; D.u *= f; D.v *= f; D.w *= f;
                real_t f2 = 1.0 / R;
                D.r *= f2; D.u *= f2; D.v *= f2; D.w *= f2;
                D.u += (rot + R * P.frame_omega) * D.r;
                real_t t = D.v;
                D.v = D.u; D.u = D.w; D.w = t;
                g.TD[g.ind(i, j)] = update_dust(P, g.CD[g.ind(i, j)], D);
            }
        }
    }
    std::swap(g.CD, g.TD);
}

// --------------------------------------------------------- orbital advection
void set_orbital_advection(orc_grid& g, real_t dt) {  // orbit.cu:74-99
    const pg_params& P = g.P;
    real_t dy0 = (P.ymax - P.ymin) / (real_t)P.yres;
    for (int k = g.k0(); k < g.k1(); ++k)
    for (int i = g.xpad; i < g.xpad + g.xres; ++i) {
        real_t dy = dy0 * g.rcyl(i, k);
        real_t rot = g.rot[g.ik(i, k)];
        int shift = (int)val(mfloor(rot * dt / dy + 0.5));
        real_t speed = shift * dy / dt;
        g.res[g.ik(i, k)] = rot - speed;
        g.inc[g.ik(i, k)] = shift;
    }
}

void shift_orbital_advection(orc_grid& g) {  // orbit.cu:17-49, 144-157
    const int jmax = g.yres;
    const int nrow = g.xres * g.zres;
#pragma omp parallel for schedule(static)
    for (int row = 0; row < nrow; ++row) {
        const int i = g.xpad + row % g.xres, k = g.zpad + row / g.xres;
        int inc = g.inc[g.ik(i, k)];
        for (int j = 0; j < jmax; ++j) {
            int jj = (j + inc) % jmax;
            if (jj < 0) jj += jmax;
            g.T[g.ind(i, jj + g.ypad, k)] = g.C[g.ind(i, j + g.ypad, k)];
            if (g.P.dust_flag) g.TD[g.ind(i, jj + g.ypad, k)] = g.CD[g.ind(i, j + g.ypad, k)];
        }
    }
    std::swap(g.C, g.T);
    if (g.P.dust_flag) std::swap(g.CD, g.TD);
}

// advection() row phases (advection.cu:1-86 gas, 175-252 dust), flat = 0
void advect_row(int geom, int n, const real_t* xa, RowScratch& s, real_t speed, real_t dt_eff,
                bool gas, std::vector<Cell>& flux, int recon = 6) {
    std::fill(s.flat.begin(), s.flat.begin() + n, 0.0);
    real_t par[3];
    for (int i = kPad; i < n + 1 - kPad; ++i) {
        Cell D{0.0, 0.0, 0.0, 0.0, 0.0};
        if (speed > 0.0) {
            real_t xl = xa[i - 1], xr = xa[i];
            real_t x0 = xr - speed * dt_eff;
            real_t q = lim01((x0 - xl) / (xr - xl));
            con_par(recon, i - 1, geom, xa, s.r.data(), s.flat.data(), par); D.r = con_aveR(recon, q, par);
            if (gas) { con_par(recon, i - 1, geom, xa, s.p.data(), s.flat.data(), par); D.p = con_aveR(recon, q, par); }
            con_par(recon, i - 1, geom, xa, s.u.data(), s.flat.data(), par); D.u = con_aveR(recon, q, par) * D.r;
            con_par(recon, i - 1, geom, xa, s.v.data(), s.flat.data(), par); D.v = con_aveR(recon, q, par) * D.r;
            con_par(recon, i - 1, geom, xa, s.w.data(), s.flat.data(), par); D.w = con_aveR(recon, q, par) * D.r;
        } else {
            real_t xl = xa[i], xr = xa[i + 1];
            real_t x0 = xl - speed * dt_eff;
            real_t q = lim01((x0 - xl) / (xr - xl));
            con_par(recon, i, geom, xa, s.r.data(), s.flat.data(), par); D.r = con_aveL(recon, q, par);
            if (gas) { con_par(recon, i, geom, xa, s.p.data(), s.flat.data(), par); D.p = con_aveL(recon, q, par); }
            con_par(recon, i, geom, xa, s.u.data(), s.flat.data(), par); D.u = con_aveL(recon, q, par) * D.r;
            con_par(recon, i, geom, xa, s.v.data(), s.flat.data(), par); D.v = con_aveL(recon, q, par) * D.r;
            con_par(recon, i, geom, xa, s.w.data(), s.flat.data(), par); D.w = con_aveL(recon, q, par) * D.r;
        }
        D.r *= speed; D.p *= speed; D.u *= speed; D.v *= speed; D.w *= speed;
        flux[i] = D;
        s.gfac[i] = geom_factor(geom, xa[i]);
    }
    for (int i = kPad; i < n - kPad; ++i) {
        real_t g0 = s.gfac[i], g1 = s.gfac[i + 1];
        Cell& D = s.Del[i];
        D.r = flux[i].r * g0 - flux[i + 1].r * g1;
        D.p = flux[i].p * g0 - flux[i + 1].p * g1;
        D.u = flux[i].u * g0 - flux[i + 1].u * g1;
        D.v = flux[i].v * g0 - flux[i + 1].v * g1;
        D.w = flux[i].w * g0 - flux[i + 1].w * g1;
    }
}

void advect_y_gas(orc_grid& g, real_t dt) {  // advection.cu:107-173, 321-338
    const pg_params& P = g.P;
    fill_y_ghosts(g, g.C);
    const int n = g.yarr;
    const int nrow = g.xres * g.zres;
#pragma omp parallel
    {
        RowScratch s;
        s.resize(n);
        std::vector<Cell> flux(n + 1);
#pragma omp for schedule(static)
        for (int row = 0; row < nrow; ++row) {
            const int i = g.xpad + row % g.xres, k = g.zpad + row / g.xres;
            real_t R = g.rcyl(i, k);
            for (int j = 0; j < n; ++j) {
                const Cell& c = g.C[g.ind(i, j, k)];
                s.r[j] = c.r; s.p[j] = c.p; s.u[j] = c.v; s.v[j] = c.w; s.w[j] = c.u;
            }
            advect_row(P.geomy, n, g.ya.data(), s, g.res[g.ik(i, k)], dt / R, true, flux, P.recon_flag);
            for (int j = g.ypad; j < n - g.ypad; ++j) {
                Cell D = s.Del[j];
                real_t f = dt / g.yv[j];
                D.r *= f; D.p *= f; D.u *= f; D.v *= f; D.w *= f;
                real_t f2 = 1.0 / R;
                D.r *= f2; D.p *= f2; D.u *= f2; D.v *= f2; D.w *= f2;
                real_t t = D.v;
                D.v = D.u; D.u = D.w; D.w = t;
                g.T[g.ind(i, j, k)] = update_cell_advect(P, g.xc(i), g.C[g.ind(i, j, k)], D);
            }
        }
    }
    std::swap(g.C, g.T);
}

void advect_y_dust(orc_grid& g, real_t dt) {  // advection.cu:254-319, 340-357
    const pg_params& P = g.P;
    fill_y_ghosts(g, g.CD);
    const int n = g.yarr;
#pragma omp parallel
    {
        RowScratch s;
        s.resize(n);
        std::vector<Cell> flux(n + 1);
#pragma omp for schedule(static)
        for (int i = g.xpad; i < g.xpad + g.xres; ++i) {
            real_t R = g.xc(i);
            for (int j = 0; j < n; ++j) {
                const Dust& c = g.CD[g.ind(i, j)];
                s.r[j] = c.r; s.u[j] = c.v; s.v[j] = c.w; s.w[j] = c.u;
            }
            advect_row(P.geomy, n, g.ya.data(), s, g.res[i], dt / R, false, flux);
            for (int j = g.ypad; j < n - g.ypad; ++j) {
                Cell D = s.Del[j];
                real_t f = dt / g.yv[j];
                D.r *= f; D.u *= f; D.v *= f; D.w *= f;
                real_t f2 = 1.0 / R;
                D.r *= f2; D.u *= f2; D.v *= f2; D.w *= f2;
                real_t t = D.v;
                D.v = D.u; D.u = D.w; D.w = t;
                Dust DD{D.r, D.u, D.v, D.w};
                g.TD[g.ind(i, j)] = update_dust(P, g.CD[g.ind(i, j)], DD);
            }
        }
    }
    std::swap(g.CD, g.TD);
}

// ---------------------------------------------------------------- viscosity
// ndim 3 evaluates the same (r, phi) stress divergence on every polar layer
// with the spherical radius and nu at the layer's angle; its z component is
// zero and the z derivatives it computes are unused (viscosity.cu:169-181)
void viscous_forces(orc_grid& g, const std::vector<Cell>& X) {  // viscosity.cu:34-197
    const pg_params& P = g.P;
    const int kb = g.sph() ? 1 : 0, ke = g.sph() ? g.zarr - 1 : 1;
    const int nrow = (g.yarr - 2) * (ke - kb);
#pragma omp parallel for schedule(static)
    for (int row = 0; row < nrow; ++row) {
        const int j = 1 + row % (g.yarr - 2), k = kb + row / (g.yarr - 2);
        for (int i = 1; i < g.xarr - 1; ++i) {
            real_t r[3], t[3], u[9], v[9], nr[9];
            for (int n = 0; n < 3; ++n) { r[n] = g.xc(i + n - 1); t[n] = g.yc(j + n - 1); }
            real_t nu0, nu1, nu2;
            if (g.sph()) {
                const real_t z = g.zc(k);
                nu0 = nu_3(P, r[0], z); nu1 = nu_3(P, r[1], z); nu2 = nu_3(P, r[2], z);
            } else {
                nu0 = nu_at(P, r[0]); nu1 = nu_at(P, r[1]); nu2 = nu_at(P, r[2]);
            }
            for (int n = 0; n < 9; ++n) {
                size_t id;
                if (n < 3) id = g.ind(i + n - 1, j, k);
                else if (n < 6) id = g.ind(i + n - 4, j - 1, k);
                else id = g.ind(i + n - 7, j + 1, k);
                nr[n] = X[id].r; u[n] = X[id].u; v[n] = X[id].v;
                if (n == 0 || n == 3 || n == 6) nr[n] *= nu0;
                else if (n == 2 || n == 5 || n == 8) nr[n] *= nu2;
                else nr[n] *= nu1;
            }
            auto d1 = [&](const real_t* a, real_t& dx, real_t& dxx) {
                dx = (a[2] - a[0]) / (r[2] - r[0]);
                dxx = ((a[2] - a[1]) / (r[2] - r[1]) - (a[1] - a[0]) / (r[1] - r[0])) / (0.5 * (r[2] - r[0]));
            };
            auto d2 = [&](const real_t* a, real_t& dy, real_t& dyy, real_t& dxy) {
                dy = (a[7] - a[4]) / (t[2] - t[0]) / r[1];
                dyy = ((a[7] - a[1]) / (t[2] - t[1]) - (a[1] - a[4]) / (t[1] - t[0])) / (0.5 * (t[2] - t[0])) / r[1] / r[1];
                dxy = (a[8] - a[6] - a[5] + a[3]) / (r[2] - r[0]) / (t[2] - t[0]) / r[1];
            };
            real_t dudx, d2udx2, dvdx, d2vdx2, dvdy, d2vdy2, d2vdxy, dudy, d2udy2, d2udxy;
            d1(u, dudx, d2udx2);
            real_t dndx = (nr[2] - nr[0]) / (r[2] - r[0]);
            d1(v, dvdx, d2vdx2);
            d2(v, dvdy, d2vdy2, d2vdxy);
            d2(u, dudy, d2udy2, d2udxy);
            real_t dndy = (nr[7] - nr[4]) / (t[2] - t[0]) / r[1];
            real_t sxx = 2.0 * nr[1] * dudx;
            real_t fx = 2.0 * dndx * dudx + 2.0 * nr[1] * d2udx2 + sxx / r[1];
            real_t syy = 2.0 * nr[1] * (dvdy + u[1] / r[1]);
            real_t sxy = nr[1] * (dudy + dvdx - v[1] / r[1]);
            real_t fy = 2.0 * dndy * dvdy + 2.0 * nr[1] * d2vdy2 + 2.0 * nr[1] * dudy / r[1] + 2.0 * dndy * u[1] / r[1];
            fx += dndy * (dudy + dvdx) + nr[1] * (d2udy2 + d2vdxy) - dndy * v[1] / r[1] - nr[1] * dvdy / r[1] - syy / r[1];
            fy += dndx * (dudy + dvdx) + nr[1] * (d2udxy + d2vdx2) - dndx * v[1] / r[1] - nr[1] * dvdx / r[1] +
                  nr[1] * v[1] / r[1] / r[1] + 2.0 * sxy / r[1];
            size_t id = g.ind(i, j, k);
            g.vis[2 * id + 0] = fx;
            g.vis[2 * id + 1] = fy;
        }
    }
}

void viscosity_eval(orc_grid& g, bool on_T) {  // viscosity.cu:489-549
    std::vector<Cell>& X = on_T ? g.T : g.C;
    fill_z_ghosts(g, X);
    fill_y_ghosts(g, X);
    fill_x_ghosts(g, X);
    viscous_forces(g, X);
}

// ----------------------------------------------------------- source terms
// apply_source_terms (sweeps.cu:597-711), ndim 2, geomx 1, cool_flag 0
void source_terms(orc_grid& g, std::vector<Cell>& in, std::vector<Dust>& in_d,
                  std::vector<Cell>& out, std::vector<Dust>& out_d, real_t mdt, real_t dt) {
    const pg_params& P = g.P;
    const bool dust = P.dust_flag != 0;
    const bool visc = P.visc_flag != 0;
    const Body b = g.planet;
    const int np = P.n_planet;
    const bool sph = g.sph();
    // ghost +- 1 in every direction (sweeps.cu:622-624)
    const int kb = sph ? g.zpad - 1 : 0, ke = sph ? g.zarr - g.zpad + 1 : 1;
    const int jb = g.ypad - 1, nj = g.yarr - 2 * g.ypad + 2;
    const int nrow = nj * (ke - kb);
#pragma omp parallel for schedule(static)
    for (int row = 0; row < nrow; ++row) {
        const int j = jb + row % nj, k = kb + row / nj;
        for (int i = g.xpad - 1; i < g.xarr - g.xpad + 1; ++i) {
            size_t id = g.ind(i, j, k);
            Cell C = in[id], T = C;
            Dust Q{0, 0, 0, 0}, D{0, 0, 0, 0};
            if (dust) { Q = in_d[id]; D = Q; }
            real_t xc = g.xc(i);
            real_t yc = g.yc(j) + g.rot[g.ik(i, k)] * 0.0;
            real_t zc = sph ? g.zc(k) : real_t(0.0);
            // get_gx / get_gy / get_gz (forces.cu:182-327), twobd_flag 0
            real_t gx, gy, gz;
            if (sph) {
                gx = -1.0 / xc / xc;
                gy = 0.0;
                gz = 0.0;
                for (int m = 0; m < np; ++m) {
                    real_t fx, fy, fz;
                    planet_force_sph(xc, yc, zc, b, fx, fy, fz);
                    gx += fx; gy += fy; gz += fz;
                }
                if (zc < P.zmin) gz *= -1.0;
            } else {
                real_t dis = msqrt(xc * xc + zc * zc);
                gx = -xc / dis / dis / dis;
                gy = 0.0;
                gz = -zc / dis / dis / dis;
                for (int m = 0; m < np; ++m) {
                    real_t fx, fy, fz;
                    planet_force_cyl(xc, yc, zc, b, fx, fy, fz);
                    gx += fx;
                }
                for (int m = 0; m < np; ++m) {
                    real_t fx, fy, fz;
                    planet_force_cyl(xc, yc, zc, b, fx, fy, fz);
                    gy += fy;
                }
                for (int m = 0; m < np; ++m) {
                    real_t fx, fy, fz;
                    planet_force_cyl(xc, yc, zc, b, fx, fy, fz);
                    gz += fz;
                }
            }
            real_t fx = 0.0, fy = 0.0, fz = 0.0;
            real_t fxd = 0.0, fyd = 0.0, fzd = 0.0, ts = 0.0;
            auto gas_forces = [&]() {  // get_g_forces (sweeps.cu:555-576), get_fx/fy/fz (forces.cu:110-180)
                if (sph) {   // the viscous z component is zero
                    fz = T.v * T.v / xc / mtan(zc);
                    if (zc < P.zmin) fz *= -1.0;
                }
                fy = 0.0;
                if (visc) fy += g.vis[2 * id + 1] / T.r;
                fx = sph ? T.v * T.v / xc + T.w * T.w / xc : T.v * T.v / xc;
                if (visc) fx += g.vis[2 * id + 0] / T.r;
            };
            auto dust_forces = [&]() {  // get_d_forces (sweeps.cu:578-594)
                fyd = 0.0 * ts + T.v - Q.v;
                fxd = (D.v * D.v / xc) * ts + T.u - Q.u;
            };
            if (dust) {
                ts = P.Stokes / (msqrt(P.gam * T.p / T.r) / P.sc_h) / T.r;  // dust.cu:1-4
                dust_forces();
            }
            gas_forces();
            if (mdt != 0.0) {
                T.u = C.u + fx * mdt + gx * mdt;
                T.v = C.v + fy * mdt + gy * mdt;
                T.w = C.w + fz * mdt + gz * mdt;
                if (dust) {
                    real_t dtau = 1.0 - mexp(-mdt / ts);
                    D.u = Q.u + fxd * dtau + gx * ts * dtau;
                    D.v = Q.v + fyd * dtau + gy * ts * dtau;
                    D.w = Q.w + fzd * dtau + gz * ts * dtau;
                    ts = P.Stokes / (msqrt(P.gam * T.p / T.r) / P.sc_h) / T.r;
                    dust_forces();
                }
                gas_forces();
            }
            Cell o;
            o.u = C.u + fx * dt + gx * dt;
            o.v = C.v + fy * dt + gy * dt;
            o.w = C.w + fz * dt + gz * dt;
            o.r = in[id].r;
            o.p = in[id].p;
            out[id] = o;
            if (dust) {
                real_t dtau = 1.0 - mexp(-dt / ts);
                Dust od;
                od.u = Q.u + fxd * dtau + gx * ts * dtau;
                od.v = Q.v + fyd * dtau + gy * ts * dtau;
                od.w = Q.w + fzd * dtau + gz * ts * dtau;
                od.r = in_d[id].r;
                out_d[id] = od;
            }
        }
    }
}

// ----------------------------------------------------------------- killwave
void killwave(orc_grid& g, real_t dt) {  // killwave.cu:8-117
    const pg_params& P = g.P;
    real_t d_in = h_at(P, P.xmin) * P.kill_width;
    real_t d_out = h_at(P, P.xmax) * P.kill_width;
    real_t inner = P.xmin + d_in, outer = P.xmax - d_out;
#pragma omp parallel for schedule(static)
    for (int j = g.ypad; j < g.ypad + g.yres; ++j) {
        for (int i = g.xpad; i < g.xpad + g.xres; ++i) {
            real_t rad = g.xc(i);
            Cell& c = g.C[g.ind(i, j)];
            for (int side = 0; side < 2; ++side) {
                bool in_zone = side == 0 ? rad < inner : rad > outer;
                if (!in_zone) continue;
                real_t f = side == 0 ? (inner - rad) / d_in : (rad - outer) / d_out;
                real_t tau = mpow(rad, 1.5);
                real_t s = (real_t)sinf((float)val(kHalfPi * f));
                f = s * s;
                f *= dt / tau;
                Cell I = cell_init(P, rad);
                c.r += f * (I.r - c.r);
                c.p += f * (I.p - c.p);
                c.u += f * (I.u - c.u);
                c.v += f * (I.v - c.v);
                c.w += f * (I.w - c.w);
            }
        }
    }
}

// ------------------------------------------------------------------- CFL
real_t cfl_rate(const orc_grid& g) {  // get_dt_lv1 (timestep.cu:18-97), ndim 2, geomy 3
    const pg_params& P = g.P;
    double best = 0.0;   // (the step size is frozen in the tangent-linear build)
    const int nrow = g.yres * g.zres;
#pragma omp parallel for reduction(max : best) schedule(static)
    for (int row = 0; row < nrow; ++row) {
        const int j = g.ypad + row % g.yres, k = g.zpad + row / g.yres;
        for (int i = g.xpad; i < g.xpad + g.xres; ++i) {
            size_t id = g.ind(i, j, k);
            const Cell& c = g.C[id];
            real_t rad = 0.5 * (g.xa[i + 1] + g.xa[i]);
            real_t rcyl = g.sph() ? rad * msin(0.5 * (g.za[k + 1] + g.za[k])) : rad;
            real_t rot = g.rot[g.ik(i, k)];
            real_t cs = msqrt(P.gam * c.p / c.r);
            real_t spd = mfabs(c.u) + cs;
            if (P.dust_flag) spd = mfmax(mfabs(g.CD[id].u), spd);
            real_t wid = mfabs(g.xa[i + 1] - g.xa[i]);
            real_t tmp = spd / wid;
            spd = mfabs(c.v - rcyl * P.frame_omega - rot) + cs;
            if (P.dust_flag) spd = mfmax(mfabs(g.CD[id].v - rcyl * P.frame_omega - rot), spd);
            wid = mfabs(g.ya[j + 1] - g.ya[j]) * rcyl;
            tmp = mfmax(spd / wid, tmp);
            if (g.sph()) {   // polar term (timestep.cu:80-88)
                spd = mfabs(c.w) + cs;
                wid = mfabs(g.za[k + 1] - g.za[k]) * rad;
                tmp = mfmax(spd / wid, tmp);
            }
            best = std::fmax(best, val(tmp));
        }
    }
    return best;
}

// -------------------------------------------------------------------- DS
void ds_step(orc_grid& g, real_t time, real_t dt) {  // solve.cu:212-277
    const pg_params& P = g.P;
    real_t hdt = 0.5 * dt;
    if (P.visc_flag) viscosity_eval(g, false);
    fill_x_ghosts(g, g.C);
    if (P.dust_flag) fill_x_ghosts_dust(g, g.CD);
    source_terms(g, g.C, g.CD, g.C, g.CD, 0.0, hdt);
    fill_x_ghosts(g, g.C);
    sweep_x_gas(g, dt);
    fill_y_ghosts(g, g.C);
    sweep_y_gas(g, dt);
    if (P.ndim > 2) {
        fill_z_ghosts(g, g.C);
        sweep_z_gas(g, dt);
    }
    if (P.dust_flag) {
        fill_x_ghosts_dust(g, g.CD);
        sweep_x_dust(g, dt);
        fill_y_ghosts(g, g.CD);
        sweep_y_dust(g, dt);
    }
    if (P.OrbAdv_flag) {
        set_orbital_advection(g, dt);
        shift_orbital_advection(g);
        advect_y_gas(g, dt);
        if (P.dust_flag) advect_y_dust(g, dt);
    }
    orbit_solution(P, g.planet, time + dt);
    if (P.visc_flag) {
        source_terms(g, g.C, g.CD, g.T, g.TD, hdt, hdt);
        viscosity_eval(g, true);
        source_terms(g, g.C, g.CD, g.C, g.CD, hdt, hdt);
    } else {
        source_terms(g, g.C, g.CD, g.C, g.CD, hdt, hdt);
    }
}

}  // namespace

// =================================================================== C ABI
extern "C" {

orc_grid* orc_create(const pg_params* Pin) {
    orc_grid* g = new orc_grid();
    g->P = *Pin;
    const pg_params& P = g->P;
    g->xres = P.xres; g->yres = P.yres;
    g->xpad = kPad; g->ypad = kPad;
    g->xarr = g->xres + 2 * g->xpad;
    g->yarr = g->yres + 2 * g->ypad;
    build_axis(g->xa, g->xv, P.xmin, P.xmax, P.xres, g->xpad, P.geomx, P.gridx);
    build_axis(g->ya, g->yv, P.ymin, P.ymax, P.yres, g->ypad, P.geomy, P.gridy);
    if (P.ndim == 3) {
        g->zres = P.zres; g->zpad = kPad; g->zarr = g->zres + 2 * g->zpad;
        build_axis(g->za, g->zv, P.zmin, P.zmax, P.zres, g->zpad, P.geomz, P.gridz);
    } else {
        build_axis(g->za, g->zv, P.zmin, P.zmax, 1, 0, P.geomz, P.gridz);
    }
    size_t N = (size_t)g->xarr * g->yarr * g->zarr;
    g->C.resize(N); g->T.resize(N); g->CD.resize(N); g->TD.resize(N);
    g->vis.assign(2 * N, 0.0);
    // fill_grid (init.cu:502-519): cell-centre values everywhere
    for (int k = 0; k < g->zarr; ++k)
    for (int j = 0; j < g->yarr; ++j)
        for (int i = 0; i < g->xarr; ++i) {
            g->C[g->ind(i, j, k)] = g->init_at(i, k);
            g->CD[g->ind(i, j, k)] = P.dust_flag ? dust_init(P, g->xc(i)) : Dust{0, 0, 0, 0};
        }
    // seeded perturbation of the interior (not in the reference); 3D keys
    // the azimuthal index with j + yres k
    for (int k = 0; k < g->zres; ++k)
    for (int j = 0; j < P.yres; ++j)
        for (int i = 0; i < P.xres; ++i) {
            Cell& c = g->C[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
            real_t x = g->xc(i + g->xpad);
            const int jk = j + P.yres * k;
            const real_t cs2 = g->sph() ? cs2_3(P, x, g->zc(k + g->zpad)) : cs2_at(P, x);
            if (P.noise_rho != 0.0) {
                c.r *= 1.0 + P.noise_rho * (2.0 * keyed_uniform(P.seed, 0, i, jk) - 1.0);
                if (P.EOS_flag == 0) c.p = cs2 * c.r;
            }
            if (P.noise_vr != 0.0)
                c.u += P.noise_vr * msqrt(P.gam * cs2) * (2.0 * keyed_uniform(P.seed, 1, i, jk) - 1.0);
        }
    g->T = g->C;
    g->TD = g->CD;
    // init_planet (planet.cu:319-403), n_planet 1; softening (planet.cu:308-317):
    // rs_fac h0 a in 2D, rs_fac (q/3)^(1/3) a in 3D
    g->planet = Body{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    g->planet.a = P.planet_radius;
    g->planet.e = P.planet_ecc;
    orbit_solution(P, g->planet, 0.0);
    g->planet.rs = P.ndim == 3 ? P.rs_fac * mpow(P.planet_mass / 3.0, 1.0 / 3.0) * P.planet_radius
                               : P.rs_fac * P.sc_h * P.planet_radius;
    // init_orbital_advection (orbit.cu:51-72), per (i, k)
    g->rot.assign((size_t)g->xarr * g->zarr, 0.0);
    g->res.assign((size_t)g->xarr * g->zarr, 0.0);
    g->inc.assign((size_t)g->xarr * g->zarr, 0);
    for (int k = 0; k < g->zarr; ++k)
    for (int i = 0; i < g->xarr; ++i) {
        real_t rad = g->xc(i);
        real_t v = g->sph() ? v_3(P, rad, g->zc(k)) : v_init(P, rad);
        g->rot[g->ik(i, k)] = P.OrbAdv_flag ? v - g->rcyl(i, k) * P.frame_omega : 0.0;
    }
    return g;
}

void orc_destroy(orc_grid* g) { delete g; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_get_state(const orc_grid* g, double* gas, double* dust) {
    size_t N = (size_t)g->xres * g->yres * g->zres;
    for (int k = 0; k < g->zres; ++k)
    for (int j = 0; j < g->yres; ++j)
        for (int i = 0; i < g->xres; ++i) {
            size_t o = (size_t)i + (size_t)g->xres * (j + (size_t)g->yres * k);
            const Cell& c = g->C[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
            if (gas) {
                gas[o] = val(c.r); gas[N + o] = val(c.p); gas[2 * N + o] = val(c.u); gas[3 * N + o] = val(c.v);
                gas[4 * N + o] = val(c.w);
            }
            if (dust) {
                const Dust& d = g->CD[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
                dust[o] = val(d.r); dust[N + o] = val(d.u); dust[2 * N + o] = val(d.v); dust[3 * N + o] = val(d.w);
            }
        }
}

void orc_set_state(orc_grid* g, const double* gas, const double* dust) {
    size_t N = (size_t)g->xres * g->yres * g->zres;
    for (int k = 0; k < g->zres; ++k)
    for (int j = 0; j < g->yres; ++j)
        for (int i = 0; i < g->xres; ++i) {
            size_t o = (size_t)i + (size_t)g->xres * (j + (size_t)g->yres * k);
            if (gas) {
                Cell& c = g->C[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
                c.r = gas[o]; c.p = gas[N + o]; c.u = gas[2 * N + o]; c.v = gas[3 * N + o]; c.w = gas[4 * N + o];
            }
            if (dust) {
                Dust& d = g->CD[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
                d.r = dust[o]; d.u = dust[N + o]; d.v = dust[2 * N + o]; d.w = dust[3 * N + o];
            }
        }
}

static void copy_out(double* dst, const std::vector<real_t>& src) {
    for (size_t k = 0; k < src.size(); ++k) dst[k] = val(src[k]);
}

void orc_get_coords(const orc_grid* g, double* xa, double* xv, double* ya, double* yv) {
    if (xa) copy_out(xa, g->xa);
    if (xv) copy_out(xv, g->xv);
    if (ya) copy_out(ya, g->ya);
    if (yv) copy_out(yv, g->yv);
}
void orc_get_zcoords(const orc_grid* g, double* za, double* zv) {
    if (za) copy_out(za, g->za);
    if (zv) copy_out(zv, g->zv);
}

void orc_get_planet(const orc_grid* g, double* b) {
    const Body& p = g->planet;
    const real_t f[10] = {p.m, p.a, p.e, p.x, p.y, p.z, p.vx, p.vy, p.vz, p.rs};
    for (int k = 0; k < 10; ++k) b[k] = val(f[k]);
}

void orc_get_rot(const orc_grid* g, double* rot) { copy_out(rot, g->rot); }

double orc_global_dt(orc_grid* g, double dt_prev) {  // timestep.cu:99-140
    double dt = g->P.CFL / val(cfl_rate(*g));
    double cap = dt_prev * 1.1;
    return dt < cap ? dt : cap;
}

void orc_solve(orc_grid* g, double t, double dt) {  // solve.cu:279-292
    ds_step(*g, t, dt);
    if (g->P.kill_flag) killwave(*g, dt);
}

// add_grid (output.cu:250-273): A += C * fac over interior cells
static void add_grid(orc_grid& g, double fac) {
    for (int k = g.k0(); k < g.k1(); ++k)
    for (int j = g.ypad; j < g.yarr - g.ypad; ++j)
        for (int i = g.xpad; i < g.xarr - g.xpad; ++i) {
            const Cell& c = g.C[g.ind(i, j, k)];
            Cell& a = g.A[g.ind(i, j, k)];
            a.r += c.r * fac;
            a.p += c.p * fac;
            a.u += c.u * fac;
            a.v += c.v * fac;
            a.w += c.w * fac;
        }
}

void orc_run(orc_grid* g, int nsteps, double* t_io, double* dt_io, double* dts) {  // main.cu:320-359
    double t = *t_io, dt = *dt_io;
    for (int n = 0; n < nsteps; ++n) {
        dt = orc_global_dt(g, dt);
        if (t + dt > g->P.end_time) dt = g->P.end_time - t;
        if (g->ave_interval > 0.0) add_grid(*g, 0.5 * dt / val(g->ave_interval));  // averaging, main.cu:344
        orc_solve(g, t, dt);
        if (g->ave_interval > 0.0) add_grid(*g, 0.5 * dt / val(g->ave_interval));  // main.cu:350
        t += dt;
        if (dts) dts[n] = dt;
    }
    *t_io = t;
    *dt_io = dt;
}

void orc_viscosity(orc_grid* g, int on_T, double* vis) {
    viscosity_eval(*g, on_T != 0);
    if (!vis) return;
    size_t N = (size_t)g->xres * g->yres * g->zres;
    for (int k = 0; k < g->zres; ++k)
    for (int j = 0; j < g->yres; ++j)
        for (int i = 0; i < g->xres; ++i) {
            size_t id = g->ind(i + g->xpad, j + g->ypad, k + g->zpad);
            size_t o = (size_t)i + (size_t)g->xres * (j + (size_t)g->yres * k);
            vis[o] = val(g->vis[2 * id]);
            vis[N + o] = val(g->vis[2 * id + 1]);
        }
}

void orc_sources(orc_grid* g, double mdt, double dt, int inplace) {
    if (inplace) source_terms(*g, g->C, g->CD, g->C, g->CD, mdt, dt);
    else source_terms(*g, g->C, g->CD, g->T, g->TD, mdt, dt);
}

void orc_sweepx(orc_grid* g, double dt) { fill_x_ghosts(*g, g->C); sweep_x_gas(*g, dt); }
void orc_sweepz(orc_grid* g, double dt) { fill_z_ghosts(*g, g->C); sweep_z_gas(*g, dt); }
void orc_sweepy(orc_grid* g, double t, double dt) { (void)t; fill_y_ghosts(*g, g->C); sweep_y_gas(*g, dt); }
void orc_sweepx_dust(orc_grid* g, double dt) { fill_x_ghosts_dust(*g, g->CD); sweep_x_dust(*g, dt); }
void orc_sweepy_dust(orc_grid* g, double t, double dt) { (void)t; fill_y_ghosts(*g, g->CD); sweep_y_dust(*g, dt); }

void orc_orbital_advection(orc_grid* g, double dt) {
    set_orbital_advection(*g, dt);
    shift_orbital_advection(*g);
    advect_y_gas(*g, dt);
    if (g->P.dust_flag) advect_y_dust(*g, dt);
}

void orc_get_orbadv(const orc_grid* g, int* inc, double* res) {
    for (int k = 0; k < g->zres; ++k)
    for (int i = 0; i < g->xres; ++i) {
        const size_t o = (size_t)i + (size_t)g->xres * k;
        if (inc) inc[o] = g->inc[g->ik(i + g->xpad, k + g->zpad)];
        if (res) res[o] = val(g->res[g->ik(i + g->xpad, k + g->zpad)]);
    }
}

void orc_killwave(orc_grid* g, double dt) { killwave(*g, dt); }

// write_check_point (output.cu:125-180) for ndim 2, with restructure_data's
// field order (output.cu:95-123); the dust w field is written too (ndim > 1
// guard of output.cu:171)
void orc_bound(orc_grid* g, int dir, double* gas, double* dust) {
    orc_grid& G = *g;
    if (dir == 0) {
        fill_x_ghosts(G, G.C);
        if (dust && G.P.dust_flag) fill_x_ghosts_dust(G, G.CD);
    } else if (dir == 1) {
        fill_y_ghosts(G, G.C);
        if (dust && G.P.dust_flag) fill_y_ghosts(G, G.CD);
    } else {
        fill_z_ghosts(G, G.C);
    }
    // the block: every cell of the padded direction, the interior of the others
    const int ni = dir == 0 ? G.xarr : G.xres, nj = dir == 1 ? G.yarr : G.yres, nk = dir == 2 ? G.zarr : G.zres;
    const int i0 = dir == 0 ? 0 : G.xpad, j0 = dir == 1 ? 0 : G.ypad, k0 = dir == 2 ? 0 : G.zpad;
    const size_t N = (size_t)ni * nj * nk;
    for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const size_t o = (size_t)i + (size_t)ni * (j + (size_t)nj * k);
            const size_t c = G.ind(i + i0, j + j0, k + k0);
            if (gas) {
                const Cell& C = G.C[c];
                gas[o] = val(C.r); gas[N + o] = val(C.p); gas[2 * N + o] = val(C.u); gas[3 * N + o] = val(C.v);
                gas[4 * N + o] = val(C.w);
            }
            if (dust && G.P.dust_flag) {
                const Dust& D = G.CD[c];
                dust[o] = val(D.r); dust[N + o] = val(D.u); dust[2 * N + o] = val(D.v); dust[3 * N + o] = val(D.w);
            }
        }
}

int orc_write_snapshot(const orc_grid* g, const char* filename, double t) {
    FILE* f = std::fopen(filename, "wb");
    if (!f) return -1;
    std::fwrite(&t, sizeof(double), 1, f);
    std::vector<double> ax(g->xres + 1), ay(g->yres + 1), az(g->zres + 1);
    for (int i = 0; i <= g->xres; ++i) ax[i] = val(g->xa[g->xpad + i]);
    for (int j = 0; j <= g->yres; ++j) ay[j] = val(g->ya[g->ypad + j]);
    for (int k = 0; k <= g->zres; ++k) az[k] = val(g->za[g->zpad + k]);
    std::fwrite(ax.data(), sizeof(double), ax.size(), f);
    std::fwrite(ay.data(), sizeof(double), ay.size(), f);
    if (g->sph()) std::fwrite(az.data(), sizeof(double), az.size(), f);
    std::vector<double> tmp((size_t)g->xres * g->yres * g->zres);
    // gas rho, p, u, v (+ w in 3D), then dust rho, u, v, w
    const int ng = g->sph() ? 5 : 4;
    int nf = ng + (g->P.dust_flag ? 4 : 0);
    for (int field = 0; field < nf; ++field) {
        for (int k = 0; k < g->zres; ++k)
        for (int j = 0; j < g->yres; ++j)
            for (int i = 0; i < g->xres; ++i) {
                size_t id = g->ind(i + g->xpad, j + g->ypad, k + g->zpad);
                const Cell& c = g->C[id];
                const Dust& d = g->CD[id];
                real_t v = 0.0;
                const int fk = field < ng ? field : 5 + (field - ng);
                switch (fk) {
                    case 0: v = c.r; break;
                    case 1: v = c.p; break;
                    case 2: v = c.u; break;
                    case 3: v = c.v; break;
                    case 4: v = c.w; break;
                    case 5: v = d.r; break;
                    case 6: v = d.u; break;
                    case 7: v = d.v; break;
                    default: v = d.w; break;
                }
                tmp[(size_t)i + (size_t)g->xres * (j + (size_t)g->yres * k)] = val(v);
            }
        std::fwrite(tmp.data(), sizeof(double), tmp.size(), f);
    }
    std::fclose(f);
    return 0;
}
void orc_evolve_planet(orc_grid* g, double t) { orbit_solution(g->P, g->planet, t); }

// init_average (output.cu:316-329) and the averaging switch of main.cu
void orc_average_begin(orc_grid* g, double interval) {
    g->A.assign(g->C.size(), Cell{0, 0, 0, 0, 0});
    g->ave_interval = interval;
}
void orc_average_end(orc_grid* g) { g->ave_interval = 0.0; }
// the average over interior cells, fields rho, p, u, v, w (x fastest)
void orc_average_get(const orc_grid* g, double* gas) {
    const size_t N = (size_t)g->xres * g->yres * g->zres;
    for (int k = 0; k < g->zres; ++k)
    for (int j = 0; j < g->yres; ++j)
        for (int i = 0; i < g->xres; ++i) {
            const Cell& a = g->A[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
            const size_t o = (size_t)i + (size_t)g->xres * (j + (size_t)g->yres * k);
            gas[o] = val(a.r);
            gas[N + o] = val(a.p);
            gas[2 * N + o] = val(a.u);
            gas[3 * N + o] = val(a.v);
            gas[4 * N + o] = val(a.w);
        }
}

// global_sum (output.cu:336-457) for one planet in 2D cylindrical geometry:
// the sum over interior cells of R gy rho dV (cal_torque), gy from
// o_star_planet_grav_cyl at dt = 0 (output/forces.cu:41-58, 106-127), which
// is the solver's star_planet_grav_cyl; summed in cell order
double orc_torque(const orc_grid* g) {
    real_t sum = 0.0;
    for (int j = g->ypad; j < g->yarr - g->ypad; ++j)
        for (int i = g->xpad; i < g->xarr - g->xpad; ++i) {
            const real_t rad = 0.5 * (g->xa[i + 1] + g->xa[i]);
            const real_t azi = 0.5 * (g->ya[j + 1] + g->ya[j]);
            real_t dv = 0.5 * (g->xa[i + 1] * g->xa[i + 1] - g->xa[i] * g->xa[i]);
            dv *= (g->ya[j + 1] - g->ya[j]);
            real_t fx, fy, fz;
            planet_force_cyl(rad, azi, 0.0, g->planet, fx, fy, fz);
            sum += rad * fy * (g->C[g->ind(i, j)].r * dv);
        }
    return val(sum);
}

#ifdef ORC_TANGENT
// tangent-linear build: the derivative parts of the interior state (fields
// as orc_get_state); orc_set_state leaves them zero
void orc_set_tangent(orc_grid* g, const double* gas, const double* dust) {
    size_t N = (size_t)g->xres * g->yres * g->zres;
    for (int k = 0; k < g->zres; ++k)
    for (int j = 0; j < g->yres; ++j)
        for (int i = 0; i < g->xres; ++i) {
            size_t o = (size_t)i + (size_t)g->xres * (j + (size_t)g->yres * k);
            if (gas) {
                Cell& c = g->C[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
                c.r.d = gas[o]; c.p.d = gas[N + o]; c.u.d = gas[2 * N + o]; c.v.d = gas[3 * N + o];
                c.w.d = gas[4 * N + o];
            }
            if (dust) {
                Dust& d = g->CD[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
                d.r.d = dust[o]; d.u.d = dust[N + o]; d.v.d = dust[2 * N + o]; d.w.d = dust[3 * N + o];
            }
        }
}
void orc_get_tangent(const orc_grid* g, double* gas, double* dust) {
    size_t N = (size_t)g->xres * g->yres * g->zres;
    for (int k = 0; k < g->zres; ++k)
    for (int j = 0; j < g->yres; ++j)
        for (int i = 0; i < g->xres; ++i) {
            size_t o = (size_t)i + (size_t)g->xres * (j + (size_t)g->yres * k);
            const Cell& c = g->C[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
            if (gas) {
                gas[o] = c.r.d; gas[N + o] = c.p.d; gas[2 * N + o] = c.u.d; gas[3 * N + o] = c.v.d;
                gas[4 * N + o] = c.w.d;
            }
            if (dust) {
                const Dust& d = g->CD[g->ind(i + g->xpad, j + g->ypad, k + g->zpad)];
                dust[o] = d.r.d; dust[N + o] = d.u.d; dust[2 * N + o] = d.v.d; dust[3 * N + o] = d.w.d;
            }
        }
}
#endif

// ------------------------------------------------------------ unit level
#ifndef ORC_TANGENT
void orc_fourth_parameters(int geom, const double* x6, const double* a5, double* par3) {
    fourth_parameters(2, geom, x6, a5, par3);
}
void orc_ppm_adjust(double flat, double* par3) { ppm_adjust(flat, par3); }
double orc_ppm_aveL(double q, const double* par3) { return ppm_aveL(q, par3); }
double orc_ppm_aveR(double q, const double* par3) { return ppm_aveR(q, par3); }
double orc_pem_aveL(double q, const double* par3) { return pem_aveL(q, par3); }
double orc_pem_aveR(double q, const double* par3) { return pem_aveR(q, par3); }
double orc_flatten(double gam, const double* r5, const double* p5, const double* u5) {
    return flatten_cell(2, 5, gam, r5, p5, u5);
}
void orc_set_state_row(int geom, double gam, int n, const double* xa, const double* dv,
                       const double* r, const double* p, const double* u, const double* v,
                       const double* w, const double* flat, int i, double dt, double* S10) {
    (void)n; (void)dv;
    State S;
    trace_state(i, geom, 6, gam, xa, r, p, u, v, w, flat, dt, 0.0, S);
    std::memcpy(S10, &S, sizeof(State));
}
void orc_riemann_hllc(double gam, int eos, const double* S10, double* out) {
    State S;
    std::memcpy(&S, S10, sizeof(State));
    double pm, sl, sm, sr, pres, uprs;
    wave_speeds(S, gam, 0.0, pm, sl, sm, sr);
    Cell F;
    hllc(S, eos, gam, pm, sl, sm, sr, F, pres, uprs);
    double o[10] = {pm, sl, sm, sr, F.r, F.p, F.u, F.v, F.w, pres};
    std::memcpy(out, o, sizeof(o));
}
void orc_riemann_row(int geom, double gam, int n, const double* xa, const double* dv,
                     const double* r, const double* p, const double* u, const double* v,
                     const double* w, double dt, double* Del5) {
    (void)dv;
    RowScratch s;
    s.resize(n);
    std::memcpy(s.r.data(), r, n * sizeof(double));
    std::memcpy(s.p.data(), p, n * sizeof(double));
    std::memcpy(s.u.data(), u, n * sizeof(double));
    std::memcpy(s.v.data(), v, n * sizeof(double));
    std::memcpy(s.w.data(), w, n * sizeof(double));
    for (int i = 0; i < n; ++i) s.Del[i] = Cell{0, 0, 0, 0, 0};
    riemann_row(geom, 0, gam, n, xa, s, dt);
    for (int i = 0; i < n; ++i) {
        Del5[i] = s.Del[i].r; Del5[n + i] = s.Del[i].p; Del5[2 * n + i] = s.Del[i].u;
        Del5[3 * n + i] = s.Del[i].v; Del5[4 * n + i] = s.Del[i].w;
    }
}
void orc_dust_flux(int geom, int n, const double* xa, const double* dv, const double* r,
                   const double* u, const double* v, const double* w, int i, double dt, double* F4) {
    (void)dv;
    std::vector<double> flat(n, 0.0);
    Dust F = dust_interface(i, geom, xa, r, u, v, w, flat.data(), dt, 0.0);
    F4[0] = F.r; F4[1] = F.u; F4[2] = F.v; F4[3] = F.w;
}
void orc_orbit_solution(const pg_params* P, double t, double* b10) {
    Body b{0, P->planet_radius, P->planet_ecc, 0, 0, 0, 0, 0, 0, P->rs_fac * P->sc_h * P->planet_radius};
    orbit_solution(*P, b, t);
    std::memcpy(b10, &b, sizeof(Body));
}
void orc_riemann_hllc_e(double gam, int eos, const double* S10, double* out) {
    State S;
    std::memcpy(&S, S10, sizeof(State));
    double pm, sl, sm, sr, pres, uprs;
    wave_speeds(S, gam, 0.0, pm, sl, sm, sr);
    Cell F;
    hllc(S, eos, gam, pm, sl, sm, sr, F, pres, uprs);
    double o[11] = {pm, sl, sm, sr, F.r, F.p, F.u, F.v, F.w, pres, uprs};
    std::memcpy(out, o, sizeof(o));
}
void orc_planet_force(int sph, const double* b10, double rad, double azi, double pol, double* f3) {
    Body b;
    std::memcpy(&b, b10, sizeof(Body));
    double fx, fy, fz;
    if (sph) planet_force_sph(rad, azi, pol, b, fx, fy, fz);
    else planet_force_cyl(rad, azi, pol, b, fx, fy, fz);
    f3[0] = fx;
    f3[1] = fy;
    f3[2] = fz;
}

#endif  // ORC_TANGENT

}  // extern "C"
