// pg_sph.hip -- the three-dimensional disk: spherical radius r (geomx 2),
// azimuth phi (geomy 4), polar angle theta (geomz 5); adiabatic gas with the
// internal-energy equation (EOS_flag 2, internal_e_flag 1), PEM4 or PPM4
// reconstruction, HLLC; orbital advection per (radius, polar layer);
// viscosity on every polar layer; spherical forces with the spline-softened
// planet.  Reference: src/solver/sweeps.cu:218-553 (update_Cell and the three
// sweeps), src/solver/riemann/*.cu, src/solver/recon/PEM.cu, recon.cu,
// src/orbit/orbit.cu:17-99, src/solver/advection/advection.cu:1-173,
// src/viscosity/viscosity.cu:34-197, src/solver/force/forces.cu:14-327,
// src/timestep/timestep.cu:18-97, src/boundary/boundary.cu:270-321,
// src/init/init.cu:7-349 (init_flag 3), src/solver/solve.cu:212-292.
//
// Device layout: five fields rho, p, u (radial), vp, w (polar) in SoA, interior
// cells only, radius fastest: field f of cell (i, j, k) at
// f*N + i + nx*(j + ny*k).  vp = v_phi - vbg(r, theta), vbg the initial
// rotation (= orbital-advection speed + R_cyl Omega_frame), so the azimuthal
// sweep's frame velocity is vp itself.  Ghosts are never stored: radial ghosts
// are the analytic initial state (boundary type 0) from tables, azimuthal
// ghosts periodic indexing, polar ghosts mirror images with w negated
// (boundary type 2).
//
// Sweeps stage a tile of R rows x S cells (S = segment + 6 ghost or halo
// cells) in LDS; thread t owns elements e = t, t + 256, ... with e = s*R + r,
// so consecutive threads hold consecutive radii of the strided sweeps (the
// coalesced direction of the global loads) and the row neighbours of an
// element sit R apart in LDS.  Phases: load -> per cell: reconstruction and
// the two characteristic-traced states -> per interface: HLLC -> per cell:
// flux and pressure-source differences, the conservative update.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/penguin_amd.h"
#include "pg_device.h"
#include "pg_internal.h"
#include "pg_sph.h"

namespace pg {

namespace {

constexpr int NF = 5;            // rho, p, u, vp, w
constexpr int SPH_BT = 256;      // threads per block of the pointwise kernels and the radial sweep
constexpr int SPH_BT2 = 512;     // threads per block of the strided sweeps (about one element each)
constexpr int SPH_BTX = 256;     // most threads per radial-sweep block (one per LDS slot)
constexpr int SPH_R = 8;         // rows per strided-sweep block
// azimuthal segment per block: with its 6 halo rows, 8 radii x 64 slots fill
// a 512-thread block (and two passes of a 256-thread one) exactly (64 rows:
// 560 slots, a partly filled extra pass; 1091 -> 1050 us for the azimuthal
// sweep, 591 -> 555 us for the orbital advection of the 384x1536x64 disk)
constexpr int SPH_TY = 58;
constexpr int SPH_ZMAX = 100;    // polar cells: the whole column is one tile (FP64: 102 KB of LDS)

template <typename real>
struct SphK {
    int nx, ny, nz;
    size_t N;
    const real* in;
    real* out;
    const real* vis;
    real* vis_out;
    // per padded radius ip = i + 3: edges (nx + 7), volumes, centres,
    // slope weights [6][nx + 6], sin-free squares of the edges (gfac)
    const real* xa;
    const real* xv;
    const real* xc;
    const real* xw;
    // per padded polar index kp = k + 3: edges, volumes, centres, weights,
    // sin of the edges (gfac), sin and cot of the centres
    const real* za;
    const real* zv;
    const real* zc;
    const real* zw;
    const real* sza;
    const real* sinz;
    const real* cotz;
    // per (padded radius, interior polar layer): background rotation, nu,
    // and the radial force balance of the background rotation
    // ebal = vbg^2 / r - 1 / r^2 (formed in FP64)
    const real* vbg;
    const real* nu;
    const real* ebal;
    // radial ghosts: [f][g][k], g = 0..2 (ip = 0..2), 3..5 (ip = nx+3..nx+5)
    const real* xg;
    // per azimuthal row j: sin and cos of phi_j / 2 (phi_j the row centre)
    const real* hphi;
    // per interior (i, k), FP64: orbital-advection speed and R_cyl
    const double* rot;
    const double* rcyl;
    real* rate;
    real yw[6];
    real dy, idy;
    double dy_d;
    real gam;
    int tx;   // radial segment of the radial sweep (sph_tx)
};

// Radial segment of the radial sweep: the row cut into the fewest equal
// parts whose slots (segment + 2 NPAD) fit one block of SPH_BTX threads, so
// every thread holds one slot and no block runs half empty (a 384-radius row
// is two 198-slot blocks of 256 threads: a CU holds more of them than of one
// 390-slot block of 448 threads, and one block's loads hide behind the
// others' arithmetic: 1343 -> 1219 us on the 384x1536x64 disk; 128-thread
// blocks of 102 slots 1240-1256 us).
inline int sph_tx(int nx) {
    const int parts = (nx + SPH_BTX - 2 * NPAD - 1) / (SPH_BTX - 2 * NPAD);
    return (nx + parts - 1) / parts;
}
inline int sph_bx(int nx) { return (sph_tx(nx) + 2 * NPAD + 63) / 64 * 64; }

template <typename real>
struct PlanetS {   // planet for the spherical force: mass, radius, angle, softening
    real m, x, y, rs, shy, chy;
};

// ----------------------------------------------------- reconstruction
// __logf of a float argument (PEM.cu:13,29): FP32 takes the hardware
// logarithm; FP64 takes the float logarithm correctly rounded (the double
// logarithm of the float argument, rounded to float), as the oracle does
template <typename real> __device__ __forceinline__ real flog(real x);
template <> __device__ __forceinline__ float flog<float>(float x) { return __logf(x); }
template <> __device__ __forceinline__ double flog<double>(double x) { return (double)(float)log((double)(float)x); }

// Parameters of a cell's sub-cell profile: slopes, mean, and (PEM4) which
// form each of get_CON_aveL / aveR takes (PEM.cu:51-93), decided once per
// cell and field since it depends on the slopes only: k = 0 the mean,
// 1 the parabola, 2 get_PEM_0, 3 get_PEM_1, with exponents eL, eR.
template <typename real>
struct Par {
    real sL, a, sR;
    real eL, eR;
    int kL, kR;
};

// the PEM4 branch choices of get_PEM_aveL / aveR (PEM.cu:51-93)
template <typename real>
__device__ __forceinline__ void pem_select(Par<real>& p) {
    if (p.sL * p.sR == real(0)) {
        p.kL = p.kR = 0;
        p.eL = p.eR = real(0);
        return;
    }
    const real rr = FM<real>::div(p.sR, p.sL), rl = FM<real>::div(p.sL, p.sR);
    p.kL = rr >= real(1) ? (rr <= real(2) ? 1 : 2) : (rl <= real(2) ? 1 : 3);
    p.eL = rr >= real(1) ? rr : rl;
    p.kR = rr > real(1) ? (rr <= real(2) ? 1 : 3) : (rl <= real(2) ? 1 : 2);
    p.eR = rr > real(1) ? rr : rl;
}

// get_PEM_0 / get_PEM_1 (PEM.cu:3-48) through one logarithm and one
// exponential: with y = x (PEM_0) or 1 - x (PEM_1), E = y^eta; PEM_0 is E,
// PEM_1 is (1 - x)(E - 1)/(eta x) + 1 or its series below eta x = 0.001 (so
// lanes of a wave that take different forms share the transcendental work).
// get_PEM_1's x >= 1 assignment is overwritten by the branch after it
// (PEM.cu:23-27) and has no counterpart.
template <typename real>
__device__ __forceinline__ real pem01(bool zero_form, real x, real S, real aB, real eta) {
    const real y = zero_form ? x : real(1) - x;
    const real E = y == real(0) ? real(0) : FM<real>::exp(eta * flog(y));
    real v;
    if (zero_form) {
        v = E;
    } else if (eta * x > real(0.001)) {
        v = FM<real>::div(y * (E - real(1)), eta * x) + real(1);
    } else {
        real t = (eta + real(1)) * x / real(2);
        v = t;
        t *= -(eta - real(1)) * x / real(3);
        v += t;
        t *= -(eta - real(2)) * x / real(4);
        v += t;
        t *= -(eta - real(3)) * x / real(5);
        v += t;
    }
    return aB + S * v;
}

// get_CON_aveL (recon.cu:424-441): mean over the first fraction q
template <int RECON, typename real>
__device__ __forceinline__ real avgL(real q, const Par<real>& p) {
    if (RECON == 6 || p.kL == 1) return aveL(q, p.sL, p.a, p.sR);
    if (p.kL == 0) return p.a;
    return pem01(p.kL == 2, q, p.sL, p.a - p.sL, p.eL);
}
// get_CON_aveR (recon.cu:405-422): mean over the cell from q to its right edge
template <int RECON, typename real>
__device__ __forceinline__ real avgR(real q, const Par<real>& p) {
    if (RECON == 6 || p.kR == 1) return aveR(real(1) - q, p.sL, p.a, p.sR);
    if (p.kR == 0) return p.a;
    return pem01(p.kR == 2, real(1) - q, -p.sR, p.a + p.sR, p.eR);
}

// 4th-order limited slopes (recon.cu:216-378); under RECON 6 also the
// flattening and the PPM monotonicity fix (PPM.cu:64-91)
template <int RECON, typename real>
__device__ __forceinline__ Par<real> params(real am2, real am1, real a0, real ap1, real ap2, const real* w, real flat) {
    Par<real> p;
    p.a = a0;
    if (RECON == 6) {
        slopes(am2, am1, a0, ap1, ap2, w[0], w[1], w[2], w[3], w[4], w[5], flat, p.sL, p.sR);
        p.kL = p.kR = 1;
        p.eL = p.eR = real(0);
        return p;
    }
    real d0 = am1 - am2, d1 = a0 - am1, d2 = ap1 - a0, d3 = ap2 - ap1;
    real l = w[0] * d0 + w[1] * d1 + w[2] * d2;
    real r = w[3] * d1 + w[4] * d2 + w[5] * d3;
    l = copysign(fmin(fabs(l), fabs(d1)), d1);
    r = copysign(fmin(fabs(r), fabs(d2)), d2);
    if (r * l <= real(0)) {
        l = real(0);
        r = real(0);
    }
    p.sL = l;
    p.sR = r;
    pem_select(p);
    return p;
}
// the same parameters about a zero mean: avgL / avgR of them is the sub-cell
// mean less the cell's own (both are affine in the mean with unit weight)
template <typename real>
__device__ __forceinline__ Par<real> zero_mean(Par<real> p) {
    p.a = real(0);
    return p;
}

// set_state (states.cu:1-135) for an adiabatic gas: the state right of the
// cell's left edge (from the cell's means over its first fractions, R side)
// and left of its right edge (L side).  P: parameters of rho, p, u, v, w;
// rc, pc, uc: the cell's values; dts: the sweep's time step over its length
// unit (dt / rad); idx = 1 / width.
// The traced pressure S[1] is returned as its deviation from the edge's
// reference pressure pref (the mean of the two cells' pressures), given the
// cell's own deviation pdev = p - pref: the pressure's sub-cell means enter
// as their differences from the cell's pressure, so the pressure
// differences of the update are formed from numbers of their own size.
// S[1] is p0 exp_lim(gam u_) - pref = (p0 - pref) + p0 (exp_lim(gam u_) - 1).
template <typename real>
__device__ __forceinline__ real exp_lim_m1(real x) {   // exp_lim(x) - 1
    return fmax(x, real(1.0e-6) - real(1));
}
template <int RECON, typename real>
__device__ __forceinline__ void trace_R(const Par<real>* P, real rc, real pc, real uc, real pdev, real gam, real dts,
                                        real idx, real* S) {
    const real hdt = real(0.5) * dts;
    real cs = FM<real>::sqrt(FM<real>::div(gam * pc, rc));
    real q = lim01((cs - fmin(uc, real(0))) * dts * idx);
    const real rmin = rc * real(1.0e-10), pmin = pc * real(1.0e-10);
    const real ur = avgL<RECON>(q, P[2]);
    const real dis = fmin(ur, real(0)) * dts;
    q = lim01(-dis * idx);
    const Par<real> Pz = zero_mean(P[1]);
    const real dp0 = avgL<RECON>(q, Pz);
    real r0 = avgL<RECON>(q, P[0]), p0 = pc + dp0, u0 = avgL<RECON>(q, P[2]);
    r0 = fmax(r0, rmin);
    const bool pfl = p0 < pmin;
    p0 = fmax(p0, pmin);
    cs = FM<real>::sqrt(FM<real>::div(gam * p0, r0));
    q = lim01((cs * dts - dis) * idx);
    const real dp_ = avgL<RECON>(q, Pz);
    real u_ = avgL<RECON>(q, P[2]);
    const real f = FM<real>::div(real(2), cs * dts) * (-hdt);
    u_ = (u_ - u0) * f;
    const real p_ = (pfl ? (pc + dp_) - p0 : dp_ - dp0) * f;
    S[0] = r0 * exp_lim(u_);
    S[1] = (pfl ? p0 - (pc - pdev) : pdev + dp0) + p0 * exp_lim_m1(gam * u_);
    S[2] = u0 + FM<real>::div(p_, r0);
    S[3] = avgL<RECON>(q, P[3]);
    S[4] = avgL<RECON>(q, P[4]);
}
template <int RECON, typename real>
__device__ __forceinline__ void trace_L(const Par<real>* P, real rc, real pc, real uc, real pdev, real gam, real dts,
                                        real idx, real* S) {
    const real hdt = real(0.5) * dts;
    real cs = FM<real>::sqrt(FM<real>::div(gam * pc, rc));
    real q = lim01(real(1) - (fmax(uc, real(0)) + cs) * dts * idx);
    const real rmin = rc * real(1.0e-10), pmin = pc * real(1.0e-10);
    const real ul = avgR<RECON>(q, P[2]);
    const real dis = fmax(ul, real(0)) * dts;
    q = lim01(real(1) - dis * idx);
    const Par<real> Pz = zero_mean(P[1]);
    const real dp0 = avgR<RECON>(q, Pz);
    real r0 = avgR<RECON>(q, P[0]), p0 = pc + dp0, u0 = avgR<RECON>(q, P[2]);
    r0 = fmax(r0, rmin);
    const bool pfl = p0 < pmin;
    p0 = fmax(p0, pmin);
    cs = FM<real>::sqrt(FM<real>::div(gam * p0, r0));
    q = lim01(real(1) - (dis + cs * dts) * idx);
    const real dp_ = avgR<RECON>(q, Pz);
    real u_ = avgR<RECON>(q, P[2]);
    const real f = FM<real>::div(real(2), cs * dts) * (-hdt);
    u_ = (u0 - u_) * f;
    const real p_ = (pfl ? p0 - (pc + dp_) : dp0 - dp_) * f;
    S[0] = r0 * exp_lim(u_);
    S[1] = (pfl ? p0 - (pc - pdev) : pdev + dp0) + p0 * exp_lim_m1(gam * u_);
    S[2] = u0 + FM<real>::div(p_, r0);
    S[3] = avgR<RECON>(q, P[3]);
    S[4] = avgR<RECON>(q, P[4]);
}

// wave_speeds (wave_speeds.cu:24-35,288-333) and HLLC_fluxes (HLLC.cu:2-77)
// with EOS_flag 2, internal_e_flag 1: mass flux; internal energy, normal and
// transverse momenta upwinded times the mass flux; interface pressure p* and
// velocity u* (the p dV work, riemann.cu:65-73).  L[1], R[1] and the
// returned p* are deviations from the interface's reference pressure pref:
// the contact speed's pressure term and p* are formed from them, the sound
// speeds and fluxes from the full pressures.
template <typename real>
__device__ __forceinline__ void hllc_adi(const real* L, const real* R, real pref, real gam, real* F, real& pres,
                                         real& uprs) {
    const real rl = L[0], pld = L[1], ul = L[2], rr = R[0], prd = R[1], ur = R[2];
    const real pl = pref + pld, pr = pref + prd;
    const real cl = FM<real>::sqrt(FM<real>::div(gam * pl, rl));
    const real cr = FM<real>::sqrt(FM<real>::div(gam * pr, rr));
    const real c_ = real(0.5) * (cr + cl), p_ = real(0.5) * (pr + pl);
    const real xm = FM<real>::div(real(0.5) * gam * (ul - ur), c_);
    const real pm = p_ * exp_lim(xm);
    const real pmd = real(0.5) * (pld + prd) + p_ * exp_lim_m1(xm);   // pm - pref
    const real srl = FM<real>::sqrt(rl), srr = FM<real>::sqrt(rr);
    const real isum = FM<real>::rcp(srl + srr);
    const real ubar = (srl * ul + srr * ur) * isum;
    const real n2 = real(0.5) * srl * srr * isum * isum;
    const real cbar = FM<real>::sqrt((cl * cl * srl + cr * cr * srr) * isum + n2 * (ur - ul) * (ur - ul));
    const real sl = ubar - cbar, sr = ubar + cbar;
    real sm = FM<real>::div(rl * ul * cl + rr * ur * cr, rl * cl + rr * cr);
    sm += FM<real>::div(pld - prd, (rl + rr) * (real(0.5) * (cl + cr)));
    // pick (HLLC.cu:2-8): sl >= 0 -> left, sr < 0 -> right, else the star
    // state on the side of the contact
    const bool supL = !signbit(sl), supR = signbit(sr), left = supL || (!supR && !signbit(sm));
    const bool star = !supL && !supR;
    const real rK = left ? rl : rr, uK = left ? ul : ur, pK = left ? pl : pr, pKd = left ? pld : prd;
    const real sK = left ? sl : sr;
    const real vK = left ? L[3] : R[3], wK = left ? L[4] : R[4];
    const real fm = star ? FM<real>::div(sm * rK * (sK - uK), sK - sm) : uK * rK;
    F[0] = fm;
    F[1] = FM<real>::div(pK, (gam - real(1)) * rK) * fm;
    F[2] = uK * fm;
    F[3] = vK * fm;
    F[4] = wK * fm;
    pres = star ? FM<real>::div(sK * pmd - sm * pKd, sK - sm) : pKd;
    uprs = star ? sm * FM<real>::div(sK * pm - uK * pK, sK * pm - sm * pK) : uK;
}

// update_Cell (sweeps.cu:218-261), EOS_flag 2, internal_e_flag 1, on the
// stored fields (rho, p, u, vp, w); D: changes of mass, internal energy and
// the momenta of u, vp, w
template <typename real>
__device__ __forceinline__ void update_cell(real* c, const real* D, real gam) {
    const real rmin = fmax(c[0] * real(1.0e-10), real(1.0e-10));
    const real pmin = fmax(c[1] * real(1.0e-10), real(1.0e-10));
    const real E = FM<real>::div(c[1], (gam - real(1)) * c[0]) * c[0];
    const real rn = c[0] + D[0];
    if (rn < rmin) {
        c[0] = rmin;
        c[1] = pmin;
        return;   // velocities kept
    }
    // (rho u + D) / (rho + D0) as u + (D - u D0) / (rho + D0): the rounding
    // falls on the change, not on the velocity itself
    const real ir = FM<real>::rcp(rn);
    c[2] += (D[2] - c[2] * D[0]) * ir;
    c[3] += (D[3] - c[3] * D[0]) * ir;
    c[4] += (D[4] - c[4] * D[0]) * ir;
    c[1] = fmax((E + D[1]) * (gam - real(1)), pmin);
    c[0] = rn;
}

// Workgroups are dealt to the 8 XCDs round robin (block b runs on XCD b % 8),
// each XCD with its own L2.  Renumbering so that XCD x takes a contiguous
// range of tiles keeps tiles that share cache lines (neighbouring radii of
// the strided sweeps) in one L2: a bijection of [0, nb).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, n = b >> 3;
    return x * q + min(x, r) + n;
}

__device__ __forceinline__ int wrapi(int j, int n) {
    j %= n;
    return j < 0 ? j + n : j;
}

template <typename real>
__device__ __forceinline__ void source_cell(const SphK<real>& a, const PlanetS<real>& pl, bool planet, int ip, int kp,
                                            int j, real rho, real u, real vp, real w, real visr, real visp, real mdt,
                                            real dt, real& un, real& vn, real& wn);

// ----------------------------------------------------------------- sweeps
// DIR 0: radius (rows (j, k), segments of a.tx cells, one thread per slot);
// 1: azimuth (R radii x SPH_TY cells per block); 2: polar angle (R radii x
// the whole column).  DIR 0 with hkick != 0 first applies the step's first
// half kick (the source terms at t with mdt = 0 over hkick = dt/2, planet pl)
// to every interior cell it loads, in place of a source-term pass; the block
// has a thread per slot, so each thread keeps its own cell's kicked state in
// registers for the update.
template <typename real, int DIR, int RECON>
__global__ __launch_bounds__(DIR == 0 ? SPH_BTX : SPH_BT2) void k3_sweep(SphK<real> a, real dt, PlanetS<real> pl,
                                                                         int planet, real hkick) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    real* sm = reinterpret_cast<real*>(sm_raw);
    constexpr int R = DIR == 0 ? 1 : SPH_R;
    const int BT = DIR == 0 ? (int)blockDim.x : SPH_BT2;
    const int L = DIR == 0 ? a.tx : DIR == 1 ? SPH_TY : a.nz;
    const int S = L + 2 * NPAD;
    const int RS = R * S;         // at most 2 BT (checked by the host)
    real* raw = sm;              // [5][RS]
    real* TLs = raw + NF * RS;   // state left of each cell's right edge
    real* TRs = TLs + NF * RS;   // state right of each cell's left edge
    // per interface, once the traced states are consumed: flux[5] in TLs,
    // p* and u* in TRs
    const int t = threadIdx.x;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const size_t N = a.N;

    // tile of this block
    int i0 = 0, j0 = 0, k0 = 0;   // first interior cell of the segment / first row
    const int bt = xcd_tile(blockIdx.x, gridDim.x);
    if (DIR == 0) {
        const int ntx = (nx + a.tx - 1) / a.tx;
        const int row = bt / ntx;
        i0 = (bt - row * ntx) * a.tx;
        j0 = row % ny;
        k0 = row / ny;
    } else if (DIR == 1) {
        const int nty = (ny + SPH_TY - 1) / SPH_TY, nig = (nx + R - 1) / R;
        const int rest = bt / nig;
        i0 = (bt - rest * nig) * R;
        j0 = (rest % nty) * SPH_TY;
        k0 = rest / nty;
    } else {
        const int nig = (nx + R - 1) / R;
        i0 = (bt % nig) * R;
        j0 = bt / nig;
    }

    // ---- load (with the sweep's field rotation and angular momenta)
    real own[NF] = {};   // DIR 0: this thread's cell as loaded (and kicked)
    for (int e = t; e < RS; e += BT) {
        const int s = e / R, r = e - (e / R) * R;
        real v[NF] = {real(1), real(1), real(0), real(0), real(0)};
        if (DIR == 0) {
            const int i = i0 + s - NPAD, ip = i + NPAD;
            if (i < nx + NPAD) {
                const int k = k0;
                real vp;
                if (i >= 0 && i < nx) {
                    const size_t c = (size_t)i + (size_t)nx * ((size_t)j0 + (size_t)ny * k);
                    v[0] = a.in[c]; v[1] = a.in[N + c]; v[2] = a.in[2 * N + c];
                    vp = a.in[3 * N + c]; v[4] = a.in[4 * N + c];
                    if (hkick != real(0)) {
                        const real visr = a.vis ? a.vis[c] : real(0), visp = a.vis ? a.vis[N + c] : real(0);
                        real un, vn, wn;
                        source_cell(a, pl, planet != 0, ip, k + NPAD, j0, v[0], v[2], vp, v[4], visr, visp, real(0),
                                    hkick, un, vn, wn);
                        v[2] = un;
                        vp = vn;
                        v[4] = wn;
                    }
                    own[0] = v[0]; own[1] = v[1]; own[2] = v[2]; own[3] = vp; own[4] = v[4];
                } else {
                    const int g = i < 0 ? ip : ip - nx;
                    const real* xg = a.xg + (size_t)g * nz + k;
                    v[0] = xg[0]; v[1] = xg[6 * nz]; v[2] = xg[12 * nz]; vp = xg[18 * nz]; v[4] = xg[24 * nz];
                }
                const real rad = a.xc[ip];
                const real rcyl = rad * a.sinz[k + NPAD];
                v[3] = (vp + a.vbg[(size_t)k * (nx + 2 * NPAD) + ip]) * rcyl;   // v R_cyl
                v[4] *= rad;                                                    // w r
            }
        } else if (DIR == 1) {
            const int i = i0 + r;
            if (i < nx) {
                const int j = wrapi(j0 + s - NPAD, ny);
                const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k0);
                v[0] = a.in[c]; v[1] = a.in[N + c];
                v[2] = a.in[3 * N + c];   // normal: vp = v - rot - R_cyl Omega
                v[3] = a.in[4 * N + c];   // w
                v[4] = a.in[2 * N + c];   // u
            }
        } else {
            const int i = i0 + r;
            if (i < nx) {
                const int kc = s - NPAD;
                const int k = kc < 0 ? -1 - kc : (kc >= nz ? 2 * nz - 1 - kc : kc);   // mirror
                const real sg = (kc < 0 || kc >= nz) ? real(-1) : real(1);
                const size_t c = (size_t)i + (size_t)nx * ((size_t)j0 + (size_t)ny * k);
                v[0] = a.in[c]; v[1] = a.in[N + c];
                v[2] = sg * a.in[4 * N + c];   // normal: w
                v[3] = a.in[2 * N + c];        // u
                const real rcyl = a.xc[i + NPAD] * a.sinz[s];
                v[4] = (a.in[3 * N + c] + a.vbg[(size_t)k * (nx + 2 * NPAD) + i + NPAD]) * rcyl;
            }
        }
        for (int f = 0; f < NF; ++f) raw[f * RS + e] = v[f];
    }
    __syncthreads();

    // ---- per cell: parameters and the two traced states
    for (int e = t; e < RS; e += BT) {
        const int s = e / R, r = e - (e / R) * R;
        if (s < 2 || s > S - 3) continue;
        real idx, dts;
        const real* w;
        real wy[6];
        if (DIR == 0) {
            const int ip = i0 + s;   // padded radius index of this cell
            if (ip > nx + 2 * NPAD - 3) continue;
            idx = FM<real>::rcp(a.xa[ip + 1] - a.xa[ip]);
            w = a.xw + ip;
            dts = dt;
        } else if (DIR == 1) {
            for (int m = 0; m < 6; ++m) wy[m] = a.yw[m];
            w = wy;
            idx = a.idy;
            const int i = min(i0 + r, nx - 1);
            dts = FM<real>::div(dt, a.xc[i + NPAD] * a.sinz[k0 + NPAD]);
        } else {
            idx = FM<real>::rcp(a.za[s + 1] - a.za[s]);
            w = a.zw + s;
            const int i = min(i0 + r, nx - 1);
            dts = FM<real>::div(dt, a.xc[i + NPAD]);
        }
        real wl[6];
        const int ws = DIR == 0 ? nx + 2 * NPAD : (DIR == 2 ? nz + 2 * NPAD : 1);
        for (int m = 0; m < 6; ++m) wl[m] = DIR == 1 ? w[m] : w[(size_t)m * ws];
        const real* q = raw + e;
        real flat = real(0);
        if (RECON == 6) {
            const real cs = FM<real>::sqrt(FM<real>::div(a.gam * q[RS], q[0]));
            const real* pp = q + RS;
            const real* uu = q + 2 * RS;
            flat = flatten5(pp[-2 * R], pp[-R], pp[0], pp[R], pp[2 * R], uu[-R], uu[0], uu[R], cs);
        }
        Par<real> P[NF];
        for (int f = 0; f < NF; ++f) {
            const real* x = q + f * RS;
            P[f] = params<RECON>(x[-2 * R], x[-R], x[0], x[R], x[2 * R], wl, flat);
        }
        // the cell's pressure less the reference pressures of its left and
        // right edges (the means of the two cells' pressures at each): the
        // differences of neighbouring pressures are exact in the working
        // precision, so these carry no rounding of the pressure itself
        const real pdl = real(0.5) * (q[RS] - q[RS - R]), pdr = real(0.5) * (q[RS] - q[RS + R]);
        real SR[NF], SL[NF];
        trace_R<RECON>(P, q[0], q[RS], q[2 * RS], pdl, a.gam, dts, idx, SR);
        trace_L<RECON>(P, q[0], q[RS], q[2 * RS], pdr, a.gam, dts, idx, SL);
        for (int f = 0; f < NF; ++f) {
            TRs[f * RS + e] = SR[f];
            TLs[f * RS + e] = SL[f];
        }
    }
    __syncthreads();

    // ---- per interface s (between cells s-1 and s): HLLC; every thread
    // reads its traced states before any flux overwrites them
    real Fr[2][NF + 2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int e = t + m * BT, s = e / R;
        if (e >= RS || s < NPAD || s > S - NPAD || (DIR == 0 && i0 + s > nx + 2 * NPAD - 3)) continue;
        real Ls[NF], Rs[NF];
        for (int f = 0; f < NF; ++f) {
            Ls[f] = TLs[f * RS + e - R];
            Rs[f] = TRs[f * RS + e];
        }
        const real pref = real(0.5) * (raw[RS + e - R] + raw[RS + e]);
        hllc_adi(Ls, Rs, pref, a.gam, Fr[m], Fr[m][NF], Fr[m][NF + 1]);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int e = t + m * BT, s = e / R;
        if (e >= RS || s < NPAD || s > S - NPAD || (DIR == 0 && i0 + s > nx + 2 * NPAD - 3)) continue;
        for (int f = 0; f < NF; ++f) TLs[f * RS + e] = Fr[m][f];
        TRs[e] = Fr[m][NF];
        TRs[RS + e] = Fr[m][NF + 1];
    }
    __syncthreads();

    // ---- per cell: flux and source differences (net_source, riemann.cu:40-87), update
    for (int e = t; e < RS; e += BT) {
        const int s = e / R, r = e - (e / R) * R;
        if (s < NPAD || s >= S - NPAD) continue;
        int i, j, k;
        real g0, g1, fac;
        if (DIR == 0) {
            i = i0 + s - NPAD; j = j0; k = k0;
            if (i >= nx) continue;
            g0 = a.xa[i + NPAD] * a.xa[i + NPAD];
            g1 = a.xa[i + NPAD + 1] * a.xa[i + NPAD + 1];
            fac = FM<real>::div(dt, a.xv[i + NPAD]);
        } else if (DIR == 1) {
            i = i0 + r; j = j0 + s - NPAD; k = k0;
            if (i >= nx || j >= ny) continue;
            g0 = g1 = real(1);
            fac = FM<real>::div(dt, a.dy * (a.xc[i + NPAD] * a.sinz[k + NPAD]));
        } else {
            i = i0 + r; j = j0; k = s - NPAD;
            if (i >= nx) continue;
            g0 = a.sza[s];
            g1 = a.sza[s + 1];
            fac = FM<real>::div(dt, a.zv[s] * a.xc[i + NPAD]);
        }
        const real* F0 = TLs + e;
        const real* F1 = TLs + e + R;
        // p* at the two edges: deviations from the edges' references, whose
        // own difference is half that of the neighbours' pressures
        const real* pq = raw + RS + e;
        const real p0d = TRs[e], p1d = TRs[e + R];
        const real du = real(0.5) * (g0 + g1) * ((p0d - p1d) + real(0.5) * (pq[-R] - pq[R]));
        const real p0s = real(0.5) * (pq[-R] + pq[0]) + p0d, p1s = real(0.5) * (pq[0] + pq[R]) + p1d;
        const real de = real(0.5) * (p0s + p1s) * (TRs[RS + e] * g0 - TRs[RS + e + R] * g1);
        real D[NF];
        D[0] = (F0[0] * g0 - F1[0] * g1) * fac;
        D[1] = (F0[RS] * g0 - F1[RS] * g1 + de) * fac;
        D[2] = (F0[2 * RS] * g0 - F1[2 * RS] * g1 + du) * fac;
        D[3] = (F0[3 * RS] * g0 - F1[3 * RS] * g1) * fac;
        D[4] = (F0[4 * RS] * g0 - F1[4 * RS] * g1) * fac;
        const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
        const real vb = a.vbg[(size_t)k * (nx + 2 * NPAD) + i + NPAD];
        real U[NF];   // changes of mass, internal energy, u, vp, w momenta
        if (DIR == 0) {
            const real rad = a.xc[i + NPAD], rcyl = rad * a.sinz[k + NPAD];
            U[0] = D[0]; U[1] = D[1]; U[2] = D[2];
            U[3] = FM<real>::div(D[3], rcyl) - vb * D[0];
            U[4] = FM<real>::div(D[4], rad);
        } else if (DIR == 1) {
            U[0] = D[0]; U[1] = D[1];
            U[3] = D[2];   // vp momentum (the frame term is the background's)
            U[4] = D[3];
            U[2] = D[4];
        } else {
            const real rcyl = a.xc[i + NPAD] * a.sinz[s];
            U[0] = D[0]; U[1] = D[1];
            U[4] = D[2];
            U[2] = D[3];
            U[3] = FM<real>::div(D[4], rcyl) - vb * D[0];
        }
        real cell[NF];
        for (int f = 0; f < NF; ++f) cell[f] = DIR == 0 ? own[f] : a.in[f * N + c];
        update_cell(cell, U, a.gam);
        for (int f = 0; f < NF; ++f) a.out[f * N + c] = cell[f];
    }
}

// --------------------------------------------------- orbital advection
// set_orbital_advection + shift (orbit.cu:17-99) and advecty
// (advection.cu:1-173) in one pass: each row (i, k) is read shifted by its
// integer increment and its residual is transported conservatively.
template <typename real, int RECON>
__global__ __launch_bounds__(SPH_BT) void k3_orbadv(SphK<real> a, double dt, int* inc_out, double* res_out) {
    extern __shared__ __align__(16) unsigned char sm_raw[];
    real* sm = reinterpret_cast<real*>(sm_raw);
    constexpr int R = SPH_R;
    const int S = SPH_TY + 2 * NPAD, RS = R * S;
    real* raw = sm;            // [5][RS]
    real* FX = raw + NF * RS;  // [5][RS]
    __shared__ real s_speed[R];
    __shared__ int s_inc[R];
    const int t = threadIdx.x;
    const int nx = a.nx, ny = a.ny;
    const size_t N = a.N;
    const int nty = (ny + SPH_TY - 1) / SPH_TY, nig = (nx + R - 1) / R;
    const int bt = xcd_tile(blockIdx.x, gridDim.x);
    const int rest = bt / nig, ty = rest % nty;
    const int i0 = (bt - rest * nig) * R, k0 = rest / nty, j0 = ty * SPH_TY;
    if (t < R) {
        const int i = min(i0 + t, nx - 1);
        const size_t ik = (size_t)i + (size_t)nx * k0;
        const double dyc = a.dy_d * a.rcyl[ik];
        const double rot = a.rot[ik];
        const int shift = (int)floor(rot * dt / dyc + 0.5);
        const double speed = shift * dyc / dt;
        s_inc[t] = shift;
        s_speed[t] = (real)(rot - speed);
        if (inc_out && ty == 0 && i0 + t < nx) {
            inc_out[ik] = shift;
            res_out[ik] = rot - speed;
        }
    }
    __syncthreads();
    for (int e = t; e < RS; e += SPH_BT) {
        const int s = e / R, r = e - (e / R) * R;
        const int i = i0 + r;
        real v[NF] = {real(1), real(1), real(0), real(0), real(0)};
        if (i < nx) {
            const int j = wrapi(j0 + s - NPAD - s_inc[r], ny);   // the shifted row
            const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k0);
            v[0] = a.in[c]; v[1] = a.in[N + c];
            v[2] = a.in[3 * N + c]; v[3] = a.in[4 * N + c]; v[4] = a.in[2 * N + c];
        }
        for (int f = 0; f < NF; ++f) raw[f * RS + e] = v[f];
    }
    __syncthreads();
    // fluxes at interfaces s in [3, S-3]: upwind sub-cell means times the speed
    for (int e = t; e < RS; e += SPH_BT) {
        const int s = e / R, r = e - (e / R) * R;
        if (s < NPAD || s > S - NPAD) continue;
        const int i = min(i0 + r, nx - 1);
        const real speed = s_speed[r];
        const real dte = FM<real>::div((real)dt, a.xc[i + NPAD] * a.sinz[k0 + NPAD]);
        const bool up = speed > real(0);
        const int cs = up ? s - 1 : s;
        const real q = up ? lim01(real(1) - speed * dte * a.idy) : lim01(-speed * dte * a.idy);
        const real* x = raw + cs * R + r;
        real D[NF];
        for (int f = 0; f < NF; ++f) {
            const real* y = x + f * RS;
            const Par<real> P = params<RECON>(y[-2 * R], y[-R], y[0], y[R], y[2 * R], a.yw, real(0));
            D[f] = up ? avgR<RECON>(q, P) : avgL<RECON>(q, P);
        }
        FX[e] = D[0] * speed;
        FX[RS + e] = D[1] * speed;
        for (int f = 2; f < NF; ++f) FX[f * RS + e] = D[f] * D[0] * speed;
    }
    __syncthreads();
    for (int e = t; e < RS; e += SPH_BT) {
        const int s = e / R, r = e - (e / R) * R;
        const int i = i0 + r, j = j0 + s - NPAD;
        if (s < NPAD || s >= S - NPAD || i >= nx || j >= ny) continue;
        const real fac = FM<real>::div((real)dt, a.dy * (a.xc[i + NPAD] * a.sinz[k0 + NPAD]));
        real D[NF];
        for (int f = 0; f < NF; ++f) D[f] = (FX[f * RS + e] - FX[f * RS + e + R]) * fac;
        // update_Cell_advect (advection.cu:88-105), EOS_flag 2: no floors
        const real* q = raw + e;
        const real rho = q[0], rn = rho + D[0], ir = FM<real>::rcp(rn);
        const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k0);
        a.out[c] = rn;
        a.out[N + c] = q[RS] + D[1];
        a.out[3 * N + c] = q[2 * RS] + (D[2] - q[2 * RS] * D[0]) * ir;   // (rho u + D) / rn, in increment form
        a.out[4 * N + c] = q[3 * RS] + (D[3] - q[3 * RS] * D[0]) * ir;
        a.out[2 * N + c] = q[4 * RS] + (D[4] - q[4 * RS] * D[0]) * ir;
    }
}

// ---------------------------------------------------------- source terms
// apply_source_terms (sweeps.cu:597-711) with the spherical forces
// (forces.cu:110-327): mode 0 the half kick (in -> out, forces at the state
// itself), modes 1/2 predictor and corrector (forces re-evaluated at the
// midpoint state, in -> out).
// gx, gy, gz: the planet's pull and the indirect term (the star's gravity
// -1/r^2 is part of the tabulated background balance ebal)
template <typename real>
__device__ __forceinline__ void sph_forces(const SphK<real>& a, const PlanetS<real>& pl, bool planet, real rad, int j,
                                           int kp, real& gx, real& gy, real& gz) {
    gx = real(0);
    gy = real(0);
    gz = real(0);
    if (!planet) return;
    // star_planet_grav (forces.cu:14-39) at dt = 0, spline softening; the
    // squared distance r^2 + a^2 - 2 a r cos(dphi) sin(theta) written as
    // (r - a)^2 + 2 a r (2 sin^2(dphi/2) + cos(dphi) cos^2(theta) / (1 + sin(theta)))
    const real sp = a.sinz[kp];
    const real cp = sp * a.cotz[kp];
    real s, c;
    const real sh = a.hphi[2 * j], ch = a.hphi[2 * j + 1];   // sin, cos of phi / 2
    const real hs = sh * pl.chy - ch * pl.shy;   // sin((phi - phi_p) / 2)
    const real hc = ch * pl.chy + sh * pl.shy;   // cos((phi - phi_p) / 2)
    s = real(2) * hs * hc;
    c = hc * hc - hs * hs;
    const real px = pl.x, m = pl.m, rs = pl.rs;
    const real Rp2 = (rad - px) * (rad - px) +
                     real(2) * px * rad * (real(2) * hs * hs + c * FM<real>::div(cp * cp, real(1) + sp));
    const real Rp = FM<real>::sqrt(Rp2);
    const real k = FM<real>::div(fmax(real(4) - real(3) * FM<real>::div(Rp, rs), real(1)), fmax(rs * rs * rs, Rp2 * Rp));
    const real ind = FM<real>::div(m, px * px);
    gx += -m * (rad - px * c * sp) * k - sp * c * ind;
    gy += -m * px * s * k + sp * s * ind;
    gz += m * px * c * cp * k - FM<real>::div(cp * c, sp) * ind;
}

// CFL rate of one cell (get_dt_lv1, timestep.cu:18-97): signal speed over
// cell width in each direction (the azimuthal frame velocity is vp)
template <typename real>
__device__ __forceinline__ real cell_rate(const SphK<real>& a, int ip, int kp, real rho, real p, real u, real vp, real w) {
    const real cs = FM<real>::sqrt(FM<real>::div(a.gam * p, rho));
    const real rad = a.xc[ip], rcyl = rad * a.sinz[kp];
    real t = FM<real>::div(fabs(u) + cs, a.xa[ip + 1] - a.xa[ip]);
    t = fmax(FM<real>::div(fabs(vp) + cs, a.dy * rcyl), t);
    return fmax(FM<real>::div(fabs(w) + cs, (a.za[kp + 1] - a.za[kp]) * rad), t);
}

// maximum of non-negative values through their bit patterns (a NaN rate
// compares above every number and is caught by the host)
__device__ __forceinline__ void atomic_max_nonneg(float* p, float v) {
    atomicMax(reinterpret_cast<unsigned int*>(p), __float_as_uint(v));
}
__device__ __forceinline__ void atomic_max_nonneg(double* p, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

constexpr int CFL_BLOCKS = 1024;

// the velocities after the source terms of one cell (sweeps.cu:597-711):
// forces at the state itself (mdt = 0) or re-evaluated at the midpoint state
// C + (f(C) + g) mdt, then C + (f + g) dt
template <typename real>
__device__ __forceinline__ void source_cell(const SphK<real>& a, const PlanetS<real>& pl, bool planet, int ip, int kp,
                                            int j, real rho, real u, real vp, real w, real visr, real visp, real mdt,
                                            real dt, real& un, real& vn, real& wn) {
    const real rad = a.xc[ip];
    real gx, gy, gz;
    sph_forces(a, pl, planet, rad, j, kp, gx, gy, gz);
    const size_t tk = (size_t)(kp - NPAD) * (a.nx + 2 * NPAD) + ip;
    const real vb = a.vbg[tk], eb = a.ebal[tk];
    const real vis0 = FM<real>::div(visr, rho), vis1 = FM<real>::div(visp, rho);
    const real ir = FM<real>::rcp(rad), cot = a.cotz[kp];
    // get_g_forces (sweeps.cu:555-576): fictitious forces of the state T at
    // azimuthal velocity vb + Tvp.  The radial centrifugal force less the
    // star's gravity is the tabulated balance of the background plus the
    // deviation's part, (vb + Tvp)^2/r - 1/r^2 = ebal + Tvp (2 vb + Tvp)/r:
    // the two large terms cancel in FP64, not in the working precision.
    auto forces = [&](real Tvp, real Tw, real& fx, real& fy, real& fz) {
        const real Tv = vb + Tvp;
        fz = Tv * Tv * ir * cot;
        fy = vis1;
        fx = eb + Tvp * (real(2) * vb + Tvp) * ir + Tw * Tw * ir + vis0;
    };
    real fx, fy, fz;
    forces(vp, w, fx, fy, fz);
    if (mdt != real(0)) {
        const real Tvp = vp + fy * mdt + gy * mdt, Tw = w + fz * mdt + gz * mdt;
        forces(Tvp, Tw, fx, fy, fz);
    }
    un = u + fx * dt + gx * dt;
    vn = vp + fy * dt + gy * dt;
    wn = w + fz * dt + gz * dt;
}

// four cells per thread (a block covers 1024 consecutive cells), their loads
// issued together
constexpr int SRC_CPT = 4;

// with `rate` set (the corrector) the CFL rate of the new state is reduced
// per block into a.rate[block % CFL_BLOCKS] (zeroed beforehand): the next
// step's dt then needs no pass of its own
template <typename real>
__global__ __launch_bounds__(256) void k3_sources(SphK<real> a, PlanetS<real> pl, int planet, real mdt, real dt,
                                                  int rate) {
    const int nx = a.nx, ny = a.ny;
    real rmax = real(0);
    const size_t N = a.N;
    const size_t c0 = (size_t)blockIdx.x * (256 * SRC_CPT) + threadIdx.x;
    real rho[SRC_CPT], u[SRC_CPT], vp[SRC_CPT], w[SRC_CPT], v0[SRC_CPT], v1[SRC_CPT];
#pragma unroll
    for (int m = 0; m < SRC_CPT; ++m) {
        const size_t c = min(c0 + (size_t)m * 256, N - 1);
        rho[m] = a.in[c]; u[m] = a.in[2 * N + c]; vp[m] = a.in[3 * N + c]; w[m] = a.in[4 * N + c];
        v0[m] = a.vis ? a.vis[c] : real(0);
        v1[m] = a.vis ? a.vis[N + c] : real(0);
    }
#pragma unroll
    for (int m = 0; m < SRC_CPT; ++m) {
        const size_t c = c0 + (size_t)m * 256;
        if (c >= N) break;
        const int i = (int)(c % nx), jk = (int)(c / nx), j = jk % ny, k = jk / ny;
        const int ip = i + NPAD, kp = k + NPAD;
        real un, vn, wn;
        source_cell(a, pl, planet != 0, ip, kp, j, rho[m], u[m], vp[m], w[m], v0[m], v1[m], mdt, dt, un, vn, wn);
        a.out[2 * N + c] = un;
        a.out[3 * N + c] = vn;
        a.out[4 * N + c] = wn;
        if (a.out != a.in) {
            a.out[c] = rho[m];
            a.out[N + c] = a.in[N + c];
        }
        if (rate) rmax = fmax(rmax, cell_rate(a, ip, kp, rho[m], a.in[N + c], un, vn, wn));
    }
    if (rate) {
        __shared__ real red[256];
        red[threadIdx.x] = rmax;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
            __syncthreads();
        }
        if (threadIdx.x == 0) atomic_max_nonneg(a.rate + blockIdx.x % CFL_BLOCKS, red[0]);
    }
}

// ------------------------------------------------------------- viscosity
// viscous_forces (viscosity.cu:34-197) on every polar layer: the (r, phi)
// stress divergence with the spherical radius, nu at the layer's angle; the
// polar component is zero.  Tiles of VX radii x VY rows of one layer with a
// one-cell halo staged in LDS (rho, u, full v); 256 threads, each computing
// VY / 4 cells of one column.
constexpr int VX = 64, VY = 16;

// PRED: the predictor (apply_source_terms with mdt = dt = dt/2 into T,
// sweeps.cu:597-711) is applied to the staged interior cells first, so the
// kernel yields the viscosity of the predicted state without storing it
// (radial ghosts stay analytic, as boundx2 makes them, boundary.cu:30-33)
template <typename real, bool PRED>
__global__ __launch_bounds__(256) void k3_viscosity(SphK<real> a, PlanetS<real> pl, int planet, real hdt) {
    constexpr int HX = VX + 2, HY = VY + 2;
    __shared__ real sr[HY * HX], su[HY * HX], sv[HY * HX];
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const size_t N = a.N;
    const int nxe = nx + 2 * NPAD;
    const int ntx = (nx + VX - 1) / VX, nty = (ny + VY - 1) / VY;
    const int b = xcd_tile(blockIdx.x, gridDim.x), k = b / (ntx * nty), rem = b - k * ntx * nty;
    const int i0 = (rem % ntx) * VX, j0 = (rem / ntx) * VY;
    const int t = threadIdx.x;
    for (int e = t; e < HX * HY; e += 256) {
        const int lx = e % HX, ly = e / HX;
        const int ii = i0 - 1 + lx, jj = wrapi(j0 - 1 + ly, ny), ip = ii + NPAD;
        real rho = real(1), uu = real(0), vp = real(0);
        if (ii >= 0 && ii < nx) {
            const size_t cc = (size_t)ii + (size_t)nx * ((size_t)jj + (size_t)ny * k);
            rho = a.in[cc]; uu = a.in[2 * N + cc]; vp = a.in[3 * N + cc];
            if (PRED) {
                real wn;
                source_cell(a, pl, planet != 0, ip, k + NPAD, jj, rho, uu, vp, a.in[4 * N + cc], a.vis[cc],
                            a.vis[N + cc], hdt, hdt, uu, vp, wn);
            }
        } else if (ii >= -1 && ii <= nx) {
            const int g = ii < 0 ? ip : ip - nx;
            const real* xg = a.xg + (size_t)g * nz + k;
            rho = xg[0]; uu = xg[12 * nz]; vp = xg[18 * nz];
        }
        sr[e] = rho;
        su[e] = uu;
        sv[e] = ip >= 0 && ip < nxe ? vp + a.vbg[(size_t)k * nxe + ip] : vp;
    }
    __syncthreads();
    const int tx = t % VX, i = i0 + tx;
    if (i >= nx) return;
    real r[3], nu[3];
    for (int n = 0; n < 3; ++n) {
        r[n] = a.xc[i + NPAD - 1 + n];
        nu[n] = a.nu[(size_t)k * nxe + i + NPAD - 1 + n];
    }
    const real dy = a.dy;
    const real ir1 = FM<real>::rcp(r[1]);
    // reciprocal spacings and derivative scales of this thread's radius,
    // formed once for all its rows
    const real i20 = FM<real>::rcp(r[2] - r[0]), i21 = FM<real>::rcp(r[2] - r[1]), i10 = FM<real>::rcp(r[1] - r[0]);
    const real i20x2 = real(2) * i20;                      // 1 / ((r2 - r0) / 2)
    const real hir = FM<real>::rcp(real(2) * dy) * ir1;    // 1 / (2 dy r)
    const real idy = FM<real>::rcp(dy);
    const real yy = (idy * idy) * (ir1 * ir1);             // 1 / (dy r)^2
    const real xy = i20 * hir;
    auto d1 = [&](const real* f, real& dx, real& dxx) {
        dx = (f[2] - f[0]) * i20;
        dxx = ((f[2] - f[1]) * i21 - (f[1] - f[0]) * i10) * i20x2;
    };
    auto d2 = [&](const real* f, real& dyv, real& dyy, real& dxy) {
        dyv = (f[7] - f[4]) * hir;
        dyy = ((f[7] - f[1]) - (f[1] - f[4])) * yy;
        dxy = (f[8] - f[6] - f[5] + f[3]) * xy;
    };
    for (int ty = t / VX; ty < VY; ty += 256 / VX) {
        const int j = j0 + ty;
        if (j >= ny) break;
        real u[9], v[9], nr[9];
        for (int n = 0; n < 9; ++n) {
            const int di = n % 3 - 1, dj = n < 3 ? 0 : (n < 6 ? -1 : 1);
            const int e = (ty + 1 + dj) * HX + tx + 1 + di;
            nr[n] = sr[e] * nu[n % 3];
            u[n] = su[e];
            v[n] = sv[e];
        }
        real dudx, d2udx2, dvdx, d2vdx2, dvdy, d2vdy2, d2vdxy, dudy, d2udy2, d2udxy;
        d1(u, dudx, d2udx2);
        const real dndx = (nr[2] - nr[0]) * i20;
        d1(v, dvdx, d2vdx2);
        d2(v, dvdy, d2vdy2, d2vdxy);
        d2(u, dudy, d2udy2, d2udxy);
        const real dndy = (nr[7] - nr[4]) * hir;
        // stresses sxx = 2 n dudx, syy = 2 n (dvdy + u/r), sxy = n (dudy +
        // dvdx - v/r) and their divergence, the terms over r gathered
        const real n1 = nr[1], n2 = real(2) * n1;
        const real vir = v[1] * ir1, uir = u[1] * ir1;
        const real tt = dudy + dvdx - vir;
        const real sxx = n2 * dudx, syy = n2 * (dvdy + uir), sxy = n1 * tt;
        const real fx = real(2) * dndx * dudx + n2 * d2udx2 + dndy * tt + n1 * (d2udy2 + d2vdxy) +
                        ir1 * (sxx - n1 * dvdy - syy);
        const real fy = real(2) * dndy * (dvdy + uir) + n2 * d2vdy2 + dndx * tt + n1 * (d2udxy + d2vdx2) +
                        ir1 * (n2 * dudy - n1 * dvdx + n1 * vir + real(2) * sxy);
        const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
        a.vis_out[c] = fx;
        a.vis_out[N + c] = fy;
    }
}

// ------------------------------------------------------------------ CFL
// get_dt_lv1 (timestep.cu:18-97): per-block maxima of the signal speed over
// the cell width in each direction
template <typename real>
__global__ __launch_bounds__(256) void k3_cfl(SphK<real> a) {
    __shared__ real red[256];
    real m = real(0);
    const size_t N = a.N;
    const int nx = a.nx, ny = a.ny;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % nx), k = (int)(c / nx / ny);
        m = fmax(m, cell_rate(a, i + NPAD, k + NPAD, a.in[c], a.in[N + c], a.in[2 * N + c], a.in[3 * N + c],
                              a.in[4 * N + c]));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) a.rate[blockIdx.x] = red[0];
}

}  // namespace

// =============================================================== host side
namespace {

const double kPi = 3.1415926535897932384626434;

// init_flag 3 profiles (init.cu:7-70, 72-349), geomx 2: x spherical radius,
// z polar angle
struct Prof3 {
    const pg_params& P;
    double cs2(double x, double z) const {
        return P.sc_h * P.sc_h / P.planet_radius * std::pow(x * std::sin(z) / P.planet_radius, -P.p_beta);
    }
    double h(double x, double z) const { return std::sqrt(cs2(x, z) * std::pow(x * std::sin(z), 3)); }
    double nu(double x, double z) const {
        if (!P.visc_flag) return 0.0;
        const double rp = P.planet_radius;
        return P.ss_alpha * std::sqrt(cs2(rp, z)) * h(rp, z) * std::pow(x * std::sin(z) / rp, P.p_alpha);
    }
    double vert(double x, double z) const { return std::exp((1.0 / x - 1.0 / (x * std::sin(z))) / cs2(x, z)); }
    double rho(double x, double z) const {
        return std::pow(x * std::sin(z) / P.planet_radius, -P.p_alpha) * vert(x, z);
    }
    double p(double x, double z) const { return cs2(x, z) * rho(x, z); }
    double u(double x, double z) const { return -1.5 * nu(x, z) / (x * std::sin(z)); }
    double v(double x, double z) const {
        const double dr = 1.1514178e-7;
        const double dpdr = (p(x + dr, z) - p(x - dr, z)) / (2.0 * dr);
        return std::sqrt(1.0 / x + (x / rho(x, z)) * dpdr);
    }
};

uint64_t smix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double keyed_u(uint64_t seed, int field, int i, int j) {
    uint64_t key = smix(seed) ^ smix(((uint64_t)field << 48) ^ ((uint64_t)j << 24) ^ (uint64_t)i);
    return (double)(smix(key) >> 11) * (1.0 / 9007199254740992.0);
}

#define SPH_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return report_error(PG_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                __FILE__, __LINE__);                                           \
    } while (0)

const char* kSphKernels[] = {"sweep_x_gas", "sweep_y_gas", "sweep_z_gas", "orbital_advection", "viscosity",
                             "sources", "cfl"};

}  // namespace

struct Sph {
    pg_params P;
    int nx = 0, ny = 0, nz = 0;
    size_t N = 0;
    int es = 4;   // bytes per real
    hipStream_t stream = nullptr;
    void* st[3] = {nullptr, nullptr, nullptr};   // state buffers A, B, predicted
    void* vis = nullptr;
    void* vis2 = nullptr;       // the fused predictor writes here, then the two swap
    void* tab = nullptr;        // all real tables, one allocation
    double* dtab = nullptr;     // FP64 tables: rot, rcyl
    void* rate = nullptr;
    void* rate_h = nullptr;     // pinned
    size_t tab_off[16] = {0};
    std::vector<double> xa, xv, ya, yv, za, zv, vbg, rot, rcyl;
    std::vector<double> xgh;    // radial ghost states (rho, p, u, v, w) [field][ghost 0..5][polar k], FP64
    Body planet{};              // planet at the time of the last evolve
    double yw[6] = {0};
    size_t bytes = 0;
    bool rate_valid = false;    // rate partials hold the current state's rate (written by the corrector)
    bool profile = false;
    std::map<std::string, std::pair<double, int>> timers;
};

namespace {

enum { TB_XA, TB_XV, TB_XC, TB_XW, TB_ZA, TB_ZV, TB_ZC, TB_ZW, TB_SZA, TB_SINZ, TB_COTZ, TB_VBG, TB_NU, TB_XG, TB_HPHI,
       TB_EBAL, TB_N };

template <typename real>
SphK<real> kargs(Sph* s, int in, int out) {
    SphK<real> a;
    std::memset(&a, 0, sizeof(a));
    a.nx = s->nx; a.ny = s->ny; a.nz = s->nz; a.N = s->N;
    a.in = in >= 0 ? (const real*)s->st[in] : nullptr;
    a.out = out >= 0 ? (real*)s->st[out] : nullptr;
    const real* T = (const real*)s->tab;
    a.xa = T + s->tab_off[TB_XA]; a.xv = T + s->tab_off[TB_XV]; a.xc = T + s->tab_off[TB_XC];
    a.xw = T + s->tab_off[TB_XW];
    a.za = T + s->tab_off[TB_ZA]; a.zv = T + s->tab_off[TB_ZV]; a.zc = T + s->tab_off[TB_ZC];
    a.zw = T + s->tab_off[TB_ZW]; a.sza = T + s->tab_off[TB_SZA]; a.sinz = T + s->tab_off[TB_SINZ];
    a.cotz = T + s->tab_off[TB_COTZ];
    a.vbg = T + s->tab_off[TB_VBG]; a.nu = T + s->tab_off[TB_NU]; a.xg = T + s->tab_off[TB_XG];
    a.ebal = T + s->tab_off[TB_EBAL];
    a.hphi = T + s->tab_off[TB_HPHI];
    a.rot = s->dtab;
    a.rcyl = s->dtab + (size_t)s->nx * s->nz;
    a.rate = (real*)s->rate;
    for (int m = 0; m < 6; ++m) a.yw[m] = (real)s->yw[m];
    const double dy = (s->P.ymax - s->P.ymin) / (double)s->ny;
    a.dy = (real)dy;
    a.idy = (real)(1.0 / dy);
    a.dy_d = dy;
    a.gam = (real)s->P.gam;
    a.tx = sph_tx(s->nx);
    return a;
}

template <typename real>
size_t sweep_lds(int dir, int nx, int nz) {
    const int R = dir == 0 ? 1 : SPH_R;
    const int L = dir == 0 ? sph_tx(nx) : dir == 1 ? SPH_TY : nz;
    return (size_t)(3 * NF) * R * (L + 2 * NPAD) * sizeof(real);
}

// launch timing (profiling): events around one launch, synchronous
struct Timed {
    Sph* s;
    const char* name;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Timed(Sph* s_, const char* n) : s(s_), name(n) {
        if (s->profile) {
            (void)hipEventCreate(&e0);
            (void)hipEventCreate(&e1);
            (void)hipEventRecord(e0, s->stream);
        }
    }
    ~Timed() {
        if (!s->profile) return;
        (void)hipEventRecord(e1, s->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        auto& t = s->timers[name];
        t.first += 1e3 * ms;
        t.second += 1;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
};

// kernels given more than 64 KB of dynamic LDS must say so once
template <typename K>
hipError_t allow_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

template <typename real>
PlanetS<real> planet_view(const Body& b);

// kick (radial sweep only): the first half kick with the planet at t over
// hkick = dt/2 applied to the cells as they are loaded
template <typename real>
int launch_sweep(Sph* s, int dir, int in, int out, double dt, const Body* kick = nullptr, double hkick = 0.0) {
    SphK<real> a = kargs<real>(s, in, out);
    PlanetS<real> pl{};
    if (kick) {
        if (dir != 0) return report_error(PG_E_INVALID, "the fused kick belongs to the radial sweep");
        a.vis = s->P.visc_flag ? (const real*)s->vis : nullptr;
        pl = planet_view<real>(*kick);
    }
    const int planet = kick && s->P.n_planet > 0 ? 1 : 0;
    const real hk = kick ? (real)hkick : real(0);
    const size_t lds = sweep_lds<real>(dir, s->nx, s->nz);
    unsigned nb;
    if (dir == 0) nb = (unsigned)(((s->nx + a.tx - 1) / a.tx) * s->ny * s->nz);
    else if (dir == 1) nb = (unsigned)(((s->ny + SPH_TY - 1) / SPH_TY) * ((s->nx + SPH_R - 1) / SPH_R) * s->nz);
    else nb = (unsigned)(((s->nx + SPH_R - 1) / SPH_R) * s->ny);
    Timed tm(s, kSphKernels[dir]);
    const bool pem = s->P.recon_flag == 5;
#define SPH_SWEEP(D)                                                                                   \
    do {                                                                                               \
        const int bt = D == 0 ? sph_bx(s->nx) : SPH_BT2;                                               \
        if (pem) {                                                                                     \
            SPH_TRY(allow_lds(k3_sweep<real, D, 5>, lds));                                             \
            hipLaunchKernelGGL((k3_sweep<real, D, 5>), dim3(nb), dim3(bt), lds, s->stream, a, (real)dt, pl, \
                               planet, hk);                                                            \
        } else {                                                                                       \
            SPH_TRY(allow_lds(k3_sweep<real, D, 6>, lds));                                             \
            hipLaunchKernelGGL((k3_sweep<real, D, 6>), dim3(nb), dim3(bt), lds, s->stream, a, (real)dt, pl, \
                               planet, hk);                                                            \
        }                                                                                              \
    } while (0)
    if (dir == 0) SPH_SWEEP(0);
    else if (dir == 1) SPH_SWEEP(1);
    else SPH_SWEEP(2);
#undef SPH_SWEEP
    SPH_TRY(hipGetLastError());
    return PG_OK;
}

template <typename real>
int launch_orbadv(Sph* s, int in, int out, double dt, int* inc_d, double* res_d) {
    SphK<real> a = kargs<real>(s, in, out);
    const size_t lds = (size_t)2 * NF * SPH_R * (SPH_TY + 2 * NPAD) * sizeof(real);
    const unsigned nb = (unsigned)(((s->ny + SPH_TY - 1) / SPH_TY) * ((s->nx + SPH_R - 1) / SPH_R) * s->nz);
    SPH_TRY(allow_lds(k3_orbadv<real, 5>, lds));
    SPH_TRY(allow_lds(k3_orbadv<real, 6>, lds));
    Timed tm(s, "orbital_advection");
    if (s->P.recon_flag == 5)
        hipLaunchKernelGGL((k3_orbadv<real, 5>), dim3(nb), dim3(SPH_BT), lds, s->stream, a, dt, inc_d, res_d);
    else
        hipLaunchKernelGGL((k3_orbadv<real, 6>), dim3(nb), dim3(SPH_BT), lds, s->stream, a, dt, inc_d, res_d);
    SPH_TRY(hipGetLastError());
    return PG_OK;
}

template <typename real>
PlanetS<real> planet_view(const Body& b) {
    PlanetS<real> pl;
    pl.m = (real)b.m; pl.x = (real)b.x; pl.y = (real)b.y; pl.rs = (real)b.rs;
    pl.shy = (real)std::sin(0.5 * b.y);
    pl.chy = (real)std::cos(0.5 * b.y);
    return pl;
}

template <typename real>
int launch_vis(Sph* s, int in) {
    SphK<real> a = kargs<real>(s, in, -1);
    a.vis_out = (real*)s->vis;
    Timed tm(s, "viscosity");
    const unsigned nb = (unsigned)(((s->nx + VX - 1) / VX) * ((s->ny + VY - 1) / VY) * s->nz);
    hipLaunchKernelGGL((k3_viscosity<real, false>), dim3(nb), dim3(256), 0, s->stream, a, PlanetS<real>{}, 0, real(0));
    SPH_TRY(hipGetLastError());
    return PG_OK;
}

// the predictor and the viscosity of the predicted state in one pass: vis
// (the start-of-step viscosity) is read, the predicted viscosity written to
// the second viscosity buffer, which the caller then swaps in
template <typename real>
int launch_predvis(Sph* s, int in, const Body& b, double hdt) {
    SphK<real> a = kargs<real>(s, in, -1);
    a.vis = (const real*)s->vis;
    a.vis_out = (real*)s->vis2;
    PlanetS<real> pl = planet_view<real>(b);
    Timed tm(s, "predictor_viscosity");
    const unsigned nb = (unsigned)(((s->nx + VX - 1) / VX) * ((s->ny + VY - 1) / VY) * s->nz);
    hipLaunchKernelGGL((k3_viscosity<real, true>), dim3(nb), dim3(256), 0, s->stream, a, pl,
                       s->P.n_planet > 0 ? 1 : 0, (real)hdt);
    SPH_TRY(hipGetLastError());
    std::swap(s->vis, s->vis2);
    return PG_OK;
}

template <typename real>
int launch_sources(Sph* s, int in, int out, const Body& b, double mdt, double dt, bool rate = false) {
    SphK<real> a = kargs<real>(s, in, out);
    a.vis = s->P.visc_flag ? (const real*)s->vis : nullptr;
    PlanetS<real> pl = planet_view<real>(b);
    if (rate) SPH_TRY(hipMemsetAsync(s->rate, 0, CFL_BLOCKS * sizeof(real), s->stream));
    Timed tm(s, "sources");
    hipLaunchKernelGGL(k3_sources<real>, dim3((unsigned)((s->N + 256 * SRC_CPT - 1) / (256 * SRC_CPT))), dim3(256), 0,
                       s->stream, a, pl, s->P.n_planet > 0 ? 1 : 0, (real)mdt, (real)dt, rate ? 1 : 0);
    SPH_TRY(hipGetLastError());
    return PG_OK;
}

template <typename real>
int rate_of(Sph* s, int in, double* rate) {
    if (!s->rate_valid) {
        SphK<real> a = kargs<real>(s, in, -1);
        Timed tm(s, "cfl");
        hipLaunchKernelGGL(k3_cfl<real>, dim3(CFL_BLOCKS), dim3(256), 0, s->stream, a);
        SPH_TRY(hipGetLastError());
    }
    SPH_TRY(hipMemcpyAsync(s->rate_h, s->rate, CFL_BLOCKS * sizeof(real), hipMemcpyDeviceToHost, s->stream));
    SPH_TRY(hipStreamSynchronize(s->stream));
    double m = 0.0;
    const real* h = (const real*)s->rate_h;
    for (int b = 0; b < CFL_BLOCKS; ++b) m = std::fmax(m, (double)h[b]);
    *rate = m;
    return PG_OK;
}

void evolve_planet(Sph* s, double t) {
    if (s->P.n_planet > 0) orbit_solution(s->planet, t, s->P.ramp_time, s->P.planet_mass, s->P.frame_omega);
}

// DS (solve.cu:212-277) for ndim 3: the state is in buffer 0 before and after
template <typename real>
int ds_step(Sph* s, double t, double dt) {
    const double hdt = 0.5 * dt;
    int rc;
    s->rate_valid = false;   // until the corrector has written the new state's rate
    const bool visc = s->P.visc_flag != 0;
    if (visc && (rc = launch_vis<real>(s, 0))) return rc;
    // first half kick fused into the radial sweep's load
    if ((rc = launch_sweep<real>(s, 0, 0, 1, dt, &s->planet, hdt))) return rc;
    if ((rc = launch_sweep<real>(s, 1, 1, 0, dt))) return rc;
    if ((rc = launch_sweep<real>(s, 2, 0, 1, dt))) return rc;
    if ((rc = launch_orbadv<real>(s, 1, 0, dt, nullptr, nullptr))) return rc;
    evolve_planet(s, t + dt);
    if (visc && (rc = launch_predvis<real>(s, 0, s->planet, hdt))) return rc;
    if ((rc = launch_sources<real>(s, 0, 0, s->planet, hdt, hdt, true))) return rc;
    s->rate_valid = true;
    return PG_OK;
}

template <typename real>
int upload_tables(Sph* s) {
    const pg_params& P = s->P;
    const int nx = s->nx, ny = s->ny, nz = s->nz, nxe = nx + 2 * NPAD, nze = nz + 2 * NPAD;
    Prof3 pr{P};
    std::vector<double> xc(nxe), zc(nze);
    for (int i = 0; i < nxe; ++i) xc[i] = 0.5 * (s->xa[i + 1] + s->xa[i]);
    for (int k = 0; k < nze; ++k) zc[k] = 0.5 * (s->za[k + 1] + s->za[k]);
    std::vector<double> xw(6 * (size_t)nxe, 0.0), zw(6 * (size_t)nze, 0.0);
    for (int ip = 2; ip < nxe - 2; ++ip) {
        double wl[3], wr[3];
        host_slope_weights(P.geomx, &s->xa[ip - 2], wl, wr);
        for (int m = 0; m < 3; ++m) {
            xw[(size_t)m * nxe + ip] = wl[m];
            xw[(size_t)(m + 3) * nxe + ip] = wr[m];
        }
    }
    for (int kp = 2; kp < nze - 2; ++kp) {
        double wl[3], wr[3];
        host_slope_weights(P.geomz, &s->za[kp - 2], wl, wr);
        for (int m = 0; m < 3; ++m) {
            zw[(size_t)m * nze + kp] = wl[m];
            zw[(size_t)(m + 3) * nze + kp] = wr[m];
        }
    }
    {
        double wl[3], wr[3];
        host_slope_weights(P.geomy, &s->ya[NPAD - 2], wl, wr);
        for (int m = 0; m < 3; ++m) {
            s->yw[m] = wl[m];
            s->yw[m + 3] = wr[m];
        }
    }
    std::vector<double> sza(nze + 1), sinz(nze), cotz(nze);
    for (int k = 0; k <= nze; ++k) sza[k] = std::sin(s->za[k]);
    for (int k = 0; k < nze; ++k) {
        sinz[k] = std::sin(zc[k]);
        cotz[k] = std::cos(zc[k]) / std::sin(zc[k]);
    }
    s->vbg.assign((size_t)nxe * nz, 0.0);
    std::vector<double> nu((size_t)nxe * nz, 0.0), xg((size_t)NF * 6 * nz, 0.0), ebal((size_t)nxe * nz, 0.0);
    s->rot.assign((size_t)nx * nz, 0.0);
    s->rcyl.assign((size_t)nx * nz, 0.0);
    s->xgh.assign((size_t)NF * 6 * nz, 0.0);
    for (int k = 0; k < nz; ++k) {
        const double z = zc[k + NPAD];
        for (int ip = 0; ip < nxe; ++ip) {
            const double vb = pr.v(xc[ip], z);
            s->vbg[(size_t)k * nxe + ip] = vb;
            nu[(size_t)k * nxe + ip] = pr.nu(xc[ip], z);
            ebal[(size_t)k * nxe + ip] = vb * vb / xc[ip] - 1.0 / (xc[ip] * xc[ip]);
        }
        for (int i = 0; i < nx; ++i) {
            const double rc = xc[i + NPAD] * std::sin(z);
            s->rcyl[(size_t)k * nx + i] = rc;
            // init_orbital_advection (orbit.cu:51-72)
            s->rot[(size_t)k * nx + i] = pr.v(xc[i + NPAD], z) - rc * P.frame_omega;
        }
        for (int g = 0; g < 6; ++g) {
            const int ip = g < 3 ? g : nx + g;
            const double x = xc[ip];
            const double f[NF] = {pr.rho(x, z), pr.p(x, z), pr.u(x, z), pr.v(x, z) - s->vbg[(size_t)k * nxe + ip], 0.0};
            for (int q = 0; q < NF; ++q) xg[((size_t)q * 6 + g) * nz + k] = f[q];
            const double fh[NF] = {f[0], f[1], f[2], pr.v(x, z), 0.0};
            for (int q = 0; q < NF; ++q) s->xgh[((size_t)q * 6 + g) * nz + k] = fh[q];
        }
    }
    std::vector<double> hphi(2 * (size_t)ny);
    const double dyh = (P.ymax - P.ymin) / ny;
    for (int j = 0; j < ny; ++j) {
        const double phi = P.ymin + (j + 0.5) * dyh;
        hphi[2 * (size_t)j] = std::sin(0.5 * phi);
        hphi[2 * (size_t)j + 1] = std::cos(0.5 * phi);
    }
    // one real allocation for the tables
    std::vector<const std::vector<double>*> parts(TB_N);
    std::vector<double> xa = s->xa, xv = s->xv, za = s->za, zv = s->zv;
    parts[TB_XA] = &xa; parts[TB_XV] = &xv; parts[TB_XC] = &xc; parts[TB_XW] = &xw;
    parts[TB_ZA] = &za; parts[TB_ZV] = &zv; parts[TB_ZC] = &zc; parts[TB_ZW] = &zw;
    parts[TB_SZA] = &sza; parts[TB_SINZ] = &sinz; parts[TB_COTZ] = &cotz;
    parts[TB_VBG] = &s->vbg; parts[TB_NU] = &nu; parts[TB_XG] = &xg; parts[TB_HPHI] = &hphi; parts[TB_EBAL] = &ebal;
    size_t tot = 0;
    for (int q = 0; q < TB_N; ++q) {
        s->tab_off[q] = tot;
        tot += (parts[q]->size() + 63) / 64 * 64;
    }
    std::vector<real> h(tot, real(0));
    for (int q = 0; q < TB_N; ++q)
        for (size_t m = 0; m < parts[q]->size(); ++m) h[s->tab_off[q] + m] = (real)(*parts[q])[m];
    SPH_TRY(hipMalloc(&s->tab, tot * sizeof(real)));
    s->bytes += tot * sizeof(real);
    SPH_TRY(hipMemcpy(s->tab, h.data(), tot * sizeof(real), hipMemcpyHostToDevice));
    std::vector<double> d(s->rot);
    d.insert(d.end(), s->rcyl.begin(), s->rcyl.end());
    SPH_TRY(hipMalloc((void**)&s->dtab, d.size() * sizeof(double)));
    s->bytes += d.size() * sizeof(double);
    SPH_TRY(hipMemcpy(s->dtab, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
    return PG_OK;
}

// host FP64 state (5, nz, ny, nx) <-> device buffer 0
template <typename real>
int put_state(Sph* s, const double* gas) {
    s->rate_valid = false;
    const size_t N = s->N;
    const int nx = s->nx, ny = s->ny, nxe = nx + 2 * NPAD;
    std::vector<real> h((size_t)NF * N);
    for (size_t c = 0; c < N; ++c) {
        const int i = (int)(c % nx), k = (int)(c / nx / ny);
        h[c] = (real)gas[c];
        h[N + c] = (real)gas[N + c];
        h[2 * N + c] = (real)gas[2 * N + c];
        h[3 * N + c] = (real)(gas[3 * N + c] - s->vbg[(size_t)k * nxe + i + NPAD]);
        h[4 * N + c] = (real)gas[4 * N + c];
    }
    SPH_TRY(hipMemcpyAsync(s->st[0], h.data(), h.size() * sizeof(real), hipMemcpyHostToDevice, s->stream));
    SPH_TRY(hipStreamSynchronize(s->stream));
    return PG_OK;
}
template <typename real>
int get_state(Sph* s, double* gas) {
    const size_t N = s->N;
    const int nx = s->nx, ny = s->ny, nxe = nx + 2 * NPAD;
    std::vector<real> h((size_t)NF * N);
    SPH_TRY(hipMemcpyAsync(h.data(), s->st[0], h.size() * sizeof(real), hipMemcpyDeviceToHost, s->stream));
    SPH_TRY(hipStreamSynchronize(s->stream));
    for (size_t c = 0; c < N; ++c) {
        const int i = (int)(c % nx), k = (int)(c / nx / ny);
        gas[c] = (double)h[c];
        gas[N + c] = (double)h[N + c];
        gas[2 * N + c] = (double)h[2 * N + c];
        gas[3 * N + c] = (double)h[3 * N + c] + s->vbg[(size_t)k * nxe + i + NPAD];
        gas[4 * N + c] = (double)h[4 * N + c];
    }
    return PG_OK;
}

template <typename real>
int initial_state(Sph* s) {
    const pg_params& P = s->P;
    Prof3 pr{P};
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const size_t N = s->N;
    std::vector<double> g((size_t)NF * N);
    for (int k = 0; k < nz; ++k) {
        const double z = 0.5 * (s->za[k + NPAD] + s->za[k + NPAD + 1]);
        for (int i = 0; i < nx; ++i) {
            const double x = 0.5 * (s->xa[i + NPAD] + s->xa[i + NPAD + 1]);
            const double rho0 = pr.rho(x, z), p0 = pr.p(x, z), u0 = pr.u(x, z), v0 = pr.v(x, z);
            const double cs2 = pr.cs2(x, z);
            for (int j = 0; j < ny; ++j) {
                const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                // seeded perturbation as the oracle's (keys i, j + ny k)
                double rho = rho0, u = u0;
                const int jk = j + ny * k;
                if (P.noise_rho != 0.0) rho *= 1.0 + P.noise_rho * (2.0 * keyed_u(P.seed, 0, i, jk) - 1.0);
                if (P.noise_vr != 0.0) u += P.noise_vr * std::sqrt(P.gam * cs2) * (2.0 * keyed_u(P.seed, 1, i, jk) - 1.0);
                g[c] = rho;
                g[N + c] = p0;
                g[2 * N + c] = u;
                g[3 * N + c] = v0;
                g[4 * N + c] = 0.0;
            }
        }
    }
    return put_state<real>(s, g.data());
}

template <typename real>
int create_t(Sph* s) {
    const pg_params& P = s->P;
    const size_t N = s->N;
    s->xa.assign(s->nx + 2 * NPAD + 1, 0.0); s->xv.assign(s->nx + 2 * NPAD, 0.0);
    s->ya.assign(s->ny + 2 * NPAD + 1, 0.0); s->yv.assign(s->ny + 2 * NPAD, 0.0);
    s->za.assign(s->nz + 2 * NPAD + 1, 0.0); s->zv.assign(s->nz + 2 * NPAD, 0.0);
    host_make_axis(s->xa.data(), s->xv.data(), P.xmin, P.xmax, s->nx, NPAD, P.geomx, P.gridx);
    host_make_axis(s->ya.data(), s->yv.data(), P.ymin, P.ymax, s->ny, NPAD, P.geomy, P.gridy);
    host_make_axis(s->za.data(), s->zv.data(), P.zmin, P.zmax, s->nz, NPAD, P.geomz, P.gridz);
    SPH_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (int b = 0; b < 3; ++b) {
        SPH_TRY(hipMalloc(&s->st[b], (size_t)NF * N * sizeof(real)));
        s->bytes += (size_t)NF * N * sizeof(real);
    }
    SPH_TRY(hipMalloc(&s->vis, 2 * N * sizeof(real)));
    SPH_TRY(hipMalloc(&s->vis2, 2 * N * sizeof(real)));
    s->bytes += 4 * N * sizeof(real);
    SPH_TRY(hipMemset(s->vis, 0, 2 * N * sizeof(real)));
    SPH_TRY(hipMalloc(&s->rate, CFL_BLOCKS * sizeof(real)));
    SPH_TRY(hipHostMalloc(&s->rate_h, CFL_BLOCKS * sizeof(real), hipHostMallocDefault));
    int rc;
    if ((rc = upload_tables<real>(s))) return rc;
    // init_planet (planet.cu:319-403): spline softening rs_fac (q/3)^(1/3) a
    // in 3D (planet.cu:308-317)
    s->planet = Body{0, P.planet_radius, P.planet_ecc, 0, 0, 0, 0, 0, 0, 0};
    evolve_planet(s, 0.0);
    s->planet.rs = P.rs_fac * std::pow(P.planet_mass / 3.0, 1.0 / 3.0) * P.planet_radius;
    return initial_state<real>(s);
}

}  // namespace

// ------------------------------------------------------------- entry points
int sph_validate(const pg_params& P) {
    if (P.ndim != 3) return report_error(PG_E_INVALID, "ndim=%d: not a 3D disk", P.ndim);
    if (P.geomx != 2 || P.geomy != 4 || P.geomz != 5)
        return report_error(PG_E_INVALID, "3D disks use geomx/y/z = 2/4/5 (spherical r, azimuth, polar angle)");
    if (P.gridx < 0 || P.gridx > 1 || P.gridy != 0 || P.gridz != 0)
        return report_error(PG_E_INVALID, "3D grids: gridx 0 or 1, uniform azimuth and polar angle");
    if (P.recon_flag != 5 && P.recon_flag != 6)
        return report_error(PG_E_INVALID, "recon_flag=%d: 3D supports 5 (PEM4) and 6 (PPM4)", P.recon_flag);
    if (P.EOS_flag != 2 || !(P.gam > 1.0))
        return report_error(PG_E_INVALID, "3D disks are adiabatic: EOS_flag 2 with gam > 1");
    if (!P.OrbAdv_flag) return report_error(PG_E_INVALID, "3D disks need OrbAdv_flag 1");
    if (P.dust_flag || P.kill_flag) return report_error(PG_E_INVALID, "3D disks: no dust, no wave killing");
    if (P.n_planet < 0 || P.n_planet > 1) return report_error(PG_E_INVALID, "n_planet=%d", P.n_planet);
    if (P.bound_lft != 0 || P.bound_rgh != 0 || P.bound_bak != 3 || P.bound_frn != 3 || P.bound_bom != 2 ||
        P.bound_top != 2)
        return report_error(PG_E_INVALID,
                            "3D boundaries: radial 0 (analytic), azimuthal 3 (periodic), polar 2 (reflecting)");
    if (P.xres < 4 || P.yres < 4 || P.zres < 4 || P.zres > SPH_ZMAX)
        return report_error(PG_E_INVALID, "3D grid %d x %d x %d: each at least 4, zres at most %d", P.xres, P.yres,
                            P.zres, SPH_ZMAX);
    if ((double)P.xres * P.yres * P.zres >= 2147483648.0)
        return report_error(PG_E_INVALID, "3D grid of 2^31 cells or more");
    if (P.precision != 4 && P.precision != 8) return report_error(PG_E_INVALID, "precision=%d", P.precision);
    if (!(P.zmin > 0.0) || !(P.zmax < kPi) || !(P.zmin < P.zmax))
        return report_error(PG_E_INVALID, "polar range must lie inside (0, pi)");
    return PG_OK;
}

int sph_create(const pg_params& P, Sph** out) {
    *out = nullptr;
    int rc;
    if ((rc = sph_validate(P))) return rc;
    Sph* s = new Sph();
    s->P = P;
    s->nx = P.xres; s->ny = P.yres; s->nz = P.zres;
    s->N = (size_t)P.xres * P.yres * P.zres;
    s->es = P.precision;
    rc = P.precision == 8 ? create_t<double>(s) : create_t<float>(s);
    if (rc) {
        sph_destroy(s);
        return rc;
    }
    *out = s;
    return PG_OK;
}

void sph_destroy(Sph* s) {
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void* p : s->st)
        if (p) (void)hipFree(p);
    if (s->vis) (void)hipFree(s->vis);
    if (s->vis2) (void)hipFree(s->vis2);
    if (s->tab) (void)hipFree(s->tab);
    if (s->dtab) (void)hipFree(s->dtab);
    if (s->rate) (void)hipFree(s->rate);
    if (s->rate_h) (void)hipHostFree(s->rate_h);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

// The analytic state of radial ghost cell g (0..2 left of the disk, 3..5
// right) at polar layer k: rho, p, u, v, w (boundary type 0).
void sph_ghost_x(const Sph* s, int g, int k, double* f5) {
    for (int q = 0; q < NF; ++q) f5[q] = s->xgh[((size_t)q * 6 + g) * s->nz + k];
}
void sph_dims(const Sph* s, int* nx, int* ny, int* nz) {
    *nx = s->nx; *ny = s->ny; *nz = s->nz;
}

int sph_get_state(Sph* s, double* gas) {
    if (!gas) return report_error(PG_E_INVALID, "null state");
    return s->es == 8 ? get_state<double>(s, gas) : get_state<float>(s, gas);
}
int sph_set_state(Sph* s, const double* gas) {
    if (!gas) return PG_OK;
    return s->es == 8 ? put_state<double>(s, gas) : put_state<float>(s, gas);
}

int sph_global_dt(Sph* s, double dt_prev, double* dt_out) {   // timestep.cu:99-140
    double rate = 0.0;
    const int rc = s->es == 8 ? rate_of<double>(s, 0, &rate) : rate_of<float>(s, 0, &rate);
    if (rc) return rc;
    if (!(rate > 0.0) || !std::isfinite(rate)) return report_error(PG_E_STATE, "CFL rate %g", rate);
    const double dt = s->P.CFL / rate, cap = dt_prev * 1.1;
    *dt_out = dt < cap ? dt : cap;
    return PG_OK;
}

int sph_solve(Sph* s, double t, double dt) {
    const int rc = s->es == 8 ? ds_step<double>(s, t, dt) : ds_step<float>(s, t, dt);
    if (rc) return rc;
    SPH_TRY(hipStreamSynchronize(s->stream));
    return PG_OK;
}

// The loop stops at end_time as the reference's does (main.cu:323-327): the
// steps left after t reaches it take dt = 0 and change nothing, as in the
// 2D device loop.
int sph_run(Sph* s, int nsteps, double* t_io, double* dt_io, double* dts) {   // main.cu:320-359
    double t = *t_io, dt = *dt_io;
    for (int n = 0; n < nsteps; ++n) {
        int rc;
        if (!(t < s->P.end_time)) {
            dt = 0.0;
            if (dts) dts[n] = 0.0;
            continue;
        }
        if ((rc = sph_global_dt(s, dt, &dt))) return rc;
        if (t + dt > s->P.end_time) dt = s->P.end_time - t;
        rc = s->es == 8 ? ds_step<double>(s, t, dt) : ds_step<float>(s, t, dt);
        if (rc) return rc;
        t += dt;
        if (dts) dts[n] = dt;
    }
    SPH_TRY(hipStreamSynchronize(s->stream));
    *t_io = t;
    *dt_io = dt;
    return PG_OK;
}

int sph_sweep(Sph* s, int dir, double dt) {
    if (dir < 0 || dir > 2) return report_error(PG_E_INVALID, "sweep direction %d", dir);
    s->rate_valid = false;
    int rc = s->es == 8 ? launch_sweep<double>(s, dir, 0, 1, dt) : launch_sweep<float>(s, dir, 0, 1, dt);
    if (rc) return rc;
    SPH_TRY(hipMemcpyAsync(s->st[0], s->st[1], (size_t)NF * s->N * s->es, hipMemcpyDeviceToDevice, s->stream));
    SPH_TRY(hipStreamSynchronize(s->stream));
    return PG_OK;
}

int sph_orbital_advection(Sph* s, double dt, int* inc, double* res) {
    s->rate_valid = false;
    const size_t nik = (size_t)s->nx * s->nz;
    int* inc_d = nullptr;
    double* res_d = nullptr;
    if (inc || res) {
        SPH_TRY(hipMalloc((void**)&inc_d, nik * sizeof(int)));
        SPH_TRY(hipMalloc((void**)&res_d, nik * sizeof(double)));
    }
    int rc = s->es == 8 ? launch_orbadv<double>(s, 0, 1, dt, inc_d, res_d)
                        : launch_orbadv<float>(s, 0, 1, dt, inc_d, res_d);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(s->st[0], s->st[1], (size_t)NF * s->N * s->es, hipMemcpyDeviceToDevice,
                                      s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess && inc) e = hipMemcpy(inc, inc_d, nik * sizeof(int), hipMemcpyDeviceToHost);
        if (e == hipSuccess && res) e = hipMemcpy(res, res_d, nik * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = report_error(PG_E_HIP, "orbital advection: %s", hipGetErrorString(e));
    }
    if (inc_d) (void)hipFree(inc_d);
    if (res_d) (void)hipFree(res_d);
    return rc;
}

int sph_sources(Sph* s, double t_planet, double mdt, double dt, int inplace) {
    s->rate_valid = false;
    Body b = s->planet;
    if (s->P.n_planet > 0) orbit_solution(b, t_planet, s->P.ramp_time, s->P.planet_mass, s->P.frame_omega);
    const int out = inplace ? 0 : 2;
    int rc = s->es == 8 ? launch_sources<double>(s, 0, out, b, mdt, dt) : launch_sources<float>(s, 0, out, b, mdt, dt);
    if (rc) return rc;
    if (!inplace)
        SPH_TRY(hipMemcpyAsync(s->st[0], s->st[2], (size_t)NF * s->N * s->es, hipMemcpyDeviceToDevice, s->stream));
    SPH_TRY(hipStreamSynchronize(s->stream));
    return PG_OK;
}

int sph_viscosity(Sph* s, int on_T, double* vis) {
    if (on_T) return report_error(PG_E_INVALID, "3D viscosity module: evaluates the current state");
    int rc = s->es == 8 ? launch_vis<double>(s, 0) : launch_vis<float>(s, 0);
    if (rc) return rc;
    SPH_TRY(hipStreamSynchronize(s->stream));
    if (!vis) return PG_OK;
    const size_t n = 2 * s->N;
    if (s->es == 8) {
        SPH_TRY(hipMemcpy(vis, s->vis, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> h(n);
        SPH_TRY(hipMemcpy(h.data(), s->vis, n * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t m = 0; m < n; ++m) vis[m] = h[m];
    }
    return PG_OK;
}

int sph_planet(Sph* s, double t, double* body10) {
    Body b = s->planet;
    if (s->P.n_planet > 0) orbit_solution(b, t, s->P.ramp_time, s->P.planet_mass, s->P.frame_omega);
    const double f[10] = {b.m, b.a, b.e, b.x, b.y, b.z, b.vx, b.vy, b.vz, b.rs};
    std::memcpy(body10, f, sizeof(f));
    return PG_OK;
}

// write_check_point (output.cu:125-180), ndim 3: t, xa, ya, za, rho, p, u, v, w
int sph_write_snapshot(Sph* s, const char* filename, double t) {
    std::vector<double> g((size_t)NF * s->N);
    int rc = sph_get_state(s, g.data());
    if (rc) return rc;
    FILE* f = std::fopen(filename, "wb");
    if (!f) return report_error(PG_E_INVALID, "cannot open %s", filename);
    bool ok = std::fwrite(&t, sizeof(double), 1, f) == 1;
    ok = ok && std::fwrite(&s->xa[NPAD], sizeof(double), s->nx + 1, f) == (size_t)s->nx + 1;
    ok = ok && std::fwrite(&s->ya[NPAD], sizeof(double), s->ny + 1, f) == (size_t)s->ny + 1;
    ok = ok && std::fwrite(&s->za[NPAD], sizeof(double), s->nz + 1, f) == (size_t)s->nz + 1;
    ok = ok && std::fwrite(g.data(), sizeof(double), g.size(), f) == g.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return report_error(PG_E_INVALID, "short write to %s", filename);
    return PG_OK;
}

int sph_read_snapshot(Sph* s, const char* filename, double* t_out) {   // load_grid (output.cu:182-248)
    FILE* f = std::fopen(filename, "rb");
    if (!f) return report_error(PG_E_INVALID, "cannot open %s", filename);
    const size_t nax = (size_t)s->nx + s->ny + s->nz + 3;
    std::vector<double> head(1 + nax), g((size_t)NF * s->N);
    bool ok = std::fread(head.data(), sizeof(double), head.size(), f) == head.size();
    ok = ok && std::fread(g.data(), sizeof(double), g.size(), f) == g.size();
    std::fclose(f);
    if (!ok) return report_error(PG_E_INVALID, "%s: truncated snapshot", filename);
    if (t_out) *t_out = head[0];
    return sph_set_state(s, g.data());
}

int sph_profile_enable(Sph* s, int on) {
    s->profile = on != 0;
    if (on) s->timers.clear();
    return PG_OK;
}
int sph_profile_get(Sph* s, const char* kernel, double* avg_us, int* count) {
    auto it = s->timers.find(kernel ? kernel : "");
    if (it == s->timers.end()) {
        if (avg_us) *avg_us = 0.0;
        if (count) *count = 0;
        return PG_OK;
    }
    if (avg_us) *avg_us = it->second.second ? it->second.first / it->second.second : 0.0;
    if (count) *count = it->second.second;
    return PG_OK;
}
int sph_device_bytes(Sph* s, size_t* bytes) {
    *bytes = s->bytes;
    return PG_OK;
}
void* sph_stream(Sph* s) { return (void*)s->stream; }

}  // namespace pg
