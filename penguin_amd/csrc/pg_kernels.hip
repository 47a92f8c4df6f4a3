// pg_kernels.hip -- gfx950 kernels of the Godunov step (2D, locally
// isothermal gas + pressureless dust, PPM4 + HLLC).
//
// Every kernel is templated on the storage type `real` (float or double);
// arithmetic runs in that type.  Data layout and tables: pg_internal.h.
//
// Kernels and the reference routines they compute:
//   k_sweep_x_gas   sweepx_inplace  (sweeps.cu:263-358) + riemann()
//   k_sweep_x_dust  sweepx_dust_inplace (dust.cu:178-257)
//   k_phi           sweepy_inplace (sweeps.cu:361-456) or the dust
//                   azimuthal sweep (dust.cu:260-353), then the orbital
//                   shift (orbit.cu:17-49) and the residual advection
//                   (advection.cu:1-319), fused over a tile of radii x rows:
//                   the shift becomes a per-radius offset of the tile's
//                   loads, so nothing but the tile is held in LDS
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

#include "pg_device.h"
#include "pg_internal.h"

namespace pg {

// ------------------------------------------------------- source terms
// 1 - exp(-x) (the drag factor dtau, sweeps.cu:673,702), as -expm1(-x):
// 1 - exp(-x) formed by a subtraction loses the leading digits when x is
// small (a long stopping time or a short step)
template <typename real>
__device__ __forceinline__ real one_minus_exp_neg(real x);
template <>
__device__ __forceinline__ float one_minus_exp_neg(float x) { return -expm1f(-x); }
template <>
__device__ __forceinline__ double one_minus_exp_neg(double x) { return -expm1(-x); }

// Source terms of one cell (apply_source_terms, sweeps.cu:555-711; the
// dust does not act back on the gas).  mode 0: the first half kick with the
// planet at t; 1/2: predictor / corrector with the planet at t+dt, the
// velocity-dependent forces evaluated at the state predicted over h.
// h64 = dt/2.  Gas inputs rho, cu, cvp (vp = deviation from vbg), viscous
// forces Vr, Vp; dust Qu, Qvp.  GAS / DUST select the outputs computed.
// Used by k_sources and by the x sweeps, which apply the first half kick to
// the cells they load.
// The per-radius table values the source terms of one cell read.
template <typename real>
struct SrcTab {
    real R, iR, vbg, ebal, ts, its;
};
template <typename real>
__device__ __forceinline__ SrcTab<real> src_tab(const DiskArgs<real>& d, int ip) {
    return {TAB(T_XC, ip), TAB(T_IR, ip), TAB(T_VBG, ip), TAB(T_EBAL, ip), TAB(T_TS, ip), TAB(T_ITS, ip)};
}
template <typename real, bool GAS, bool DUST>
__device__ __forceinline__ void cell_sources(const DiskArgs<real>& d, int mode, const SrcTab<real>& tb, int j,
                                             double h64, real rho, real cu, real cvp, real Vr, real Vp, real Qu,
                                             real Qvp, real& nu, real& nvp, real& ndu, real& ndvp) {
    const real h = (real)h64;
    const real mdt = mode == 0 ? real(0) : h;
    const PlanetK& pk = mode == 0 ? d.ctl->pk0 : d.ctl->pk1;
    const real R = tb.R;
    const real iR = tb.iR;
    const real vbg = tb.vbg;
    const real ebal = tb.ebal;
    real px = real(0), py = real(0);
    if (d.n_planet > 0) planet_acc(pk, d.phitab + 4 * j, R, px, py);
    const real irho = FM<real>::rcp(rho);
    real vx = real(0), vy = real(0);
    if (d.visc) { vx = Vr * irho; vy = Vp * irho; }
    const real ay = vy + py;               // fy + gy, independent of the state
    real Tu = cu, Tvp = cvp;
    real Du = Qu, Dvp = Qvp, ts = 0, its = 0;
    if (DUST) {
        ts = tb.ts * irho;                    // St / (c_s/h0) / rho (dust.cu:1-4)
        its = rho * tb.its;
    }
    // fx + gx at azimuthal velocity vbg + vp: centrifugal - stellar
    // gravity (tabulated at vbg) + deviation part + planet (+ viscosity)
    auto axd = [&](real vp) { return ebal + vp * (real(2) * vbg + vp) * iR + px; };
    auto axg = [&](real vp) { return axd(vp) + vx; };
    if (mode == 1 || mode == 2) {
        Tu = cu + axg(cvp) * mdt;
        Tvp = cvp + ay * mdt;
        if (DUST) {
            const real dtau = one_minus_exp_neg(mdt * its);
            Du = Qu + (axd(Dvp) * ts + cu - Qu) * dtau;
            Dvp = Qvp + ((cvp - Qvp) + py * ts) * dtau;
        }
    }
    if (GAS) {
        nu = cu + axg(Tvp) * h;
        nvp = cvp + ay * h;
    }
    if (DUST) {
        const real dtau = one_minus_exp_neg(h * its);
        ndu = Qu + (axd(Dvp) * ts + Tu - Qu) * dtau;
        ndvp = Qvp + ((Tvp - Qvp) + py * ts) * dtau;
    }
}

// Threads of an x-sweep block for a tile of TX cells: TX itself for the
// tiles that are whole waves (the first NPAD*2 threads then take a second
// slot in each phase), or the TX + 2 NPAD slots rounded up to whole waves
// for the tiles sized to fill them (122 + 6 = two waves, 250 + 6 = four,
// 506 + 6 = eight: one slot per thread, no phase run twice by one wave).
// A 1024-cell tile has 512 threads, two slots each (the first NPAD*2 a
// third): a 1024-radius row in one workgroup, the per-wave set-up shared by
// two cells, and no workgroup for the few radii a 506-cell tiling leaves.
template <int TX>
struct XBlock {
    static constexpr int BT = TX == 1024 ? 512 : TX % 64 == 0 ? TX : (TX + 2 * NPAD + 63) / 64 * 64;
    static constexpr int S = TX + 2 * NPAD;                                // LDS slots
    static constexpr int NFX = (S - 2 - NPAD + BT - 1) / BT;               // interfaces per thread, at most
};
// LDS arrays of S slots: the gas tile 12, the dust tile 10 (their fluxes
// are held in registers until the slopes are consumed, then stored over
// the slope arrays)
constexpr int XLDS_GAS = 12, XLDS_DUST = 10;

// ------------------------------------------------------------ x sweep, gas
// One block = TX interior cells of one azimuthal row (the whole row when
// nr <= 1024) plus 3 ghost/halo cells each side, staged in LDS; thread t
// owns slot t (interior cell t-3 of the tile).  Phases (barrier between):
// load -> flattening and parabolas -> interface fluxes -> conservative
// update (a cell's flattening coefficient is used by its own parabolas
// only, PPM.cu:64-78, so the two share a phase).
// Density and pressure are held as deviations from the radial background
// (rho_b, p_b = c_s^2 rho_b); their differences across an edge are the
// deviations' difference plus the tabulated background difference, and
// the interface pressures are deviations from the mean of the backgrounds
// on either side, whose own difference is the tabulated T_PSRC term.
// TX is a compile-time constant so every LDS array offset is an immediate.
// kick != 0: apply the first half kick (cell_sources mode 0: planet at t,
// the viscous forces of the step start) to every interior cell as it is
// loaded, in place of a separate source-term pass.
template <typename real, int TX>
__device__ __forceinline__ void sweep_x_gas_tile(const LaunchSet<real>& ls, int2 w, int par, int kick, real* sm) {
    const DiskArgs<real>& d = ls.disks[w.x];
    constexpr int S = XBlock<TX>::S;
    real* s_r = sm;             // rho - rho_b (p - p_b = c_s^2 (rho - rho_b) is formed where it is read)
    real* s_u = sm + S;
    real* s_L = sm + 2 * S;
    real* s_vp = sm + 3 * S;
    real* s_sl = sm + 4 * S;    // 4 fields: r, p, u, L
    real* s_sr = sm + 8 * S;
    real* s_Fr = s_sl;          // the fluxes, once the parabolas are consumed
    real* s_Fu = s_sl + S;
    real* s_FL = s_sl + 2 * S;
    real* s_pr = s_sl + 3 * S;  // interface pressure - its background reference
    const int nx = d.nx;
    const size_t N = (size_t)nx * d.ny;
    const int t = threadIdx.x;
    const int tiles = (nx + TX - 1) / TX;
    const int j = w.y / tiles;
    const int x0 = (w.y - j * tiles) * TX;
    const int nt = min(TX, nx - x0);
    const int ns = nt + 2 * NPAD;
    // a wave with no slot in any phase (the last, narrow tile of a row)
    // ends here: a barrier waits on the waves of the workgroup that have
    // not ended, and the ended waves free their slots on the CU for other
    // workgroups (the condition is uniform over each wave)
    if ((t & ~63) >= ns) return;
    const real dt = (real)d.ctl->dt;
    const PG_GLOBAL real* __restrict__ in = gmem(d.gas[gsel(par)]);
    const size_t row = (size_t)j * nx;
    // The narrow tiles of one slot per thread in every phase (122, 250):
    // the per-radius table values of the thread's slots are loaded here,
    // with the state, so that no later phase waits on a table load.  The
    // thread's parabola slot is radius ia = x0 + t + 2, its interface and
    // its update slot ia + 1 (indices ia - 2 .. ia + 2, inside the padded
    // tables for every slot of the tile).  Measured: the x sweep of 64
    // uneven disks on 122-radius tiles 321 -> 282 us; the 506-radius tiles
    // of a 1024-radius disk, whose eight waves hide the table loads behind
    // each other, 91 -> 99 us with the 24 more VGPRs, so they load as they go.
    constexpr bool ONE = XBlock<TX>::BT >= XBlock<TX>::S && TX <= 250;
    real hw[6] = {}, hcs[5] = {}, htp[4] = {}, htr[4] = {}, hgr[2] = {}, hc0[2] = {}, hix[2] = {}, hxa[2] = {};
    real hixv = real(0), hps = real(0), hir = real(0), hvb = real(0);
    if (ONE && t < ns - 4) {
        const int ia = x0 + t + 2;
        hw[0] = TAB(T_WL0, ia); hw[1] = TAB(T_WL1, ia); hw[2] = TAB(T_WL2, ia);
        hw[3] = TAB(T_WR0, ia); hw[4] = TAB(T_WR1, ia); hw[5] = TAB(T_WR2, ia);
#pragma unroll
        for (int m = 0; m < 5; ++m) hcs[m] = TAB(T_CS2, ia - 2 + m);
#pragma unroll
        for (int m = 0; m < 4; ++m) { htp[m] = TAB(T_TBP, ia - 1 + m); htr[m] = TAB(T_TBR, ia - 1 + m); }
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            hgr[m] = TAB(T_G_R, ia + m); hc0[m] = TAB(T_C0, ia + m); hix[m] = TAB(T_IDX, ia + m);
            hxa[m] = TAB(T_XA, ia + 1 + m);
        }
        hixv = TAB(T_IXV, ia + 1); hps = TAB(T_PSRC, ia + 1); hir = TAB(T_IR, ia + 1); hvb = TAB(T_VBG, ia + 1);
    }

    for (int s = t; s < ns; s += XBlock<TX>::BT) {
        const int i = x0 - NPAD + s;
        const int ip = i + NPAD;
        real rd, u, vp;
        if (i < 0 || i >= nx) {
            rd = real(0); u = TAB(T_G_U, ip); vp = TAB(T_G_V, ip);   // the background itself
        } else {
            rd = in[row + i]; u = in[N + row + i]; vp = in[2 * N + row + i];
            if (kick) {
                real Vr = real(0), Vp = real(0), nu, nvp, unused0, unused1;
                if (d.visc) { Vr = gmem(d.vis)[row + i]; Vp = gmem(d.vis)[N + row + i]; }
                cell_sources<real, true, false>(d, 0, src_tab(d, ip), j, 0.5 * d.ctl->dt, TAB(T_G_R, ip) + rd, u, vp,
                                                Vr, Vp, real(0), real(0), nu, nvp, unused0, unused1);
                u = nu;
                vp = nvp;
            }
        }
        s_r[s] = rd;
        s_u[s] = u;
        s_vp[s] = vp;
        s_L[s] = TAB(T_XC, ip) * (TAB(T_VBG, ip) + vp);
    }
    __syncthreads();
    if (d.out_kick) {   // the kicked state of the tile's own cells
        PG_GLOBAL real* k1 = gmem(d.out_kick);
        for (int s = t + NPAD; s < ns - NPAD; s += XBlock<TX>::BT) {
            const size_t o = row + x0 - NPAD + s;
            k1[o] = s_r[s]; k1[N + o] = s_u[s]; k1[2 * N + o] = s_vp[s];
        }
    }
    for (int s = t + 2; s < ns - 2; s += XBlock<TX>::BT) {
        const int ip = x0 + s;
        const real wl0 = ONE ? hw[0] : TAB(T_WL0, ip), wl1 = ONE ? hw[1] : TAB(T_WL1, ip);
        const real wl2 = ONE ? hw[2] : TAB(T_WL2, ip), wr0 = ONE ? hw[3] : TAB(T_WR0, ip);
        const real wr1 = ONE ? hw[4] : TAB(T_WR1, ip), wr2 = ONE ? hw[5] : TAB(T_WR2, ip);
        real cs[5], tp[4], tr[4];
#pragma unroll
        for (int m = 0; m < 5; ++m) cs[m] = ONE ? hcs[m] : TAB(T_CS2, ip - 2 + m);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            tp[m] = ONE ? htp[m] : TAB(T_TBP, ip - 1 + m);
            tr[m] = ONE ? htr[m] : TAB(T_TBR, ip - 1 + m);
        }
        // differences across the edges ip-1 .. ip+2 (edge e between cells e-1, e)
        const real p0 = cs[0] * s_r[s - 2], p1 = cs[1] * s_r[s - 1];
        const real p2 = cs[2] * s_r[s], p3 = cs[3] * s_r[s + 1];
        const real p4 = cs[4] * s_r[s + 2];
        const real dp0 = tp[0] + (p1 - p0), dp1 = tp[1] + (p2 - p1);
        const real dp2 = tp[2] + (p3 - p2), dp3 = tp[3] + (p4 - p3);
        const real r0 = s_r[s - 2], r1 = s_r[s - 1], r2 = s_r[s], r3 = s_r[s + 1], r4 = s_r[s + 2];
        const real dr0 = tr[0] + (r1 - r0), dr1 = tr[1] + (r2 - r1);
        const real dr2 = tr[2] + (r3 - r2), dr3 = tr[3] + (r4 - r3);
        const real pc = cs[2] * (ONE ? hgr[0] : TAB(T_G_R, ip)) + p2;
        const real fl = flatten_d(dp2, dp1 + dp2 + dp3, dp1, dp0 + dp1 + dp2, pc, s_u[s - 1], s_u[s], s_u[s + 1]);
        real l, r;
        slopes_d(dr0, dr1, dr2, dr3, wl0, wl1, wl2, wr0, wr1, wr2, fl, l, r);
        s_sl[s] = l; s_sr[s] = r;
        slopes_d(dp0, dp1, dp2, dp3, wl0, wl1, wl2, wr0, wr1, wr2, fl, l, r);
        s_sl[S + s] = l; s_sr[S + s] = r;
        slopes(s_u[s - 2], s_u[s - 1], s_u[s], s_u[s + 1], s_u[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, fl, l, r);
        s_sl[2 * S + s] = l; s_sr[2 * S + s] = r;
        slopes(s_L[s - 2], s_L[s - 1], s_L[s], s_L[s + 1], s_L[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, fl, l, r);
        s_sl[3 * S + s] = l; s_sr[3 * S + s] = r;
    }
    __syncthreads();
    real fx[XBlock<TX>::NFX][4];
#pragma unroll
    for (int q = 0; q < XBlock<TX>::NFX; ++q) {
        const int k = t + NPAD + q * XBlock<TX>::BT;
        if (k >= ns - 2) break;
        const int ipk = x0 + k;
        const real rbl = ONE ? hgr[0] : TAB(T_G_R, ipk - 1), rbr = ONE ? hgr[1] : TAB(T_G_R, ipk);
        const real cs2l = ONE ? hcs[2] : TAB(T_CS2, ipk - 1), cs2r = ONE ? hcs[3] : TAB(T_CS2, ipk);
        const real pbl = cs2l * rbl, pbr = cs2r * rbr;
        real rl, rld, pld, ul, vl, rr, rrd, prd, ur, vr, cl, cr;
        trace_left_d(ONE ? hc0[0] : TAB(T_C0, ipk - 1), dt, ONE ? hix[0] : TAB(T_IDX, ipk - 1), rbl, pbl, s_r[k - 1], cs2l * s_r[k - 1], s_u[k - 1],
                     s_sl[k - 1], s_sr[k - 1], s_sl[S + k - 1], s_sr[S + k - 1],
                     s_sl[2 * S + k - 1], s_sr[2 * S + k - 1],
                     s_sl[3 * S + k - 1], s_L[k - 1], s_sr[3 * S + k - 1], rl, rld, pld, ul, vl, cl);
        trace_right_d(ONE ? hc0[1] : TAB(T_C0, ipk), dt, ONE ? hix[1] : TAB(T_IDX, ipk), rbr, pbr, s_r[k], cs2r * s_r[k], s_u[k],
                      s_sl[k], s_sr[k], s_sl[S + k], s_sr[S + k], s_sl[2 * S + k], s_sr[2 * S + k],
                      s_sl[3 * S + k], s_L[k], s_sr[3 * S + k], rr, rrd, prd, ur, vr, cr);
        real Fr, Fu, FL, presd;
        hllc_iso_d(cl, cr, pbl, ONE ? htp[2] : TAB(T_TBP, ipk), rl, pld, ul, vl, rr, prd, ur, vr, Fr, Fu, FL, presd);
        const real g = ONE ? hxa[0] : TAB(T_XA, ipk);  // eval_gfac(1, xa) = edge radius
        fx[q][0] = Fr * g;
        fx[q][1] = Fu * g;
        fx[q][2] = FL * g;
        fx[q][3] = presd;
    }
    __syncthreads();   // every parabola read: the fluxes go over them
#pragma unroll
    for (int q = 0; q < XBlock<TX>::NFX; ++q) {
        const int k = t + NPAD + q * XBlock<TX>::BT;
        if (k >= ns - 2) break;
        s_Fr[k] = fx[q][0];
        s_Fu[k] = fx[q][1];
        s_FL[k] = fx[q][2];
        s_pr[k] = fx[q][3];
    }
    __syncthreads();
    PG_GLOBAL real* __restrict__ out = gmem(d.gas[gsel(par) ^ 1]);
    for (int s = t + NPAD; s < ns - NPAD; s += XBlock<TX>::BT) {
        const int ip = x0 + s;
        const int i = ip - NPAD;
        const real g0 = ONE ? hxa[0] : TAB(T_XA, ip), g1 = ONE ? hxa[1] : TAB(T_XA, ip + 1);
        const real f = dt * (ONE ? hixv : TAB(T_IXV, ip));
        const real Dr = (s_Fr[s] - s_Fr[s + 1]) * f;
        // net_source (riemann.cu:59-62): the deviations' part plus the
        // background's, T_PSRC = (g0 + g1)/2 (pref[ip] - pref[ip+1])
        const real Du = (s_Fu[s] - s_Fu[s + 1] + real(0.5) * (g0 + g1) * (s_pr[s] - s_pr[s + 1]) +
                         (ONE ? hps : TAB(T_PSRC, ip))) * f;
        const real DL = (s_FL[s] - s_FL[s + 1]) * f;
        const real rb = ONE ? hgr[1] : TAB(T_G_R, ip);
        const real rd = s_r[s], u = s_u[s], vp = s_vp[s];
        const real r = rb + rd;
        const real rn = r + Dr;
        real rdn = rd + Dr;
        const real rmin = fmax(r * real(1.0e-10), real(1.0e-10));
        real un = u, vn = vp;
        if (dt == real(0)) {
            rdn = rd;   // a step of size 0 (the loop past end_time) leaves the cell as it is
        } else if (rn < rmin) {   // update_Cell's floor (sweeps.cu:238-246): old velocities kept
            rdn = floor_dev(rmin, rb);
        } else {
            const real irn = FM<real>::rcp(rn);
            // (r u + Du) / (r + Dr) as u + (Du - u Dr) / (r + Dr): the
            // rounding falls on the change, not on the velocity itself
            un = u + (Du - u * Dr) * irn;
            vn = vp + ((DL * (ONE ? hir : TAB(T_IR, ip)) - (ONE ? hvb : TAB(T_VBG, ip)) * Dr) - vp * Dr) * irn;
        }
        out[row + i] = rdn;
        out[N + row + i] = un;
        out[2 * N + row + i] = vn;
    }
}

// ----------------------------------------------------------- x sweep, dust
// kick != 0: apply the first half kick (drag toward the gas and the
// planet's pull, cell_sources mode 0) to every interior cell as it is
// loaded; the drag reads the gas state before its own kick and sweep, the
// gas buffer gpre (the gas x sweep reads it and writes the other one).
template <typename real, int TX>
__device__ __forceinline__ void sweep_x_dust_tile(const LaunchSet<real>& ls, int2 w, int par, int gpre, int kick,
                                                  real* sm) {
    const DiskArgs<real>& d = ls.disks[w.x];
    constexpr int S = XBlock<TX>::S;
    real* s_r = sm;
    real* s_u = sm + S;
    real* s_L = sm + 2 * S;
    real* s_vp = sm + 3 * S;
    real* s_sl = sm + 4 * S;  // 3 fields: r, u, L
    real* s_sr = sm + 7 * S;
    real* s_Fr = s_sl;        // the fluxes, once the parabolas are consumed
    real* s_Fu = s_sl + S;
    real* s_FL = s_sl + 2 * S;
    const int nx = d.nx;
    const size_t N = (size_t)nx * d.ny;
    const int t = threadIdx.x;
    const int tiles = (nx + TX - 1) / TX;
    const int j = w.y / tiles;
    const int x0 = (w.y - j * tiles) * TX;
    const int nt = min(TX, nx - x0);
    const int ns = nt + 2 * NPAD;
    if ((t & ~63) >= ns) return;   // as in the gas tile
    const real dt = (real)d.ctl->dt;
    const PG_GLOBAL real* __restrict__ in = gmem(d.dust[dsel(par)]);
    const size_t row = (size_t)j * nx;
    // the narrow tiles load the table values of the thread's slots here, as
    // the gas tile does (parabola slot ia = x0 + t + 2, interface and update
    // slot ia + 1)
    constexpr bool ONE = XBlock<TX>::BT >= XBlock<TX>::S && TX <= 250;
    real hw[6] = {}, hix[2] = {};
    real hxa = real(0), hixv = real(0), hir = real(0), hvb = real(0);
    if (ONE && t < ns - 4) {
        const int ia = x0 + t + 2;
        hw[0] = TAB(T_WL0, ia); hw[1] = TAB(T_WL1, ia); hw[2] = TAB(T_WL2, ia);
        hw[3] = TAB(T_WR0, ia); hw[4] = TAB(T_WR1, ia); hw[5] = TAB(T_WR2, ia);
        hix[0] = TAB(T_IDX, ia); hix[1] = TAB(T_IDX, ia + 1);
        hxa = TAB(T_XA, ia + 1); hixv = TAB(T_IXV, ia + 1); hir = TAB(T_IR, ia + 1); hvb = TAB(T_VBG, ia + 1);
    }

    for (int s = t; s < ns; s += XBlock<TX>::BT) {
        const int i = x0 - NPAD + s;
        const int ip = i + NPAD;
        real r, u, vp;
        if (i < 0 || i >= nx) {
            r = TAB(T_D_R, ip); u = TAB(T_D_U, ip); vp = TAB(T_D_V, ip);
        } else {
            r = in[row + i]; u = in[N + row + i]; vp = in[2 * N + row + i];
            if (kick) {
                const PG_GLOBAL real* __restrict__ g0 = gmem(d.gas[gpre]);
                const real rho = TAB(T_G_R, ip) + g0[row + i], cu = g0[N + row + i], cvp = g0[2 * N + row + i];
                real ndu, ndvp, unused0, unused1;
                cell_sources<real, false, true>(d, 0, src_tab(d, ip), j, 0.5 * d.ctl->dt, rho, cu, cvp, real(0), real(0),
                                                u, vp, unused0, unused1, ndu, ndvp);
                u = ndu;
                vp = ndvp;
            }
        }
        s_r[s] = r;
        s_u[s] = u;
        s_vp[s] = vp;
        s_L[s] = TAB(T_XC, ip) * (TAB(T_VBG, ip) + vp);
    }
    __syncthreads();
    for (int s = t + 2; s < ns - 2; s += XBlock<TX>::BT) {
        const int ip = x0 + s;
        const real wl0 = ONE ? hw[0] : TAB(T_WL0, ip), wl1 = ONE ? hw[1] : TAB(T_WL1, ip);
        const real wl2 = ONE ? hw[2] : TAB(T_WL2, ip), wr0 = ONE ? hw[3] : TAB(T_WR0, ip);
        const real wr1 = ONE ? hw[4] : TAB(T_WR1, ip), wr2 = ONE ? hw[5] : TAB(T_WR2, ip);
        real l, r;
        slopes(s_r[s - 2], s_r[s - 1], s_r[s], s_r[s + 1], s_r[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, real(0), l, r);
        s_sl[s] = l; s_sr[s] = r;
        slopes(s_u[s - 2], s_u[s - 1], s_u[s], s_u[s + 1], s_u[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, real(0), l, r);
        s_sl[S + s] = l; s_sr[S + s] = r;
        slopes(s_L[s - 2], s_L[s - 1], s_L[s], s_L[s + 1], s_L[s + 2], wl0, wl1, wl2, wr0, wr1, wr2, real(0), l, r);
        s_sl[2 * S + s] = l; s_sr[2 * S + s] = r;
    }
    __syncthreads();
    real fx[XBlock<TX>::NFX][3];
#pragma unroll
    for (int q = 0; q < XBlock<TX>::NFX; ++q) {
        const int k = t + NPAD + q * XBlock<TX>::BT;
        if (k >= ns - 2) break;
        const int ipk = x0 + k;
        real Fr, Fu, FL;
        dust_iface(dt, ONE ? hix[0] : TAB(T_IDX, ipk - 1), s_r[k - 1], s_sl[k - 1], s_sr[k - 1],
                   s_u[k - 1], s_sl[S + k - 1], s_sr[S + k - 1],
                   s_L[k - 1], s_sl[2 * S + k - 1], s_sr[2 * S + k - 1],
                   ONE ? hix[1] : TAB(T_IDX, ipk), s_r[k], s_sl[k], s_sr[k], s_u[k], s_sl[S + k], s_sr[S + k],
                   s_L[k], s_sl[2 * S + k], s_sr[2 * S + k], Fr, Fu, FL);
        const real g = ONE ? hxa : TAB(T_XA, ipk);
        fx[q][0] = Fr * g;
        fx[q][1] = Fu * g;
        fx[q][2] = FL * g;
    }
    __syncthreads();   // every parabola read: the fluxes go over them
#pragma unroll
    for (int q = 0; q < XBlock<TX>::NFX; ++q) {
        const int k = t + NPAD + q * XBlock<TX>::BT;
        if (k >= ns - 2) break;
        s_Fr[k] = fx[q][0];
        s_Fu[k] = fx[q][1];
        s_FL[k] = fx[q][2];
    }
    __syncthreads();
    PG_GLOBAL real* __restrict__ out = gmem(d.dust[dsel(par) ^ 1]);
    const real rfloor = d.D_G_ratio * real(1.0e-10);
    for (int s = t + NPAD; s < ns - NPAD; s += XBlock<TX>::BT) {
        const int ip = x0 + s;
        const int i = ip - NPAD;
        const real f = dt * (ONE ? hixv : TAB(T_IXV, ip));
        const real Dr = (s_Fr[s] - s_Fr[s + 1]) * f;
        const real Du = (s_Fu[s] - s_Fu[s + 1]) * f;
        const real DL = (s_FL[s] - s_FL[s + 1]) * f;
        const real r = s_r[s], u = s_u[s], vp = s_vp[s];
        real rn = r + Dr;
        const real rmin = fmax(r * real(1.0e-10), rfloor);
        real un = u, vn = vp;
        if (dt == real(0)) {
            rn = r;   // a step of size 0 leaves the cell as it is
        } else if (rn < rmin) {
            rn = rmin;
        } else {
            const real irn = FM<real>::rcp(rn);
            // (r u + Du) / (r + Dr) as u + (Du - u Dr) / (r + Dr): the
            // rounding falls on the change, not on the velocity itself
            un = u + (Du - u * Dr) * irn;
            vn = vp + ((DL * (ONE ? hir : TAB(T_IR, ip)) - (ONE ? hvb : TAB(T_VBG, ip)) * Dr) - vp * Dr) * irn;
        }
        out[row + i] = rn;
        out[N + row + i] = un;
        out[2 * N + row + i] = vn;
    }
}

// The x sweeps as kernels: one fluid per launch, or both in one launch
// (blocks [0, n_x) the gas rows, [n_x, 2 n_x) the dust rows), which the
// gas and dust sweeps allow because they write different buffers and the
// dust's drag reads the gas buffer the gas sweep only reads.
template <typename real, int TX>
__global__ __launch_bounds__(XBlock<TX>::BT) void k_sweep_x_gas(LaunchSet<real> ls, int par, int kick) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    sweep_x_gas_tile<real, TX>(ls, ls.wl_x[blockIdx.x], par, kick, reinterpret_cast<real*>(smem_raw));
}
// par: the parity after the gas sweep (the pre-sweep gas is the other buffer)
template <typename real, int TX>
__global__ __launch_bounds__(XBlock<TX>::BT) void k_sweep_x_dust(LaunchSet<real> ls, int par, int kick) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    sweep_x_dust_tile<real, TX>(ls, ls.wl_x[blockIdx.x], par, gsel(par) ^ 1, kick, reinterpret_cast<real*>(smem_raw));
}
template <typename real, int TX>
__global__ __launch_bounds__(XBlock<TX>::BT) void k_sweep_x_both(LaunchSet<real> ls, int par, int kick) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    real* sm = reinterpret_cast<real*>(smem_raw);
    const int b = blockIdx.x;
    if (b < ls.n_x) sweep_x_gas_tile<real, TX>(ls, ls.wl_x[b], par, kick, sm);
    else sweep_x_dust_tile<real, TX>(ls, ls.wl_x[b - ls.n_x], par, gsel(par), kick, sm);
}


// Tile geometry of the azimuthal pass for a thread count: TR radii x TY
// output rows.  The rows of one radius are independent of the other radii,
// so a tile is TR independent azimuthal strips side by side; TR consecutive
// radii make each row of the tile one contiguous segment in HBM.
// With 512 threads an FP32 tile is 8 radii x 128 rows (each thread takes two
// cells of every stage; the 12-row halo is 9 % of the rows swept, against
// 19 % for 16 x 64 in the same 40 KB of LDS: 73 us against 80 on the
// 1024x3072 disk; 4 x 256, with 16-byte row segments in HBM, 96 us); an
// FP64 tile, twice the LDS per cell, 8 x 64.
template <typename real, int B> struct PhiTile;
template <typename real> struct PhiTile<real, 256> { static constexpr int TR = 16, TY = 16; };
template <typename real> struct PhiTile<real, 512> { static constexpr int TR = 8, TY = sizeof(real) == 4 ? 128 : 64; };
template <typename real> struct PhiTile<real, 1024> { static constexpr int TR = 16, TY = 64; };

// Azimuthal pass of one fluid over a tile, fused with the orbital shift and
// the residual advection: sweepy_inplace (sweeps.cu:361-456) or
// sweepy_dust_inplace (dust.cu:260-353), then shift_orbital_advection
// (orbit.cu:17-49) and advecty / advecty_dust (advection.cu:1-357).
//
// Output row y (after the shift) of radius i is the residual advection of
// shifted cells y-3..y+3, i.e. of swept cells y-inc-3..y-inc+3, which need
// input rows y-inc-6..y-inc+6.  So the tile loads, for each radius, TY+12
// input rows starting at y0-6-inc(i): LDS row k of that radius holds input
// row y0-6-inc+k, which the shift carries to row y0-6+k.  Stage A sweeps LDS
// rows [3, TY+9); stage B advects rows [6, TY+6) and writes them to rows
// y0..y0+TY-1, the same rows for every radius, so the stores are contiguous.
// flags: bit 0 = azimuthal sweep, bit 1 = orbital shift + residual advection.
// The gas density is held as its deviation from the ring's background
// rho_b(R), constant along the strip: differences and slopes are those of
// the deviations, the interface pressures deviations from c_s^2 rho_b, and
// the residual advection's mass flux the deviation's (the background's
// uniform flux cancels in every difference).
// LDS of one tile: 9 arrays of E cells (state 3, slopes 6).  A flux stage
// keeps its interface fluxes in registers until every thread has read the
// slopes, then stores them over the slope arrays for the update.
template <typename real, int B> struct PhiLds {
    static constexpr int NK = PhiTile<real, B>::TY + 4 * NPAD;   // LDS rows per radius
    static constexpr int E = NK * PhiTile<real, B>::TR;          // LDS cells per array
    static constexpr int WORDS = 9 * E;
};
template <typename real, bool DUST, int B>
__device__ __forceinline__ void phi_tile(const LaunchSet<real>& ls, int2 w, int par, int flags, real* lds) {
    constexpr int TR = PhiTile<real, B>::TR, TY = PhiTile<real, B>::TY;
    constexpr int NK = PhiLds<real, B>::NK;
    constexpr int E = PhiLds<real, B>::E;
    real* const c0 = lds;
    real* const c1 = lds + E;
    real* const c2 = lds + 2 * E;
    real* const sl0 = lds + 3 * E;
    real* const sr0 = lds + 4 * E;
    real* const sl1 = lds + 5 * E;
    real* const sr1 = lds + 6 * E;
    real* const sl2 = lds + 7 * E;
    real* const sr2 = lds + 8 * E;
    real* const F0 = sl0;   // fluxes, once the slopes are consumed
    real* const F1 = sr0;
    real* const F2 = sl1;
    real* const F3 = sr1;
    if (w.y < 0) return;   // padding entry of the work list (whole block)
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int ntr = (nx + TR - 1) / TR;
    const int ty = w.y / ntr;
    const int i0 = (w.y - ty * ntr) * TR;
    const int y0 = ty * TY;
    const int t = threadIdx.x;
    const bool sweep_on = flags & 1;
    const bool orbadv = (flags & 2) && d.orbadv;
    // per-lane radius: lanes of one wave cover TR radii x (64/TR) rows
    const int ri = t % TR;
    const int i = i0 + ri;
    const bool live = i < nx;
    const int ic = live ? i : nx - 1;          // clamped for table reads
    const int ip = ic + NPAD;
    const real iR = TAB(T_IR, ip);
    const real dt = (real)d.ctl->dt;
    const real dts = dt * iR;                  // set_state / dust_flow receive dt/rad
    const real idy = d.idy;
    const real fup = dt * idy * iR;            // Del.multiply(dt/yv/rad_cyl), sweeps.cu:412
    const real dti = dts * idy;                // dt idx of the traced states
    const int inc = orbadv ? gmem(d.inc)[ic] : 0;
    const PG_GLOBAL real* __restrict__ in = gmem(DUST ? d.dust[dsel(par)] : d.gas[gsel(par)]);
    PG_GLOBAL real* __restrict__ out = gmem(DUST ? d.dust[dsel(par) ^ 1] : d.gas[gsel(par) ^ 1]);
    const real wy[6] = {d.wLy[0], d.wLy[1], d.wLy[2], d.wRy[0], d.wRy[1], d.wRy[2]};
    const real rfloor = d.D_G_ratio * real(1.0e-10);
    const real rbg = DUST ? real(0) : TAB(T_G_R, ip);   // what the stored density is relative to
    // the sound speed of the radius, read once (the gas interfaces use them)
    const real cs2 = DUST ? real(0) : TAB(T_CS2, ip), cs0 = DUST ? real(0) : TAB(T_C0, ip);
    const real ic0 = DUST ? real(0) : TAB(T_IC0, ip);
    // B is a multiple of TR, so lane t keeps radius ri in every loop below
    static_assert(B % TR == 0, "tile");

    // ---- load: TY+12 rows of each radius, shifted back by its inc ----
    const int jbase = y0 - 2 * NPAD - inc;
    const int jb = wrap(jbase + t / TR, ny);   // row of this lane's first element
    constexpr int KSTEP = B / TR;              // rows advanced per loop trip
    int jj = jb;
    for (int e = t; e < E; e += B) {
        real r = DUST ? real(1) : real(0), vp = real(0), u = real(0);
        if (live) {
            const size_t o = (size_t)jj * nx + i;
            r = in[o];
            vp = in[2 * N + o];   // azimuthal: normal to the interfaces
            u = in[N + o];        // radial: transverse here
        }
        c0[e] = r; c1[e] = vp; c2[e] = u;
        jj += KSTEP;
        while (jj >= ny) jj -= ny;   // periodic in azimuth (more than once only when ny < TY+12)
    }
    __syncthreads();
    // ---- stage A: azimuthal sweep of LDS rows [3, TY+9) ----
    for (int e = t + 2 * TR; e < (NK - 2) * TR; e += B) {
        const real a0 = c0[e - 2 * TR], a1 = c0[e - TR], a2 = c0[e], a3 = c0[e + TR], a4 = c0[e + 2 * TR];
        const real d0 = a1 - a0, d1 = a2 - a1, d2 = a3 - a2, d3 = a4 - a3;
        // p = c_s^2 rho with c_s^2 fixed along the strip: flatten's
        // pressure ratios are the density ratios (the dust is not flattened)
        const real f = DUST ? real(0)
                            : flatten_d(d2, d1 + d2 + d3, d1, d0 + d1 + d2, rbg + a2, c1[e - TR], c1[e], c1[e + TR]);
        real l, r;
        slopes_d(d0, d1, d2, d3, wy[0], wy[1], wy[2], wy[3], wy[4], wy[5], f, l, r);
        sl0[e] = l; sr0[e] = r;
        slopes(c1[e - 2 * TR], c1[e - TR], c1[e], c1[e + TR], c1[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4], wy[5], f, l, r);
        sl1[e] = l; sr1[e] = r;
        slopes(c2[e - 2 * TR], c2[e - TR], c2[e], c2[e + TR], c2[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4], wy[5], f, l, r);
        sl2[e] = l; sr2[e] = r;
    }
    __syncthreads();
    // interfaces [NPAD, NK-2) of every radius: at most NFA per thread
    constexpr int FA0 = NPAD * TR, FA1 = (NK - 2) * TR, NFA = (FA1 - FA0 + B - 1) / B;
    real fa[NFA][4];
#pragma unroll
    for (int k = 0; k < NFA; ++k) {
        const int e = t + FA0 + k * B;
        if (e >= FA1) break;
        const int m = e - TR;   // the cell left of interface e
        if (DUST) {
            dust_iface(dts, idy, c0[m], sl0[m], sr0[m], c1[m], sl1[m], sr1[m], c2[m], sl2[m], sr2[m],
                       idy, c0[e], sl0[e], sr0[e], c1[e], sl1[e], sr1[e], c2[e], sl2[e], sr2[e],
                       fa[k][0], fa[k][1], fa[k][2]);
            fa[k][3] = real(0);
        } else {
            real rl, rld, ul, vl, rr, rrd, ur, vr;
            // the pressure parabola is c_s^2 times the density parabola
            trace_left_cd(cs0, ic0, dti, rbg, c0[m], c1[m], sl0[m], sr0[m], sl1[m], sr1[m], sl2[m], c2[m],
                          sr2[m], rl, rld, ul, vl);
            trace_right_cd(cs0, ic0, dti, rbg, c0[e], c1[e], sl0[e], sr0[e], sl1[e], sr1[e], sl2[e], c2[e],
                           sr2[e], rr, rrd, ur, vr);
            hllc_cd(cs0, ic0, cs2, rbg, rl, rld, ul, vl, rr, rrd, ur, vr, fa[k][0], fa[k][1], fa[k][2], fa[k][3]);
        }
    }
    __syncthreads();   // every slope read: the fluxes go over them
#pragma unroll
    for (int k = 0; k < NFA; ++k) {
        const int e = t + FA0 + k * B;
        if (e >= FA1) break;
        F0[e] = fa[k][0]; F1[e] = fa[k][1]; F2[e] = fa[k][2];
        if (!DUST) F3[e] = fa[k][3];
    }
    __syncthreads();
    for (int e = t + NPAD * TR; e < (NK - NPAD) * TR; e += B) {
        const real Dr = (F0[e] - F0[e + TR]) * fup;
        const real Du = DUST ? (F1[e] - F1[e + TR]) * fup
                             : (F1[e] - F1[e + TR] + (F3[e] - F3[e + TR])) * fup;
        const real Dw = (F2[e] - F2[e + TR]) * fup;
        const real rd = c0[e], vp = c1[e], u = c2[e];
        const real r = rbg + rd, rn = r + Dr;   // the density itself
        real rdn = rd + Dr;                     // and what is stored
        const real rmin = fmax(r * real(1.0e-10), DUST ? rfloor : real(1.0e-10));
        real vn = vp, un = u;
        if (!sweep_on || dt == real(0)) {   // no sweep, or a step of size 0
            rdn = rd;
        } else if (rn < rmin) {
            rdn = floor_dev(rmin, rbg);
        } else {
            // gas: (rho v + Du + (rot + R Omega_f) Dr) / rho' - vbg == (rho vp + Du) / rho'
            const real irn = FM<real>::rcp(rn);
            vn = vp + (Du - vp * Dr) * irn;   // (r vp + Du) / rn, in increment form
            un = u + (Dw - u * Dr) * irn;
        }
        c0[e] = rdn; c1[e] = vn; c2[e] = un;   // each thread rewrites only its own cells
    }
    __syncthreads();
    const int rows = min(TY, ny - y0);
    if (!DUST && d.out_swept && live) {
        // LDS row k holds swept row y0 - 2 NPAD - inc + k; rows k in
        // [2 NPAD, 2 NPAD + rows) of every tile cover each row once
        PG_GLOBAL real* sw = gmem(d.out_swept);
        for (int e = t + 2 * NPAD * TR; e < (2 * NPAD + rows) * TR; e += B) {
            const size_t o = (size_t)wrap(y0 - inc + e / TR - 2 * NPAD, ny) * nx + i;
            sw[o] = c0[e]; sw[N + o] = c2[e]; sw[2 * N + o] = c1[e];
        }
    }
    if (!orbadv) {
        for (int e = t + 2 * NPAD * TR; e < (2 * NPAD + TY) * TR; e += B) {
            const int y = e / TR - 2 * NPAD;
            if (!live || y >= rows) continue;
            const size_t o = (size_t)(y0 + y) * nx + i;
            out[o] = c0[e]; out[N + o] = c2[e]; out[2 * N + o] = c1[e];
        }
        return;
    }
    // ---- stage B: residual advection at speed res (advection.cu:1-105) ----
    const real speed = gmem(d.res)[ic];
    const real q = lim01(fabs(speed) * dts * idy);   // |speed| dt / (R dy)
    for (int e = t + (2 * NPAD - 1) * TR; e < (2 * NPAD + TY + 1) * TR; e += B) {
        real l, r;
        slopes(c0[e - 2 * TR], c0[e - TR], c0[e], c0[e + TR], c0[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4], wy[5], real(0), l, r);
        sl0[e] = l; sr0[e] = r;
        slopes(c1[e - 2 * TR], c1[e - TR], c1[e], c1[e + TR], c1[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4], wy[5], real(0), l, r);
        sl1[e] = l; sr1[e] = r;
        slopes(c2[e - 2 * TR], c2[e - TR], c2[e], c2[e + TR], c2[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4], wy[5], real(0), l, r);
        sl2[e] = l; sr2[e] = r;
    }
    __syncthreads();
    constexpr int FB0 = 2 * NPAD * TR, FB1 = (2 * NPAD + TY + 1) * TR, NFB = (FB1 - FB0 + B - 1) / B;
    real fb[NFB][3];
#pragma unroll
    for (int k = 0; k < NFB; ++k) {
        const int e = t + FB0 + k * B;
        if (e >= FB1) break;
        real Dr, Dv, Du;   // Dr: upwind mean of the stored density (gas: its deviation)
        if (speed > real(0)) {
            const int m = e - TR;
            Dr = aveR(q, sl0[m], c0[m], sr0[m]);
            const real ra = rbg + Dr;
            Dv = aveR(q, sl1[m], c1[m], sr1[m]) * ra;
            Du = aveR(q, sl2[m], c2[m], sr2[m]) * ra;
        } else {
            Dr = aveL(q, sl0[e], c0[e], sr0[e]);
            const real ra = rbg + Dr;
            Dv = aveL(q, sl1[e], c1[e], sr1[e]) * ra;
            Du = aveL(q, sl2[e], c2[e], sr2[e]) * ra;
        }
        fb[k][0] = Dr * speed; fb[k][1] = Dv * speed; fb[k][2] = Du * speed;
    }
    __syncthreads();   // every slope read: the fluxes go over them
#pragma unroll
    for (int k = 0; k < NFB; ++k) {
        const int e = t + FB0 + k * B;
        if (e >= FB1) break;
        F0[e] = fb[k][0]; F1[e] = fb[k][1]; F2[e] = fb[k][2];
    }
    __syncthreads();
    for (int e = t + 2 * NPAD * TR; e < (2 * NPAD + TY) * TR; e += B) {
        const int y = e / TR - 2 * NPAD;
        if (!live || y >= rows) continue;
        const real Dr = (F0[e] - F0[e + TR]) * fup;
        const real Dv = (F1[e] - F1[e + TR]) * fup;
        const real Du = (F2[e] - F2[e + TR]) * fup;
        const real rd = c0[e], r = rbg + rd;
        real rn = rd + Dr;   // stored form
        real vn = c1[e], un = c2[e];
        // update_Dust floors (advection.cu:315); update_Cell_advect has none
        // (advection.cu:88-105)
        if (dt == real(0)) {
            rn = rd;   // a step of size 0 moves nothing
        } else if (DUST && rn < fmax(r * real(1.0e-10), rfloor)) {
            rn = fmax(r * real(1.0e-10), rfloor);
        } else {
            const real irn = FM<real>::rcp(r + Dr);
            vn = c1[e] + (Dv - c1[e] * Dr) * irn;   // (r v + Dv) / (r + Dr), in increment form
            un = c2[e] + (Du - c2[e] * Dr) * irn;
        }
        const size_t o = (size_t)(y0 + y) * nx + i;
        out[o] = rn;
        out[N + o] = un;
        out[2 * N + o] = vn;
    }
}

// The azimuthal pass of one fluid, or of both in one launch (blocks
// [0, n_ring) the gas tiles, [n_ring, 2 n_ring) the dust tiles; the two
// fluids read and write different buffers).
template <typename real, bool DUST, int B>
__global__ __launch_bounds__(B) void k_phi(LaunchSet<real> ls, int par, int flags) {
    __shared__ real lds[PhiLds<real, B>::WORDS];
    phi_tile<real, DUST, B>(ls, ls.wl_ring[blockIdx.x], par, flags, lds);
}
template <typename real, int B>
__global__ __launch_bounds__(B) void k_phi_both(LaunchSet<real> ls, int par, int flags) {
    __shared__ real lds[PhiLds<real, B>::WORDS];
    const int b = blockIdx.x;
    if (b < ls.n_ring) phi_tile<real, false, B>(ls, ls.wl_ring[b], par, flags, lds);
    else phi_tile<real, true, B>(ls, ls.wl_ring[b - ls.n_ring], par, flags, lds);
}

// ----------------------------------------------------------------- viscosity
// Pointwise and stencil kernels work on tiles of PT_X radii x TH rows (the
// launch set's pt_rows as a template parameter); thread t covers radius
// t % PT_X of rows t / PT_X + PT_TY * k.  The tile id of a work-list entry
// is ty * ntx + tx.
static_assert(PT_X * PT_TY == 256 && PT_YMAX % PT_TY == 0, "pointwise tile");
__device__ __forceinline__ int pt_ntx(int nx) { return (nx + PT_X - 1) / PT_X; }

// 9-point stencil of viscous_forces (viscosity.cu:34-197, ndim 2) at the cell
// of padded radius ip from its neighbourhood (n = 3b + a: radius i-1+a, row
// j, j-1, j+1 for b = 0, 1, 2) of rho*nu, u and vp.  The derivatives of
// v_phi = vbg(R) + vp are split into the tabulated background part and the
// stored deviation (the difference operators are linear, so this is the
// same arithmetic rearranged).
// The per-radius table values of the stencil: reciprocal differences of the
// centre grid (viscosity.cu:9-31), 1/R and the background rotation's
// value and derivatives.
template <typename real>
struct VisTab {
    real i20, i21, i10, ir, dvbg, d2vbg, vbg;
};
template <typename real>
__device__ __forceinline__ VisTab<real> vis_tab(const DiskArgs<real>& d, int ip) {
    return {TAB(T_I20, ip), TAB(T_I21, ip), TAB(T_I10, ip), TAB(T_IR, ip),
            TAB(T_DVBG, ip), TAB(T_D2VBG, ip), TAB(T_VBG, ip)};
}
template <typename real>
__device__ __forceinline__ void vis_cell(const DiskArgs<real>& d, const VisTab<real>& tb, const real* nr,
                                         const real* u, const real* v, real& fx_out, real& fy_out) {
    const real i20 = tb.i20, i21 = tb.i21, i10 = tb.i10;
    const real ir = tb.ir;
    const real idy = d.idy;
    // scale factors of the differences, per radius (formed once per thread
    // when the stencil runs over several rows of one radius)
    const real hir = real(0.5) * idy * ir;        // 1/((t2 - t0) R)
    const real yy = (idy * idy) * (ir * ir);      // 1/(dy R)^2
    const real xy = i20 * hir;
    const real i20x2 = real(2) * i20;
    // radial derivatives
    const real dudx = (u[2] - u[0]) * i20;
    const real d2udx2 = ((u[2] - u[1]) * i21 - (u[1] - u[0]) * i10) * i20x2;
    const real dvdx = tb.dvbg + (v[2] - v[0]) * i20;
    const real d2vdx2 = tb.d2vbg + ((v[2] - v[1]) * i21 - (v[1] - v[0]) * i10) * i20x2;
    const real dndx = (nr[2] - nr[0]) * i20;
    // azimuthal and mixed derivatives
    const real dvdy = (v[7] - v[4]) * hir;
    const real d2vdy2 = (v[7] - real(2) * v[1] + v[4]) * yy;
    const real d2vdxy = (v[8] - v[6] - v[5] + v[3]) * xy;
    const real dudy = (u[7] - u[4]) * hir;
    const real d2udy2 = (u[7] - real(2) * u[1] + u[4]) * yy;
    const real d2udxy = (u[8] - u[6] - u[5] + u[3]) * xy;
    const real dndy = (nr[7] - nr[4]) * hir;
    // stresses sxx = 2 n dudx, syy = 2 n (dvdy + u/R), sxy = n (dudy + dvdx -
    // v/R) and their divergence in polar coordinates, the terms over R
    // gathered
    const real n1 = nr[1], n2 = real(2) * n1;
    const real vir = (tb.vbg + v[1]) * ir, uir = u[1] * ir;
    const real t = dudy + dvdx - vir;
    const real sxx = n2 * dudx, syy = n2 * (dvdy + uir), sxy = n1 * t;
    const real fx = real(2) * dndx * dudx + n2 * d2udx2 + dndy * t + n1 * (d2udy2 + d2vdxy) +
                    ir * (sxx - n1 * dvdy - syy);
    const real fy = real(2) * dndy * (dvdy + uir) + n2 * d2vdy2 + dndx * t + n1 * (d2udxy + d2vdx2) +
                    ir * (n2 * dudy - n1 * dvdx + n1 * vir + real(2) * sxy);
    fx_out = fx;
    fy_out = fy;
}

// which = 0: viscosity of the current state; 1: of the other buffer (the
// predictor, viscosity_tensor_evaluation2).  Result in d.vis.
template <typename real, int TH>
__global__ __launch_bounds__(256) void k_viscosity(LaunchSet<real> ls, int par, int which) {
    const int2 w = ls.wl_pt[blockIdx.x];
    if (w.y < 0) return;   // padding entry of the work list (whole block)
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int ntx = pt_ntx(nx);
    const int ty = w.y / ntx;
    const int i = (w.y - ty * ntx) * PT_X + (int)(threadIdx.x % PT_X);
    const PG_GLOBAL real* __restrict__ st = gmem(d.gas[gsel(par) ^ which]);
    if (i >= nx) return;
    const int ip = i + NPAD;
    real nu[3];
    for (int n = 0; n < 3; ++n) nu[n] = TAB(T_NU, ip - 1 + n);
#pragma unroll
    for (int k = 0; k < TH / PT_TY; ++k) {
        const int j = ty * TH + (int)(threadIdx.x / PT_X) + PT_TY * k;
        if (j >= ny) break;
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
                    rr = TAB(T_G_R, ii + NPAD) + st[o]; uu = st[N + o]; vv = st[2 * N + o];
                }
                nr[n] = rr * nu[a];
                u[n] = uu;
                v[n] = vv;  // deviation vp
            }
        }
        real fx, fy;
        vis_cell(d, vis_tab(d, ip), nr, u, v, fx, fy);
        size_t o = (size_t)j * nx + i;
        PG_GLOBAL real* vis = gmem(d.vis);
        vis[o] = fx;
        vis[N + o] = fy;
    }
}

// ------------------------------------------------------------- sources

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

// CFL rate of one cell (get_dt_lv1, timestep.cu:18-97): the largest signal
// speed over the cell width in each direction, the dust's included.  A
// velocity that is not finite gives +inf, so that a broken state is not lost
// in the maxima (fmax drops NaN) and the step that reads the rate stops.
template <typename real>
__device__ __forceinline__ real cell_rate_t(const DiskArgs<real>& d, real cs, real idx, real ir, real u, real vp,
                                            real du, real dvp) {
    real sx = fabs(u) + cs, sy = fabs(vp) + cs;
    if (d.has_dust) { sx = fmax(fabs(du), sx); sy = fmax(fabs(dvp), sy); }
    const real r = fmax(sx * idx, sy * d.idy * ir);
    const real all = fabs(u) + fabs(vp) + (d.has_dust ? fabs(du) + fabs(dvp) : real(0));
    return all < real(INFINITY) ? r : real(INFINITY);
}
template <typename real>
__device__ __forceinline__ real cell_rate(const DiskArgs<real>& d, int ip, real u, real vp, real du, real dvp) {
    return cell_rate_t(d, TAB(T_C0, ip), TAB(T_IDX, ip), TAB(T_IR, ip), u, vp, du, dvp);
}

// Pointwise source terms.  mode: 0 first half kick (planet at t, in place),
// 1 predictor (planet at t+dt, gas written to gasT), 2 corrector (planet at
// t+dt, in place) optionally followed by wave killing and the CFL rate,
// 3 wave killing alone.
template <typename real, int TH>
__global__ __launch_bounds__(256) void k_sources(LaunchSet<real> ls, int par, int mode, int do_kill, int do_rate) {
    const int2 w = ls.wl_pt[blockIdx.x];
    if (w.y < 0) return;   // padding entry of the work list (whole block)
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int ntx = pt_ntx(nx);
    const int ty = w.y / ntx;
    const int i = (w.y - ty * ntx) * PT_X + (int)(threadIdx.x % PT_X);
    PG_GLOBAL real* __restrict__ gas = gmem(d.gas[gsel(par)]);
    PG_GLOBAL real* __restrict__ gasT = gmem(d.gas[gsel(par) ^ 1]);
    PG_GLOBAL real* __restrict__ dust = gmem(d.dust[dsel(par)]);
    real rate = real(0);
#pragma unroll
    for (int k = 0; k < TH / PT_TY; ++k) {
        const int j = ty * TH + (int)(threadIdx.x / PT_X) + PT_TY * k;
        if (i >= nx || j >= ny) break;
        const int ip = i + NPAD;
        const size_t o = (size_t)j * nx + i;
        const real h = (real)(0.5 * d.ctl->dt);
        const real rd = gas[o], cu = gas[N + o], cvp = gas[2 * N + o];
        const real rho = TAB(T_G_R, ip) + rd;
        real Vr = real(0), Vp = real(0), Qu = real(0), Qvp = real(0);
        if (d.visc) { Vr = gmem(d.vis)[o]; Vp = gmem(d.vis)[N + o]; }
        if (d.has_dust) { Qu = dust[N + o]; Qvp = dust[2 * N + o]; }
        real nu_ = cu, nvp = cvp, ndu = Qu, ndvp = Qvp;
        if (mode != 3) {
            if (d.has_dust && mode != 1)
                cell_sources<real, true, true>(d, mode, src_tab(d, ip), j, 0.5 * d.ctl->dt, rho, cu, cvp, Vr, Vp, Qu, Qvp, nu_, nvp, ndu, ndvp);
            else
                cell_sources<real, true, false>(d, mode, src_tab(d, ip), j, 0.5 * d.ctl->dt, rho, cu, cvp, Vr, Vp, Qu, Qvp, nu_, nvp, ndu, ndvp);
        }
        real nr_ = rd;   // stored form: the deviation from rho_b
        if (mode == 1) {
            gasT[o] = nr_; gasT[N + o] = nu_; gasT[2 * N + o] = nvp;
        } else {
            if (do_kill) {   // killwave on the corrected state (solve.cu:287-289), toward rho_b
                const real kk = TAB(T_KILL, ip) * (real)d.ctl->dt;
                nr_ += kk * (real(0) - nr_);
                nu_ += kk * (TAB(T_G_U, ip) - nu_);
                nvp += kk * (real(0) - nvp);
            }
            if (do_kill) gas[o] = nr_;   // rho changes only by wave killing
            gas[N + o] = nu_; gas[2 * N + o] = nvp;
            if (d.has_dust) { dust[N + o] = ndu; dust[2 * N + o] = ndvp; }
            if (do_rate) rate = fmax(rate, cell_rate(d, ip, nu_, nvp, ndu, ndvp));
        }
    }
    if (do_rate) block_max_store(d.rate_part + w.y, rate);
}

// Stencil tile of the fused kernels: the PT_X x TH cells of the tile and a
// one-cell halo, (PT_X+2) x (TH+2) LDS slots; slot (lx, ly) is radius
// i0-1+lx, row j0-1+ly (periodic).  Radial ghosts hold the analytic state.
constexpr int ST_X = PT_X + 2;

// The 9-point viscosity of the tile's own cells from a staged state.
template <typename real, int TH>
__device__ __forceinline__ void vis_from_tile(const DiskArgs<real>& d, const real* r_, const real* u_,
                                              const real* v_, int i0, int j0, PG_GLOBAL real* out) {
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int lx = (int)(threadIdx.x % PT_X) + 1;
    const int i = i0 + lx - 1;
    if (i >= nx) return;
    const int ip = i + NPAD;
    real nu[3];
    for (int n = 0; n < 3; ++n) nu[n] = TAB(T_NU, ip - 1 + n);
    const VisTab<real> vt = vis_tab(d, ip);   // the thread's radius: one load per table
#pragma unroll
    for (int k = 0; k < TH / PT_TY; ++k) {
        const int ly = (int)(threadIdx.x / PT_X) + PT_TY * k + 1;
        const int j = j0 + ly - 1;
        if (j >= ny) break;
        const int rows[3] = {ly, ly - 1, ly + 1};
        real nr[9], u[9], v[9];
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int n = 3 * b + a, e = rows[b] * ST_X + lx - 1 + a;
                nr[n] = r_[e] * nu[a];
                u[n] = u_[e];
                v[n] = v_[e];
            }
        real fx, fy;
        vis_cell(d, vt, nr, u, v, fx, fy);
        const size_t o = (size_t)j * nx + i;
        out[o] = fx;
        out[N + o] = fy;
    }
}

// Predictor + viscosity of the predicted state (apply_source_terms with
// mdt != 0 into T, then viscosity_tensor_evaluation2; solve.cu:282-285):
// the predicted state of the tile and its halo is formed in LDS from the
// corrected-sweep state S and the viscous forces of the step start
// (d.vis), and only its viscous forces (into d.vis2) leave the kernel.
template <typename real, int TH>
__global__ __launch_bounds__(256) void k_predvis(LaunchSet<real> ls, int par) {
    __shared__ real Tr[ST_X * (TH + 2)], Tu[ST_X * (TH + 2)], Tv[ST_X * (TH + 2)];
    const int2 w = ls.wl_pt[blockIdx.x];
    if (w.y < 0) return;
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int ntx = pt_ntx(nx);
    const int ty = w.y / ntx;
    constexpr int th = TH;
    const int i0 = (w.y - ty * ntx) * PT_X, j0 = ty * th;
    const PG_GLOBAL real* __restrict__ gas = gmem(d.gas[gsel(par)]);
    const PG_GLOBAL real* __restrict__ vis = gmem(d.vis);
    // Slot (lx, ly): radius i = i0 - 1 + lx, row j0 - 1 + ly.  The radius'
    // table values are loaded once per column (a slot past the radial ghost
    // feeds no output and holds the analytic state of radius nx).
    auto slot = [&](int lx) {
        const int i = i0 - 1 + lx;
        return make_int2(i, min(i, nx) + NPAD);
    };
    auto pred = [&](int lx, int ly, int i, const SrcTab<real>& tb, real rb, real gu, real gv) {
        const int e = ly * ST_X + lx, j = wrap(j0 - 1 + ly, ny);
        real r, u, v;
        if (i < 0 || i >= nx) {
            r = rb; u = gu; v = gv;
        } else {
            const size_t o = (size_t)j * nx + i;
            const real rho = rb + gas[o], cu = gas[N + o], cvp = gas[2 * N + o];
            real nu_, nvp, unused0, unused1;
            cell_sources<real, true, false>(d, 1, tb, j, 0.5 * d.ctl->dt, rho, cu, cvp, vis[o], vis[N + o], real(0),
                                            real(0), nu_, nvp, unused0, unused1);
            r = rho; u = nu_; v = nvp;
        }
        Tr[e] = r; Tu[e] = u; Tv[e] = v;
    };
    const int tl = threadIdx.x;
    if constexpr (sizeof(real) == 8) {
        {   // thread tl: column tl % PT_X + 1, rows tl / PT_X + PT_TY k of the tile and its halo rows
            const int lx = tl % PT_X + 1;
            const int2 c = slot(lx);
            const SrcTab<real> tb = src_tab(d, c.y);
            const real rb = TAB(T_G_R, c.y), gu = TAB(T_G_U, c.y), gv = TAB(T_G_V, c.y);
            for (int ly = tl / PT_X; ly < th + 2; ly += PT_TY) pred(lx, ly, c.x, tb, rb, gu, gv);
        }
        {   // the two halo columns: one slot for each of the threads 2 PT_X .. 2 PT_X + 2 (TH + 2) - 1
            static_assert(2 * PT_X + 2 * (TH + 2) <= PT_X * PT_TY, "the halo columns need a thread per slot");
            const int hs = tl - 2 * PT_X;
            if (hs >= 0 && hs < 2 * (th + 2)) {
                const int side = hs >= th + 2, ly = hs - side * (th + 2), lx = side ? PT_X + 1 : 0;
                const int2 c = slot(lx);
                pred(lx, ly, c.x, src_tab(d, c.y), TAB(T_G_R, c.y), TAB(T_G_U, c.y), TAB(T_G_V, c.y));
            }
        }
    } else {
        // FP32: the slots in row order, a radius' tables read per slot (the
        // column order above measured 27.0 us against 24.9 on the 1024x3072
        // disk in FP32, and 57.6 against 62.9 in FP64; the ghost state is
        // read for the ghost slots only: 30.2 us when every slot read it)
        for (int e = tl; e < ST_X * (th + 2); e += PT_X * PT_TY) {
            const int lx = e % ST_X, ly = e / ST_X;
            const int i = i0 - 1 + lx, j = wrap(j0 - 1 + ly, ny);
            real r, u, v;
            if (i < 0 || i >= nx) {
                const int ip = min(i, nx) + NPAD;   // slots past the ghost feed no output
                r = TAB(T_G_R, ip); u = TAB(T_G_U, ip); v = TAB(T_G_V, ip);
            } else {
                const size_t o = (size_t)j * nx + i;
                const real rho = TAB(T_G_R, i + NPAD) + gas[o], cu = gas[N + o], cvp = gas[2 * N + o];
                real nu_, nvp, unused0, unused1;
                cell_sources<real, true, false>(d, 1, src_tab(d, i + NPAD), j, 0.5 * d.ctl->dt, rho, cu, cvp, vis[o],
                                                vis[N + o], real(0), real(0), nu_, nvp, unused0, unused1);
                r = rho; u = nu_; v = nvp;
            }
            Tr[e] = r; Tu[e] = u; Tv[e] = v;
        }
    }
    __syncthreads();
    if (d.out_pred) {   // the predicted state of the tile's own cells (its density is the input's)
        PG_GLOBAL real* tp = gmem(d.out_pred);
        const int lx = (int)(threadIdx.x % PT_X) + 1, i = i0 + lx - 1;
        for (int ly = (int)(threadIdx.x / PT_X) + 1; ly <= th; ly += PT_TY) {
            const int j = j0 + ly - 1;
            if (i >= nx || j >= ny) break;
            const size_t o = (size_t)j * nx + i;
            const int e = ly * ST_X + lx;
            tp[o] = gas[o]; tp[N + o] = Tu[e]; tp[2 * N + o] = Tv[e];
        }
    }
    vis_from_tile<real, TH>(d, Tr, Tu, Tv, i0, j0, gmem(d.vis2));
}

// Corrector (apply_source_terms_inplace with the viscous forces of the
// predicted state, d.vis2; solve.cu:286-287), wave killing
// (solve.cu:288-289) and the CFL rate of the result (timestep.cu:18-97),
// then the viscous forces of the corrected state for the next step's first
// half kick (into d.vis).  The corrected state of the halo is recomputed
// from its own inputs, so the tile needs no neighbour's result; the gas
// result goes to the other buffer of the pair, because neighbouring tiles
// read this tile's cells as their halo input.  (The dust is corrected in
// place: its halo results are not used.)
template <typename real, int TH>
__global__ __launch_bounds__(256) void k_corrvis(LaunchSet<real> ls, int par, int do_kill) {
    __shared__ real Sr[ST_X * (TH + 2)], Su[ST_X * (TH + 2)], Sv[ST_X * (TH + 2)];
    const int2 w = ls.wl_pt[blockIdx.x];
    if (w.y < 0) return;
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int ntx = pt_ntx(nx);
    const int ty = w.y / ntx;
    constexpr int th = TH;
    const int i0 = (w.y - ty * ntx) * PT_X, j0 = ty * th;
    const PG_GLOBAL real* __restrict__ gas = gmem(d.gas[gsel(par)]);
    PG_GLOBAL real* __restrict__ gasN = gmem(d.gas[gsel(par) ^ 1]);
    PG_GLOBAL real* __restrict__ dust = gmem(d.dust[dsel(par)]);
    const PG_GLOBAL real* __restrict__ vis2 = gmem(d.vis2);
    const real h = (real)(0.5 * d.ctl->dt);
    real rate = real(0);
    // Per-radius table values of a slot column (as in k_predvis).
    struct ColTab {
        SrcTab<real> src;
        real rb, gu, gv, kill, c0, idx;
    };
    auto col_tab = [&](int ipc) {
        return ColTab{src_tab(d, ipc), TAB(T_G_R, ipc), TAB(T_G_U, ipc), TAB(T_G_V, ipc),
                      TAB(T_KILL, ipc), TAB(T_C0, ipc), TAB(T_IDX, ipc)};
    };
    auto corr = [&](int lx, int ly, const ColTab& c) {
        const int e = ly * ST_X + lx;
        const int i = i0 - 1 + lx, jr = j0 - 1 + ly, j = wrap(jr, ny);
        real r, u, v;
        if (i < 0 || i >= nx) {
            r = c.rb; u = c.gu; v = c.gv;
        } else {
            const size_t o = (size_t)j * nx + i;
            const real rb = c.rb;
            const real rd = gas[o], cu = gas[N + o], cvp = gas[2 * N + o];
            const real rho = rb + rd;
            real Qu = real(0), Qvp = real(0);
            if (d.has_dust) { Qu = dust[N + o]; Qvp = dust[2 * N + o]; }
            real nu_ = cu, nvp = cvp, ndu = Qu, ndvp = Qvp;
            if (d.has_dust)
                cell_sources<real, true, true>(d, 2, c.src, j, 0.5 * d.ctl->dt, rho, cu, cvp, vis2[o], vis2[N + o], Qu,
                                               Qvp, nu_, nvp, ndu, ndvp);
            else
                cell_sources<real, true, false>(d, 2, c.src, j, 0.5 * d.ctl->dt, rho, cu, cvp, vis2[o], vis2[N + o], Qu,
                                                Qvp, nu_, nvp, ndu, ndvp);
            real nrd = rd;   // the deviation from rho_b, which wave killing relaxes to 0
            if (do_kill) {
                const real kk = c.kill * (real)d.ctl->dt;
                nrd += kk * (real(0) - nrd);
                nu_ += kk * (c.gu - nu_);
                nvp += kk * (real(0) - nvp);
            }
            r = rb + nrd; u = nu_; v = nvp;
            // the tile's own cells are written back, with their CFL rate
            const bool own = lx >= 1 && lx <= PT_X && ly >= 1 && ly <= th && jr < ny;
            if (own) {
                gasN[o] = nrd;
                gasN[N + o] = nu_;
                gasN[2 * N + o] = nvp;
                if (d.has_dust) { dust[N + o] = ndu; dust[2 * N + o] = ndvp; }
                rate = fmax(rate, cell_rate_t(d, c.c0, c.idx, c.src.iR, nu_, nvp, ndu, ndvp));
            }
        }
        Sr[e] = r; Su[e] = u; Sv[e] = v;
    };
    const int tl = threadIdx.x;
    {   // thread tl: column tl % PT_X + 1, rows tl / PT_X + PT_TY k of the tile and its halo rows
        const int lx = tl % PT_X + 1;
        const ColTab c = col_tab(min(i0 - 1 + lx, nx) + NPAD);
        for (int ly = tl / PT_X; ly < th + 2; ly += PT_TY) corr(lx, ly, c);
    }
    {   // the two halo columns: one slot for each of the threads 2 PT_X .. 2 PT_X + 2 (TH + 2) - 1
        static_assert(2 * PT_X + 2 * (TH + 2) <= PT_X * PT_TY, "the halo columns need a thread per slot");
        const int hs = tl - 2 * PT_X;
        if (hs >= 0 && hs < 2 * (th + 2)) {
            const int side = hs >= th + 2, ly = hs - side * (th + 2), lx = side ? PT_X + 1 : 0;
            corr(lx, ly, col_tab(min(i0 - 1 + lx, nx) + NPAD));
        }
    }
    block_max_store(d.rate_part + w.y, rate);   // includes the barrier the tile needs
    __syncthreads();
    vis_from_tile<real, TH>(d, Sr, Su, Sv, i0, j0, gmem(d.vis));
}

// CFL rate of the current state (get_dt_lv1, timestep.cu:18-97).
template <typename real, int TH>
__global__ __launch_bounds__(256) void k_cfl(LaunchSet<real> ls, int par) {
    const int2 w = ls.wl_pt[blockIdx.x];
    if (w.y < 0) return;   // padding entry of the work list (whole block)
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int ntx = pt_ntx(nx);
    const int ty = w.y / ntx;
    const int i = (w.y - ty * ntx) * PT_X + (int)(threadIdx.x % PT_X);
    const PG_GLOBAL real* __restrict__ gas = gmem(d.gas[gsel(par)]);
    const PG_GLOBAL real* __restrict__ dust = gmem(d.dust[dsel(par)]);
    real rate = real(0);
#pragma unroll
    for (int k = 0; k < TH / PT_TY; ++k) {
        const int j = ty * TH + (int)(threadIdx.x / PT_X) + PT_TY * k;
        if (i >= nx || j >= ny) break;
        const int ip = i + NPAD;
        const size_t o = (size_t)j * nx + i;
        const real du = d.has_dust ? dust[N + o] : real(0), dvp = d.has_dust ? dust[2 * N + o] : real(0);
        rate = fmax(rate, cell_rate(d, ip, gas[N + o], gas[2 * N + o], du, dvp));
    }
    block_max_store(d.rate_part + w.y, rate);
}

// ---------------------------------------------------------- diagnostics
// Time average of the gas: add_grid (output.cu:250-273) with the weight
// dt/2 / interval that averaging (output.cu:299-314) applies before and
// after each step (main.cu:344-350), accumulated in FP64.  The average holds
// the stored fields rho - rho_b, u and vp; the sum of the weights (for
// rho = rho_b + the deviation, p = cs2 rho and v = vbg + vp) accumulates in
// the control block, added once per disk by its tile 0.
template <typename real, int TH>
__global__ __launch_bounds__(256) void k_average(LaunchSet<real> ls, int par) {
    const int2 w = ls.wl_pt[blockIdx.x];
    if (w.y < 0) return;   // padding entry of the work list (whole block)
    const DiskArgs<real>& d = ls.disks[w.x];
    if (!d.ave) return;
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const double fac = 0.5 * d.ctl->dt / d.ctl->ave_interval;
    if (w.y == 0 && threadIdx.x == 0) d.ctl->ave_w += fac;
    const int ntx = pt_ntx(nx);
    const int ty = w.y / ntx;
    const int i = (w.y - ty * ntx) * PT_X + (int)(threadIdx.x % PT_X);
    if (i >= nx) return;
    const PG_GLOBAL real* __restrict__ st = gmem(d.gas[gsel(par)]);
    double* __restrict__ ave = d.ave;
#pragma unroll
    for (int k = 0; k < TH / PT_TY; ++k) {
        const int j = ty * TH + (int)(threadIdx.x / PT_X) + PT_TY * k;
        if (j >= ny) break;
        const size_t o = (size_t)j * nx + i;
        for (int f = 0; f < 3; ++f) ave[f * N + o] += fac * (double)st[f * N + o];
    }
}

// Block sum of a double into this block's slot of a partial array.
__device__ __forceinline__ void block_sum_store(double* part, double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    __shared__ double wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) wsum[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < (int)((blockDim.x + 63) >> 6); ++k) m += wsum[k];
        part[0] = m;
    }
}

// Torque of the planet on the gas: the sum over interior cells of
// R gy rho dV (get_sum_lv1 + cal_torque, output.cu:336-402), with gy the
// azimuthal planet acceleration of o_star_planet_grav_cyl at dt = 0
// (output/forces.cu:41-58, 106-127; its indirect term included) for the
// planet at the end of the last step, in FP64.  One partial per tile.
template <typename real, int TH>
__global__ __launch_bounds__(256) void k_torque(LaunchSet<real> ls, int par, double* part) {
    const int2 w = ls.wl_pt[blockIdx.x];
    if (w.y < 0) return;   // padding entry of the work list (whole block)
    const DiskArgs<real>& d = ls.disks[w.x];
    const int nx = d.nx, ny = d.ny;
    const int ntx = pt_ntx(nx);
    const int ty = w.y / ntx;
    const int i = (w.y - ty * ntx) * PT_X + (int)(threadIdx.x % PT_X);
    const PG_GLOBAL real* __restrict__ rho = gmem(d.gas[gsel(par)]);
    const PlanetK pk = d.ctl->pk1;
    double acc = 0.0;
    if (i < nx) {
        const double R = d.rot[nx + i];
        const double rb = d.rot[2 * nx + i];   // FP64 background density
        const double dv = (double)TAB(T_XV, i + NPAD) * d.dy_d;
#pragma unroll
        for (int k = 0; k < TH / PT_TY; ++k) {
            const int j = ty * TH + (int)(threadIdx.x / PT_X) + PT_TY * k;
            if (j >= ny) break;
            const double sj = (double)d.phitab[4 * j], cj = (double)d.phitab[4 * j + 1];
            const double sn = sj * pk.cphi - cj * pk.sphi;   // sin(phi - phi_p)
            const double cs = cj * pk.cphi + sj * pk.sphi;   // cos(phi - phi_p)
            const double Rp = sqrt(R * R + pk.x * pk.x - 2.0 * pk.x * R * cs + pk.rs * pk.rs);
            const double kk = 1.0 / (Rp * Rp * Rp);
            const double gy = -pk.m * pk.x * sn * kk + sn * pk.m_over_x2;
            acc += R * gy * ((rb + (double)rho[(size_t)j * nx + i]) * dv);
        }
    }
    block_sum_store(part + w.y, acc);
}

// Max of the partials into the control block (host-driven global_dt).
template <typename real>
__global__ __launch_bounds__(1024) void k_rate_reduce(LaunchSet<real> ls) {
    const DiskArgs<real>& d = ls.disks[blockIdx.x];
    double m = reduce_partials(d.rate_part, d.rate_nblk);
    if (threadIdx.x == 0) d.ctl->rate_bits = (unsigned long long)__double_as_longlong(m);
}

// Step prologue, one block of 1024 threads.  mode 0: t and dt were set by
// the host.  mode 1 (device loop): finish the previous step (t += dt), take
// dt = min(CFL/rate, 1.1 dt_prev) clipped to end_time, record it.  Then the
// planet at t and t+dt and the orbital shift of every radius.
// The device loop stops where the reference loop does (main.cu:323-327):
// once t reaches end_time dt is 0, and every pass of a step of size 0
// leaves the state as it is (no shift, no flux, no kick).  A CFL rate that
// is not finite and positive (a state with NaN or infinite velocities)
// sets ctl->err with the step's index and also makes dt 0, so the loop
// freezes there and the host reports it.
// mode 2 (replay of a recorded step, one disk): t = td[0], dt = td[1].
// The loads that do not depend on the new step size (the control block, the
// planets, each thread's first radius of the shift) are issued before the
// rate reduction, and the two planets (lanes 0 and 1 of wave 0, one
// evaluation for both) need no barrier before the shifts, which do not read
// them.
// The prologue of one disk, by the calling block (any size).
template <typename real>
__device__ __forceinline__ void prologue_disk(const DiskArgs<real>& d, int mode, int do_orbadv, const double* td) {
    double* dts = d.dts;
    StepCtl* ctl = d.ctl;
    __shared__ double s_dt, s_t;
    const int tid = threadIdx.x;
    const bool oa = d.orbadv && do_orbadv;
    double rt0 = 0.0, ry0 = 0.0;
    if (oa && tid < d.nx) {
        rt0 = d.rot[tid];
        ry0 = d.rot[d.nx + tid];
    }
    double t0 = 0.0, dt0 = 0.0, tend = 0.0, cfl = 0.0;
    int n0 = 0, err0 = 0;
    Body b{};
    if (tid == 0) {
        t0 = ctl->t;
        dt0 = ctl->dt;
        n0 = ctl->nstep;
        tend = ctl->end_time;
        cfl = d.CFL;
        err0 = ctl->err;
    }
    if (tid < 2) b = tid == 0 ? ctl->planet0 : ctl->planet1;
    double rate = mode == 1 ? reduce_partials(d.rate_part, d.rate_nblk) : 0.0;
    if (tid == 0) {
        double t = t0, dt = dt0;
        if (mode == 2) {
            t = td[0];
            dt = td[1];
            ctl->t = t;
            ctl->dt = dt;
        }
        if (mode == 1) {
            if (n0 > 0) t += dt;
            ctl->rate_bits = (unsigned long long)__double_as_longlong(rate);
            double dtn = cfl / rate;
            double cap = dt * 1.1;
            dtn = dtn < cap ? dtn : cap;
            if (t + dtn > tend) dtn = tend > t ? tend - t : 0.0;
            if (err0 || !(rate > 0.0 && rate < INFINITY)) {
                if (!err0) {
                    ctl->err = 1;
                    ctl->err_step = n0;
                }
                dtn = 0.0;
            }
            dt = dtn;
            ctl->t = t;
            ctl->dt = dt;
            ctl->dt_prev = dt;
            if (dts && n0 < d.dts_cap) dts[n0] = dt;
            ctl->nstep = n0 + 1;
        }
        s_t = t;
        s_dt = dt;
    }
    __syncthreads();
    // the planet at t and at t + dt, on two lanes
    if (tid < 2) {
        orbit_solution(b, tid == 0 ? s_t : s_t + s_dt, d.ramp_time, d.planet_mass, (double)d.frame_omega);
        const PlanetK v = planet_kernel_view(b);
        if (tid == 0) { ctl->planet0 = b; ctl->pk0 = v; }
        else { ctl->planet1 = b; ctl->pk1 = v; }
    }
    if (oa) {
        const double dt = s_dt;
        for (int i = tid; i < d.nx; i += blockDim.x) {
            // set_orbital_advection (orbit.cu:74-99), FP64
            double dy = d.dy_d * (i == tid ? ry0 : d.rot[d.nx + i]);
            double rt = i == tid ? rt0 : d.rot[i];
            int shift = 0;
            double res = 0.0;   // a step of size 0 moves nothing
            if (dt > 0.0) {
                shift = (int)floor(rt * dt / dy + 0.5);
                res = rt - shift * dy / dt;
            }
            d.inc[i] = shift;
            d.res[i] = (real)res;
        }
    }
}

template <typename real>
__global__ __launch_bounds__(1024) void k_prologue(LaunchSet<real> ls, int mode, int do_orbadv, const double* td) {
    prologue_disk<real>(ls.disks[blockIdx.x], mode, do_orbadv, td);
}

// Replays of n recorded steps at once: block b sets the control block and
// orbital shift of descriptor b * stride from t = td[2b], dt = td[2b + 1]
// (the replays depend on the recorded times only, not on any state).
template <typename real>
__global__ __launch_bounds__(1024) void k_prologue_replay(LaunchSet<real> ls, int stride, const double* td) {
    prologue_disk<real>(ls.disks[(size_t)blockIdx.x * stride], 2, 1, td + 2 * (size_t)blockIdx.x);
}

// ------------------------------------------------------------- launchers
template <typename Kern>
static hipError_t set_lds(Kern k, size_t lds) {
    if (lds <= 65536) return hipSuccess;
    return hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

// which: 0 the gas, 1 the dust, 2 both in one launch
template <typename real, int TX>
static hipError_t launch_sweep_x(const LaunchSet<real>& ls, int par, int which, int kick, hipStream_t s) {
    const size_t S = XBlock<TX>::S;
    if (which == 1) {
        size_t lds = XLDS_DUST * S * sizeof(real);
        hipError_t e = set_lds(k_sweep_x_dust<real, TX>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_sweep_x_dust<real, TX>), dim3(ls.n_x), dim3(XBlock<TX>::BT), lds, s, ls, par, kick);
    } else if (which == 0) {
        size_t lds = XLDS_GAS * S * sizeof(real);
        hipError_t e = set_lds(k_sweep_x_gas<real, TX>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_sweep_x_gas<real, TX>), dim3(ls.n_x), dim3(XBlock<TX>::BT), lds, s, ls, par, kick);
    } else {
        size_t lds = XLDS_GAS * S * sizeof(real);
        hipError_t e = set_lds(k_sweep_x_both<real, TX>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_sweep_x_both<real, TX>), dim3(2 * ls.n_x), dim3(XBlock<TX>::BT), lds, s, ls, par, kick);
    }
    return hipGetLastError();
}

template <typename real>
static hipError_t sweep_x_any(const LaunchSet<real>& ls, int par, int which, int kick, hipStream_t s) {
    switch (ls.tile_x) {
        case 122: return launch_sweep_x<real, 122>(ls, par, which, kick, s);
        case 250: return launch_sweep_x<real, 250>(ls, par, which, kick, s);
        case 506: return launch_sweep_x<real, 506>(ls, par, which, kick, s);
        case 128: return launch_sweep_x<real, 128>(ls, par, which, kick, s);
        case 384: return launch_sweep_x<real, 384>(ls, par, which, kick, s);
        case 512: return launch_sweep_x<real, 512>(ls, par, which, kick, s);
        case 1024: return launch_sweep_x<real, 1024>(ls, par, which, kick, s);
        default: return launch_sweep_x<real, 256>(ls, par, which, kick, s);
    }
}

template <typename real, int B>
static hipError_t launch_ring(const LaunchSet<real>& ls, int par, int flags, int which, hipStream_t s) {
    if (which == 1)
        hipLaunchKernelGGL((k_phi<real, true, B>), dim3(ls.n_ring), dim3(B), 0, s, ls, par, flags);
    else if (which == 0)
        hipLaunchKernelGGL((k_phi<real, false, B>), dim3(ls.n_ring), dim3(B), 0, s, ls, par, flags);
    else
        hipLaunchKernelGGL((k_phi_both<real, B>), dim3(2 * ls.n_ring), dim3(B), 0, s, ls, par, flags);
    return hipGetLastError();
}

template <typename real>
static hipError_t ring_any(const LaunchSet<real>& ls, int par, int flags, int which, hipStream_t s) {
    switch (ls.ring_b) {
        case 256: return launch_ring<real, 256>(ls, par, flags, which, s);
        case 1024: return launch_ring<real, 1024>(ls, par, flags, which, s);
        default: return launch_ring<real, 512>(ls, par, flags, which, s);
    }
}

// Launch a pointwise/stencil kernel with the tile height of the launch set
// as a compile-time constant (4, 8 or 16 rows).
#define PT_LAUNCH(K, ls, s, ...)                                                                       \
    do {                                                                                               \
        if ((ls).pt_rows == 4)                                                                         \
            hipLaunchKernelGGL((K<real, 4>), dim3((ls).n_pt), dim3(256), 0, s, ls, __VA_ARGS__);       \
        else if ((ls).pt_rows == 8)                                                                    \
            hipLaunchKernelGGL((K<real, 8>), dim3((ls).n_pt), dim3(256), 0, s, ls, __VA_ARGS__);       \
        else                                                                                           \
            hipLaunchKernelGGL((K<real, 16>), dim3((ls).n_pt), dim3(256), 0, s, ls, __VA_ARGS__);      \
    } while (0)

template <typename real>
struct Launch {
    static int phi_tile_radii(int B) {
        return B == 256 ? PhiTile<real, 256>::TR : B == 1024 ? PhiTile<real, 1024>::TR : PhiTile<real, 512>::TR;
    }
    static int phi_tile_rows(int B) {
        return B == 256 ? PhiTile<real, 256>::TY : B == 1024 ? PhiTile<real, 1024>::TY : PhiTile<real, 512>::TY;
    }
    static hipError_t sweep_x_gas(const LaunchSet<real>& ls, int par, int kick, hipStream_t s) { return sweep_x_any(ls, par, 0, kick, s); }
    static hipError_t sweep_x_dust(const LaunchSet<real>& ls, int par, int kick, hipStream_t s) { return sweep_x_any(ls, par, 1, kick, s); }
    static hipError_t sweep_x_both(const LaunchSet<real>& ls, int par, int kick, hipStream_t s) { return sweep_x_any(ls, par, 2, kick, s); }
    static hipError_t ring_gas(const LaunchSet<real>& ls, int par, int flags, hipStream_t s) { return ring_any(ls, par, flags, 0, s); }
    static hipError_t ring_dust(const LaunchSet<real>& ls, int par, int flags, hipStream_t s) { return ring_any(ls, par, flags, 1, s); }
    static hipError_t ring_both(const LaunchSet<real>& ls, int par, int flags, hipStream_t s) { return ring_any(ls, par, flags, 2, s); }
    static hipError_t viscosity(const LaunchSet<real>& ls, int par, int which, hipStream_t s) {
        PT_LAUNCH(k_viscosity, ls, s, par, which);
        return hipGetLastError();
    }
    static hipError_t sources(const LaunchSet<real>& ls, int par, int mode, int kill, int rate, hipStream_t s) {
        PT_LAUNCH(k_sources, ls, s, par, mode, kill, rate);
        return hipGetLastError();
    }
    static hipError_t predvis(const LaunchSet<real>& ls, int par, hipStream_t s) {
        PT_LAUNCH(k_predvis, ls, s, par);
        return hipGetLastError();
    }
    static hipError_t corrvis(const LaunchSet<real>& ls, int par, int kill, hipStream_t s) {
        PT_LAUNCH(k_corrvis, ls, s, par, kill);
        return hipGetLastError();
    }
    static hipError_t average(const LaunchSet<real>& ls, int par, hipStream_t s) {
        PT_LAUNCH(k_average, ls, s, par);
        return hipGetLastError();
    }
    static hipError_t torque(const LaunchSet<real>& ls, int par, double* part, hipStream_t s) {
        PT_LAUNCH(k_torque, ls, s, par, part);
        return hipGetLastError();
    }
    static hipError_t cfl(const LaunchSet<real>& ls, int par, hipStream_t s) {
        PT_LAUNCH(k_cfl, ls, s, par);
        return hipGetLastError();
    }
    static hipError_t rate_reduce(const LaunchSet<real>& ls, hipStream_t s) {
        hipLaunchKernelGGL(k_rate_reduce<real>, dim3(ls.ndisk), dim3(1024), 0, s, ls);
        return hipGetLastError();
    }
    static hipError_t prologue(const LaunchSet<real>& ls, int mode, int do_orbadv, hipStream_t s) {
        hipLaunchKernelGGL(k_prologue<real>, dim3(ls.ndisk), dim3(1024), 0, s, ls, mode, do_orbadv, (const double*)nullptr);
        return hipGetLastError();
    }
    static hipError_t prologue_replay(const LaunchSet<real>& ls, const double* td, hipStream_t s) {
        hipLaunchKernelGGL(k_prologue<real>, dim3(1), dim3(1024), 0, s, ls, 2, 1, td);
        return hipGetLastError();
    }
    static hipError_t prologue_replay_many(const LaunchSet<real>& ls, int stride, int n, const double* td,
                                           hipStream_t s) {
        hipLaunchKernelGGL(k_prologue_replay<real>, dim3(n), dim3(1024), 0, s, ls, stride, td);
        return hipGetLastError();
    }
};

template struct Launch<float>;
template struct Launch<double>;

}  // namespace pg
