// pg_adjoint.hip -- reverse-mode (adjoint) kernels of the Godunov step of a
// 2D locally isothermal gas disk (no dust, no wave killing).
//
// The step is split as the forward step runs it (solve.cu:212-292):
//   V0 = viscosity(S0); S1 = kick(S0, V0); S2 = x-sweep(S1);
//   SA = azimuthal sweep(S2); S3 = shift + residual advection(SA);
//   T = predictor(S3, V0); V1 = viscosity(T); S4 = corrector(S3, V1)
// and each kernel here is the transpose of one of these maps at the forward
// point, with the step size, the orbital shift and every branch of the
// forward arithmetic (limiters, floors, the HLLC upwind choice) frozen.
// Each kernel recomputes the forward values of its tile from the stored
// input of its pass, then runs the phases backwards.  Adjoints of stencil
// inputs are gathered (each thread sums what its cell receives), so results
// do not depend on thread order; the two parameter adjoints (planet mass,
// viscosity alpha) are summed per workgroup into slots that the host adds up.
#include <hip/hip_runtime.h>

#include <cmath>

#include "pg_adjoint_fn.h"
#include "pg_device.h"
#include "pg_internal.h"

namespace pg {

// Sum a per-thread value over the block and add it to this block's slot of
// a parameter-adjoint array (slot = linear block index; launches on one
// stream run one after the other, so each slot has one writer at a time and
// no atomics are needed).  The host sums the slots at the end.
__device__ __forceinline__ void block_sum_add(double* slots, double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) wsum[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < (int)((blockDim.x + 63) >> 6); ++k) s += wsum[k];
        if (s != 0.0) slots[blockIdx.x + (size_t)gridDim.x * blockIdx.y] += s;
    }
}

template <typename real>
__device__ __forceinline__ void put(real* li, size_t o, real v, int acc) {
    li[o] = acc ? li[o] + v : v;
}

// ------------------------------------------------------- x sweep adjoint
// Tile of TX output cells of one row; forward values on cells x0-6 ..
// x0+TX+5 (LDS slot s <-> cell x0-6+s).  Cells beyond the radial ghost
// zone take the outermost ghost's values: they feed nothing the outputs
// depend on, and finite values keep their zero adjoints zero.
// Block of the x-sweep adjoint: one thread per LDS slot of a TX-cell tile
// (TX + 12 slots, rounded up to whole waves).
template <int TX> struct AdjXBlock { static constexpr int B = (TX + 4 * NPAD + 63) / 64 * 64; };

template <typename real, int TX>
__global__ __launch_bounds__(AdjXBlock<TX>::B) void k_adj_x(AdjArgs<real> A) {
    const DiskArgs<real>& d = *A.d;
    constexpr int S = TX + 4 * NPAD;
    constexpr int BT = AdjXBlock<TX>::B;    // one thread per slot: every phase is one trip
    constexpr int NT = (S + BT - 1) / BT;   // slots per thread (1)
    __shared__ real r_[S], p_[S], u_[S], vp_[S], L_[S], fl_[S];
    // W: parabolas (8), fluxes (4) and update adjoints (4) of the forward and
    // update phases, then the difference adjoints (16) and flattening
    // adjoint (1) of the parabola reversal
    __shared__ real W[17][S];
    // per-interface adjoints sent to the left cell (those sent to the right
    // cell stay in the registers of the thread that owns it); CL[0..4] then
    // hold the flattening adjoints of the pressures
    __shared__ real CL[TI_N][S];
    real (*sl_)[S] = W;
    real (*sr_)[S] = W + 4;
    real *Fr_ = W[8], *Fu_ = W[9], *FL_ = W[10], *pr_ = W[11];
    real *cA_ = W[12], *cB_ = W[13], *cC_ = W[14], *cP_ = W[15];
    const int nx = d.nx;
    const size_t N = (size_t)nx * d.ny;
    const int t = threadIdx.x;
    const int j = blockIdx.y;
    const int x0 = blockIdx.x * TX;
    const int nt = min(TX, nx - x0);
    const int ns = nt + 4 * NPAD;
    const real dt = (real)d.ctl->dt;
    const size_t row = (size_t)j * nx;
    const int ipmax = nx + 2 * NPAD - 1;
    auto tip = [&](int s) { return min(max(x0 - 2 * NPAD + s + NPAD, 0), ipmax); };   // clamped table index
    auto inside = [&](int s) { const int g = x0 - 2 * NPAD + s; return g >= 0 && g < nx; };

    // ---- forward recompute (k_sweep_x_gas) ----
    for (int s = t; s < ns; s += BT) {
        const int g = x0 - 2 * NPAD + s;
        const int ip = tip(s);
        real r, u, vp;
        if (g < 0 || g >= nx) {
            r = TAB(T_G_R, ip); u = TAB(T_G_U, ip); vp = TAB(T_G_V, ip);
        } else {
            r = TAB(T_G_R, ip) + A.s[row + g]; u = A.s[N + row + g]; vp = A.s[2 * N + row + g];
        }
        r_[s] = r; p_[s] = TAB(T_CS2, ip) * r; u_[s] = u; vp_[s] = vp;
        L_[s] = TAB(T_XC, ip) * (TAB(T_VBG, ip) + vp);
    }
    __syncthreads();
    for (int s = t + 2; s < ns - 2; s += BT)
        fl_[s] = flatten5(p_[s - 2], p_[s - 1], p_[s], p_[s + 1], p_[s + 2], u_[s - 1], u_[s], u_[s + 1], real(1));
    __syncthreads();
    const real* f4[4] = {r_, p_, u_, L_};
    for (int s = t + 2; s < ns - 2; s += BT) {
        const int ip = tip(s);
        const real w0 = TAB(T_WL0, ip), w1 = TAB(T_WL1, ip), w2 = TAB(T_WL2, ip);
        const real v0 = TAB(T_WR0, ip), v1 = TAB(T_WR1, ip), v2 = TAB(T_WR2, ip);
        for (int f = 0; f < 4; ++f) {
            const real* a = f4[f];
            real l, r;
            slopes(a[s - 2], a[s - 1], a[s], a[s + 1], a[s + 2], w0, w1, w2, v0, v1, v2, fl_[s], l, r);
            sl_[f][s] = l; sr_[f][s] = r;
        }
    }
    __syncthreads();
    auto state_in = [&](int s, real* in) {
        in[TI_R] = r_[s]; in[TI_P] = p_[s]; in[TI_U] = u_[s];
        in[TI_RSL] = sl_[0][s]; in[TI_RSR] = sr_[0][s];
        in[TI_PSL] = sl_[1][s]; in[TI_PSR] = sr_[1][s];
        in[TI_USL] = sl_[2][s]; in[TI_USR] = sr_[2][s];
        in[TI_VSL] = sl_[3][s]; in[TI_V] = L_[s]; in[TI_VSR] = sr_[3][s];
    };
    // forward fluxes of interfaces [3, ns-3] (interface k: between slots k-1 and k)
    for (int k = t + NPAD; k < ns - 2; k += BT) {
        const int ipl = tip(k - 1), ipr = tip(k);
        real inL[TI_N], inR[TI_N];
        state_in(k - 1, inL);
        state_in(k, inR);
        real rl, pl, ul, vl, cl, rr, pr, ur, vr, cr;
        trace_left(TAB(T_C0, ipl), dt, TAB(T_IDX, ipl), inL[0], inL[1], inL[2], inL[3], inL[4], inL[5], inL[6],
                   inL[7], inL[8], inL[9], inL[10], inL[11], rl, pl, ul, vl, cl);
        trace_right(TAB(T_C0, ipr), dt, TAB(T_IDX, ipr), inR[0], inR[1], inR[2], inR[3], inR[4], inR[5], inR[6],
                    inR[7], inR[8], inR[9], inR[10], inR[11], rr, pr, ur, vr, cr);
        real Fr, Fu, FL, pres;
        hllc_iso(cl, cr, rl, pl, ul, vl, rr, pr, ur, vr, Fr, Fu, FL, pres);
        const real g = TAB(T_XA, ipr);
        Fr_[k] = Fr * g; Fu_[k] = Fu * g; FL_[k] = FL * g; pr_[k] = pres;
    }
    __syncthreads();
    // ---- reverse of the update (cells [3, ns-3)); the cell adjoints stay
    //      in registers (thread t owns slots t + q BT in every later phase)
    real ar[NT], ap[NT], au[NT], aL[NT], avp[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int s = t + q * BT;
        real lr = 0, lu = 0, lvp = 0, cA = 0, cB = 0, cC = 0, cP = 0;
        if (s >= NPAD && s < ns - NPAD && inside(s)) {
            const int g = x0 - 2 * NPAD + s;
            const int ip = g + NPAD;
            const real lrn = A.lo[row + g], lun = A.lo[N + row + g], lvn = A.lo[2 * N + row + g];
            const real g0 = TAB(T_XA, ip), g1 = TAB(T_XA, ip + 1);
            const real f = dt * TAB(T_IXV, ip);
            const real Dr = (Fr_[s] - Fr_[s + 1]) * f;
            const real Du = (Fu_[s] - Fu_[s + 1] + real(0.5) * (g0 + g1) * (pr_[s] - pr_[s + 1])) * f;
            const real DL = (FL_[s] - FL_[s + 1]) * f;
            const real r = r_[s], u = u_[s], vp = vp_[s];
            const real rn = r + Dr;
            const real rmin = fmax(r * real(1.0e-10), real(1.0e-10));
            if (rn < rmin) {
                lr = r * real(1.0e-10) > real(1.0e-10) ? lrn * real(1.0e-10) : real(0);
                lu = lun;
                lvp = lvn;
            } else {
                const real irn = FM<real>::rcp(rn);
                const real iR = TAB(T_IR, ip), vbg = TAB(T_VBG, ip);
                const real un = (r * u + Du) * irn;
                const real vn = (r * vp + (DL * iR - vbg * Dr)) * irn;
                const real lrt = lrn - (lun * un + lvn * vn) * irn;
                lr = lrt + (lun * u + lvn * vp) * irn;
                lu = lun * r * irn;
                lvp = lvn * r * irn;
                const real lDr = lrt - lvn * vbg * irn, lDu = lun * irn, lDL = lvn * iR * irn;
                cA = lDr * f; cB = lDu * f; cC = lDL * f; cP = lDu * real(0.5) * (g0 + g1) * f;
            }
        }
        ar[q] = lr; ap[q] = real(0); au[q] = lu; aL[q] = real(0); avp[q] = lvp;
        if (s < S) { cA_[s] = cA; cB_[s] = cB; cC_[s] = cC; cP_[s] = cP; }
    }
    __syncthreads();
    // ---- reverse of every interface in one pass: interface k (slots k-1 |
    //      k) is reversed by the thread owning slot k, which keeps what it
    //      sends to that slot and writes what it sends to slot k-1 into CL ----
    real lR[NT][TI_N];
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int k = t + q * BT;
        real lL[TI_N];
        for (int m = 0; m < TI_N; ++m) { lL[m] = real(0); lR[q][m] = real(0); }
        if (k >= S) break;
        const bool iface = k >= NPAD && k < ns - 2;
        const real lFs = iface ? cA_[k] - cA_[k - 1] : real(0), lUs = iface ? cB_[k] - cB_[k - 1] : real(0);
        const real lLs = iface ? cC_[k] - cC_[k - 1] : real(0), lpres = iface ? cP_[k] - cP_[k - 1] : real(0);
        if (lFs != real(0) || lUs != real(0) || lLs != real(0) || lpres != real(0)) {
            const int ipl = tip(k - 1), ipr = tip(k);
            const real g = TAB(T_XA, ipr);
            real inL[TI_N], inR[TI_N];
            state_in(k - 1, inL);
            state_in(k, inR);
            const real cL0 = TAB(T_C0, ipl), cR0 = TAB(T_C0, ipr), idxL = TAB(T_IDX, ipl), idxR = TAB(T_IDX, ipr);
            real S8[8], cl, cr;
            trace_left(cL0, dt, idxL, inL[0], inL[1], inL[2], inL[3], inL[4], inL[5], inL[6], inL[7], inL[8],
                       inL[9], inL[10], inL[11], S8[0], S8[1], S8[2], S8[3], cl);
            trace_right(cR0, dt, idxR, inR[0], inR[1], inR[2], inR[3], inR[4], inR[5], inR[6], inR[7], inR[8],
                        inR[9], inR[10], inR[11], S8[4], S8[5], S8[6], S8[7], cr);
            real lS8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, lcl = 0, lcr = 0;
            hllc_iso_adj(cl, cr, S8, lFs * g, lUs * g, lLs * g, lpres, lS8, lcl, lcr);
            trace_left_adj(cL0, dt, idxL, inL, lS8[0], lS8[1], lS8[2], lS8[3], lcl, lL);
            trace_right_adj(cR0, dt, idxR, inR, lS8[4], lS8[5], lS8[6], lS8[7], lcr, lR[q]);
        }
        for (int m = 0; m < TI_N; ++m) CL[m][k] = lL[m];
    }
    __syncthreads();
    // ---- gather per cell (right cell of interface s, left cell of
    //      interface s+1), then the reverse of its parabolas: adjoints of
    //      the neighbour differences and of flattening (into W) ----
    real (*ld_)[4][S] = reinterpret_cast<real (*)[4][S]>(&W[0][0]);
    real* lfl_ = W[16];
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int s = t + q * BT;
        if (s >= S) break;
        real gsum[TI_N];
        for (int m = 0; m < TI_N; ++m) gsum[m] = lR[q][m];   // zero unless interface s exists
        if (s + 1 >= NPAD && s + 1 <= ns - 3)
            for (int m = 0; m < TI_N; ++m) gsum[m] += CL[m][s + 1];
        ar[q] += gsum[TI_R]; ap[q] += gsum[TI_P]; au[q] += gsum[TI_U]; aL[q] += gsum[TI_V];
        const real asl[4] = {gsum[TI_RSL], gsum[TI_PSL], gsum[TI_USL], gsum[TI_VSL]};
        const real asr[4] = {gsum[TI_RSR], gsum[TI_PSR], gsum[TI_USR], gsum[TI_VSR]};
        real lflat = real(0);
        real ldf[4][4];
        for (int f = 0; f < 4; ++f)
            for (int m = 0; m < 4; ++m) ldf[f][m] = real(0);
        if (s >= 2 && s < ns - 2) {
            const int ip = tip(s);
            const real w0 = TAB(T_WL0, ip), w1 = TAB(T_WL1, ip), w2 = TAB(T_WL2, ip);
            const real v0 = TAB(T_WR0, ip), v1 = TAB(T_WR1, ip), v2 = TAB(T_WR2, ip);
            for (int f = 0; f < 4; ++f) {
                if (asl[f] == real(0) && asr[f] == real(0)) continue;
                const real* a = f4[f];
                slopes_adj(a[s - 2], a[s - 1], a[s], a[s + 1], a[s + 2], w0, w1, w2, v0, v1, v2, fl_[s],
                           asl[f], asr[f], ldf[f], lflat);
            }
        }
        // (W's parabolas and fluxes are dead: the interface pass is done)
        for (int f = 0; f < 4; ++f)
            for (int m = 0; m < 4; ++m) ld_[f][m][s] = ldf[f][m];
        lfl_[s] = lflat;
    }
    __syncthreads();
    // gather the differences into the cells; flattening adjoint per cell
    real (*lfp_)[S] = CL;
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int s = t + q * BT;
        if (s >= S) break;
        real add[4];
        for (int f = 0; f < 4; ++f) {
            real v = real(0);
            if (s + 2 < S) v -= ld_[f][0][s + 2];
            if (s + 1 < S) v += ld_[f][0][s + 1] - ld_[f][1][s + 1];
            v += ld_[f][1][s] - ld_[f][2][s];
            if (s >= 1) v += ld_[f][2][s - 1] - ld_[f][3][s - 1];
            if (s >= 2) v += ld_[f][3][s - 2];
            add[f] = v;
        }
        ar[q] += add[0]; ap[q] += add[1]; au[q] += add[2]; aL[q] += add[3];
        real lp5[5] = {0, 0, 0, 0, 0};
        if (s >= 2 && s < ns - 2 && lfl_[s] != real(0))
            flatten5_adj(p_[s - 2], p_[s - 1], p_[s], p_[s + 1], p_[s + 2], u_[s - 1], u_[s], u_[s + 1], lfl_[s], lp5);
        for (int m = 0; m < 5; ++m) lfp_[m][s] = lp5[m];
    }
    __syncthreads();
    // ---- gather the flattening adjoints, chain p = cs2 rho and L = R (vbg + vp), store ----
    double lalpha = 0.0;
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int s = t + q * BT;
        if (s >= ns) break;
        real lp = ap[q];
        if (s + 2 < S) lp += lfp_[0][s + 2];
        if (s + 1 < S) lp += lfp_[1][s + 1];
        lp += lfp_[2][s];
        if (s >= 1) lp += lfp_[3][s - 1];
        if (s >= 2) lp += lfp_[4][s - 2];
        const int g = x0 - 2 * NPAD + s;
        const int ip = tip(s);
        if (s >= 2 * NPAD && s < 2 * NPAD + nt) {
            const size_t o = row + g;
            put(A.li, o, ar[q] + TAB(T_CS2, ip) * lp, A.accumulate);
            put(A.li, N + o, au[q], A.accumulate);
            put(A.li, 2 * N + o, avp[q] + TAB(T_XC, ip) * aL[q], A.accumulate);
        } else if ((g < 0 && g >= -NPAD) || (g >= nx && g < nx + NPAD)) {
            // radial ghost: u = -1.5 nu(R)/R is proportional to alpha
            lalpha += (double)au[q] * (double)TAB(T_G_U, ip);
        }
    }
    if (A.alpha > 0.0) block_sum_add(A.lparam + A.nslot, lalpha / A.alpha);
}

// ------------------------------------------------- azimuthal sweep adjoint
// Tile of TR radii x TY output rows, no shift (the pass SA = sweep(S2));
// LDS row k <-> row y0-6+k, rows [0, TY+12).
// FP32: 16 radii x 32 rows, 512 threads, two cells per thread (the 12-row
// halo adds 38 % to the rows swept, against 75 % for 16-row tiles); FP64
// 8 x 16 within the LDS of a CU.
template <typename real> struct AdjPhiTile {
    static constexpr int TR = sizeof(real) == 8 ? 8 : 16, TY = sizeof(real) == 8 ? 16 : 32;
};

template <typename real>
__global__ __launch_bounds__(AdjPhiTile<real>::TR * AdjPhiTile<real>::TY) void k_adj_phi(AdjArgs<real> A) {
    constexpr int TR = AdjPhiTile<real>::TR, TY = AdjPhiTile<real>::TY, B = TR * TY;
    constexpr int NK = TY + 4 * NPAD, E = NK * TR;
    constexpr int NT = (E + B - 1) / B;   // cells per thread
    __shared__ real c0[E], c1[E], c2[E], fl[E];
    // W: parabolas (6), fluxes (4) and update adjoints (3) of the forward and
    // update phases, then the difference adjoints (12) and flattening adjoint
    // (1) of the parabola reversal
    __shared__ real W[13][E];
    // per-interface adjoints sent to the left cell (those sent to the right
    // cell stay in the registers of the thread that owns it); CL[0..4] then
    // hold the flattening adjoints of the pressures
    __shared__ real CL[TC_N][E];
    real (*sl)[E] = W;
    real (*sr)[E] = W + 3;
    real *F0 = W[6], *F1 = W[7], *F2 = W[8], *F3 = W[9];
    real *cA = W[10], *cB = W[11], *cC = W[12];
    const DiskArgs<real>& d = *A.d;
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int t = threadIdx.x;
    const int ri = t % TR;
    const int i = blockIdx.x * TR + ri;
    const bool live = i < nx;
    const int ic = live ? i : nx - 1;
    const int ip = ic + NPAD;
    const int y0 = blockIdx.y * TY;
    const real iR = TAB(T_IR, ip);
    const real dt = (real)d.ctl->dt;
    const real dts = dt * iR, idy = d.idy, fup = dt * idy * iR;
    const real cs2 = TAB(T_CS2, ip), cs0 = TAB(T_C0, ip), ic0 = TAB(T_IC0, ip);
    const real wy[6] = {d.wLy[0], d.wLy[1], d.wLy[2], d.wRy[0], d.wRy[1], d.wRy[2]};
    const int jbase = y0 - 2 * NPAD;
    static_assert(B % TR == 0, "tile");
    for (int e = t; e < E; e += B) {
        real r = real(1), vp = real(0), u = real(0);
        if (live) {
            const size_t o = (size_t)wrap(jbase + e / TR, ny) * nx + i;
            r = TAB(T_G_R, ip) + A.s[o]; vp = A.s[2 * N + o]; u = A.s[N + o];
        }
        c0[e] = r; c1[e] = vp; c2[e] = u;
    }
    __syncthreads();
    for (int e = t + 2 * TR; e < (NK - 2) * TR; e += B)
        fl[e] = flatten5(c0[e - 2 * TR], c0[e - TR], c0[e], c0[e + TR], c0[e + 2 * TR], c1[e - TR], c1[e], c1[e + TR], cs0);
    __syncthreads();
    real* cf[3] = {c0, c1, c2};
    for (int e = t + 2 * TR; e < (NK - 2) * TR; e += B) {
        for (int f = 0; f < 3; ++f) {
            const real* a = cf[f];
            real l, r;
            slopes(a[e - 2 * TR], a[e - TR], a[e], a[e + TR], a[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4], wy[5], fl[e], l, r);
            sl[f][e] = l; sr[f][e] = r;
        }
    }
    __syncthreads();
    auto state_in = [&](int e, real* in) {
        in[TC_R] = c0[e]; in[TC_U] = c1[e]; in[TC_RSL] = sl[0][e]; in[TC_RSR] = sr[0][e];
        in[TC_USL] = sl[1][e]; in[TC_USR] = sr[1][e]; in[TC_VSL] = sl[2][e]; in[TC_V] = c2[e]; in[TC_VSR] = sr[2][e];
    };
    // forward fluxes of interfaces [3, NK-3] (interface k: between rows k-1 and k)
    for (int e = t + NPAD * TR; e < (NK - 2) * TR; e += B) {
        real inL[TC_N], inR[TC_N];
        state_in(e - TR, inL);
        state_in(e, inR);
        real rl, ul, vl, rr, ur, vr;
        trace_left_c(cs0, ic0, dts, idy, inL[0], inL[1], inL[2], inL[3], inL[4], inL[5], inL[6], inL[7], inL[8], rl, ul, vl);
        trace_right_c(cs0, ic0, dts, idy, inR[0], inR[1], inR[2], inR[3], inR[4], inR[5], inR[6], inR[7], inR[8], rr, ur, vr);
        real Fr, Fu, Fw, pres;
        hllc_c(cs0, ic0, cs2, rl, ul, vl, rr, ur, vr, Fr, Fu, Fw, pres);
        F0[e] = Fr; F1[e] = Fu; F2[e] = Fw; F3[e] = pres;
    }
    __syncthreads();
    // reverse of the update, rows [3, NK-3); the cell adjoints stay in
    // registers (thread t owns cells t + q B in every later phase)
    real a0[NT], a1[NT], a2[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int e = t + q * B;
        const int k = e / TR;
        real l0 = 0, l1 = 0, l2 = 0, ca = 0, cb = 0, cc = 0;
        if (e < E && live && k >= NPAD && k < NK - NPAD) {
            const size_t o = (size_t)wrap(jbase + k, ny) * nx + i;
            const real lrn = A.lo[o], lun = A.lo[N + o], lvn = A.lo[2 * N + o];
            const real Dr = (F0[e] - F0[e + TR]) * fup;
            const real Du = (F1[e] - F1[e + TR] + (F3[e] - F3[e + TR])) * fup;
            const real Dw = (F2[e] - F2[e + TR]) * fup;
            const real r = c0[e], vp = c1[e], u = c2[e];
            const real rn = r + Dr;
            const real rmin = fmax(r * real(1.0e-10), real(1.0e-10));
            if (rn < rmin) {
                l0 = r * real(1.0e-10) > real(1.0e-10) ? lrn * real(1.0e-10) : real(0);
                l1 = lvn;
                l2 = lun;
            } else {
                const real irn = FM<real>::rcp(rn);
                const real vn = (r * vp + Du) * irn, un = (r * u + Dw) * irn;
                const real lrt = lrn - (lvn * vn + lun * un) * irn;
                l0 = lrt + (lvn * vp + lun * u) * irn;
                l1 = lvn * r * irn;
                l2 = lun * r * irn;
                ca = lrt * fup; cb = lvn * irn * fup; cc = lun * irn * fup;
            }
        }
        a0[q] = l0; a1[q] = l1; a2[q] = l2;
        if (e < E) { cA[e] = ca; cB[e] = cb; cC[e] = cc; }
    }
    __syncthreads();
    // reverse of every interface in one pass: interface e (rows k-1 | k)
    // is reversed by the thread owning cell e, which keeps what it sends to
    // that cell and writes what it sends to cell e - TR into CL
    real lR[NT][TC_N];
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int e = t + q * B;
        real lL[TC_N];
        for (int m = 0; m < TC_N; ++m) { lL[m] = real(0); lR[q][m] = real(0); }
        if (e >= E) break;
        const bool iface = e >= NPAD * TR && e < (NK - 2) * TR;
        const real lFr = iface ? cA[e] - cA[e - TR] : real(0), lFu = iface ? cB[e] - cB[e - TR] : real(0);
        const real lFw = iface ? cC[e] - cC[e - TR] : real(0);
        const real lpres = lFu;   // the pressure difference enters Du like the momentum flux
        if (lFr != real(0) || lFu != real(0) || lFw != real(0)) {
            real inL[TC_N], inR[TC_N];
            state_in(e - TR, inL);
            state_in(e, inR);
            real Sc[6];
            trace_left_c(cs0, ic0, dts, idy, inL[0], inL[1], inL[2], inL[3], inL[4], inL[5], inL[6], inL[7], inL[8],
                         Sc[0], Sc[1], Sc[2]);
            trace_right_c(cs0, ic0, dts, idy, inR[0], inR[1], inR[2], inR[3], inR[4], inR[5], inR[6], inR[7], inR[8],
                          Sc[3], Sc[4], Sc[5]);
            real lSc[6] = {0, 0, 0, 0, 0, 0};
            hllc_c_adj(cs0, ic0, cs2, Sc, lFr, lFu, lFw, lpres, lSc);
            trace_left_c_adj(cs0, ic0, dts, idy, inL, lSc[0], lSc[1], lSc[2], lL);
            trace_right_c_adj(cs0, ic0, dts, idy, inR, lSc[3], lSc[4], lSc[5], lR[q]);
        }
        for (int m = 0; m < TC_N; ++m) CL[m][e] = lL[m];
    }
    __syncthreads();
    // gather per cell (right cell of interface k, left cell of interface
    // k+1), then the reverse of its parabolas: adjoints of the neighbour
    // differences and of the flattening coefficient (into W)
    real (*ld)[4][E] = reinterpret_cast<real (*)[4][E]>(&W[0][0]);
    real* lfl = W[12];
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int e = t + q * B;
        if (e >= E) break;
        const int k = e / TR;
        real g[TC_N];
        for (int m = 0; m < TC_N; ++m) g[m] = lR[q][m];   // zero unless interface e exists
        if (k + 1 >= NPAD && k + 1 <= NK - 3)
            for (int m = 0; m < TC_N; ++m) g[m] += CL[m][e + TR];
        a0[q] += g[TC_R]; a1[q] += g[TC_U]; a2[q] += g[TC_V];
        const real asl[3] = {g[TC_RSL], g[TC_USL], g[TC_VSL]}, asr[3] = {g[TC_RSR], g[TC_USR], g[TC_VSR]};
        real lflat = real(0);
        real ldf[3][4];
        for (int f = 0; f < 3; ++f)
            for (int m = 0; m < 4; ++m) ldf[f][m] = real(0);
        if (k >= 2 && k < NK - 2) {
            for (int f = 0; f < 3; ++f) {
                if (asl[f] == real(0) && asr[f] == real(0)) continue;
                const real* a = cf[f];
                slopes_adj(a[e - 2 * TR], a[e - TR], a[e], a[e + TR], a[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4],
                           wy[5], fl[e], asl[f], asr[f], ldf[f], lflat);
            }
        }
        // (W's parabolas and fluxes are dead: the interface pass is done)
        for (int f = 0; f < 3; ++f)
            for (int m = 0; m < 4; ++m) ld[f][m][e] = ldf[f][m];
        lfl[e] = lflat;
    }
    __syncthreads();
    real (*lfp)[E] = CL;
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int e = t + q * B;
        if (e >= E) break;
        const int k = e / TR;
        real add[3];
        for (int f = 0; f < 3; ++f) {
            real v = real(0);
            if (k + 2 < NK) v -= ld[f][0][e + 2 * TR];
            if (k + 1 < NK) v += ld[f][0][e + TR] - ld[f][1][e + TR];
            v += ld[f][1][e] - ld[f][2][e];
            if (k >= 1) v += ld[f][2][e - TR] - ld[f][3][e - TR];
            if (k >= 2) v += ld[f][3][e - 2 * TR];
            add[f] = v;
        }
        a0[q] += add[0]; a1[q] += add[1]; a2[q] += add[2];
        real lp5[5] = {0, 0, 0, 0, 0};
        if (k >= 2 && k < NK - 2 && lfl[e] != real(0))
            flatten5_adj(c0[e - 2 * TR], c0[e - TR], c0[e], c0[e + TR], c0[e + 2 * TR], c1[e - TR], c1[e], c1[e + TR],
                         lfl[e], lp5);
        for (int m = 0; m < 5; ++m) lfp[m][e] = lp5[m];
    }
    __syncthreads();
    const int rows = min(TY, ny - y0);
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int e = t + q * B;
        const int k = e / TR;
        const int y = k - 2 * NPAD;
        if (e >= E || !live || y < 0 || y >= rows) continue;
        real lr = a0[q] + lfp[0][e + 2 * TR] + lfp[1][e + TR] + lfp[2][e] + lfp[3][e - TR] + lfp[4][e - 2 * TR];
        const size_t o = (size_t)(y0 + y) * nx + i;
        put(A.li, o, lr, A.accumulate);
        put(A.li, N + o, a2[q], A.accumulate);
        put(A.li, 2 * N + o, a1[q], A.accumulate);
    }
}

// ----------------------------- orbital shift + residual advection adjoint
// Forward: S3[y] = advect(SA shifted by inc)[y], i.e. post-shift row y
// takes pre-shift row y - inc.  Tile of TR radii x TY pre-shift rows y0..;
// LDS row k <-> pre-shift row y0-6+k <-> post-shift row y0-6+k+inc.
template <typename real>
__global__ __launch_bounds__(AdjPhiTile<real>::TR * AdjPhiTile<real>::TY) void k_adj_adv(AdjArgs<real> A) {
    constexpr int TR = AdjPhiTile<real>::TR, TY = AdjPhiTile<real>::TY, B = TR * TY;
    constexpr int NK = TY + 4 * NPAD, E = NK * TR;
    constexpr int NT = (E + B - 1) / B;   // cells per thread
    __shared__ real c0[E], c1[E], c2[E];
    // W: parabolas (6), fluxes (3) and update adjoints (3), then the
    // adjoints of the neighbour differences (12)
    __shared__ real W[12][E];
    real (*sl)[E] = W;
    real (*sr)[E] = W + 3;
    real *F0 = W[6], *F1 = W[7], *F2 = W[8];
    real *cA = W[9], *cB = W[10], *cC = W[11];
    const DiskArgs<real>& d = *A.d;
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int t = threadIdx.x;
    const int ri = t % TR;
    const int i = blockIdx.x * TR + ri;
    const bool live = i < nx;
    const int ic = live ? i : nx - 1;
    const int ip = ic + NPAD;
    const int y0 = blockIdx.y * TY;
    const real iR = TAB(T_IR, ip);
    const real dt = (real)d.ctl->dt;
    const real dts = dt * iR, idy = d.idy, fup = dt * idy * iR;
    const real wy[6] = {d.wLy[0], d.wLy[1], d.wLy[2], d.wRy[0], d.wRy[1], d.wRy[2]};
    const int inc = d.orbadv ? d.inc[ic] : 0;
    const real speed = d.res[ic];
    const real q = lim01(fabs(speed) * dts * idy);
    const int jbase = y0 - 2 * NPAD;
    static_assert(B % TR == 0, "tile");
    for (int e = t; e < E; e += B) {
        real r = real(1), vp = real(0), u = real(0);
        if (live) {
            const size_t o = (size_t)wrap(jbase + e / TR, ny) * nx + i;
            r = TAB(T_G_R, ip) + A.s[o]; vp = A.s[2 * N + o]; u = A.s[N + o];
        }
        c0[e] = r; c1[e] = vp; c2[e] = u;
    }
    __syncthreads();
    real* cf[3] = {c0, c1, c2};
    for (int e = t + 2 * TR; e < (NK - 2) * TR; e += B) {
        for (int f = 0; f < 3; ++f) {
            const real* a = cf[f];
            real l, r;
            slopes(a[e - 2 * TR], a[e - TR], a[e], a[e + TR], a[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4], wy[5], real(0), l, r);
            sl[f][e] = l; sr[f][e] = r;
        }
    }
    __syncthreads();
    // upwind cell of interface e: e-TR when speed > 0, else e
    const int up = speed > real(0) ? -TR : 0;
    for (int e = t + NPAD * TR; e < (NK - 2) * TR; e += B) {
        const int m = e + up;
        real Dr, Dv, Du;
        if (speed > real(0)) {
            Dr = aveR(q, sl[0][m], c0[m], sr[0][m]);
            Dv = aveR(q, sl[1][m], c1[m], sr[1][m]) * Dr;
            Du = aveR(q, sl[2][m], c2[m], sr[2][m]) * Dr;
        } else {
            Dr = aveL(q, sl[0][m], c0[m], sr[0][m]);
            Dv = aveL(q, sl[1][m], c1[m], sr[1][m]) * Dr;
            Du = aveL(q, sl[2][m], c2[m], sr[2][m]) * Dr;
        }
        F0[e] = Dr * speed; F1[e] = Dv * speed; F2[e] = Du * speed;
    }
    __syncthreads();
    // reverse of the update (update_Cell_advect, no floor), cells [3, NK-3);
    // the cell adjoints stay in registers (thread t owns cells t + q B)
    real a0[NT], a1[NT], a2[NT];
#pragma unroll
    for (int qq = 0; qq < NT; ++qq) {
        const int e = t + qq * B;
        const int k = e / TR;
        real l0 = 0, l1 = 0, l2 = 0, ca = 0, cb = 0, cc = 0;
        if (e < E && live && k >= NPAD && k < NK - NPAD) {
            const size_t o = (size_t)wrap(jbase + k + inc, ny) * nx + i;
            const real lrn = A.lo[o], lun = A.lo[N + o], lvn = A.lo[2 * N + o];
            const real Dr = (F0[e] - F0[e + TR]) * fup;
            const real Dv = (F1[e] - F1[e + TR]) * fup;
            const real Du = (F2[e] - F2[e + TR]) * fup;
            const real r = c0[e];
            const real rn = r + Dr;
            const real irn = FM<real>::rcp(rn);
            const real vn = (r * c1[e] + Dv) * irn, un = (r * c2[e] + Du) * irn;
            const real lrt = lrn - (lvn * vn + lun * un) * irn;
            l0 = lrt + (lvn * c1[e] + lun * c2[e]) * irn;
            l1 = lvn * r * irn;
            l2 = lun * r * irn;
            ca = lrt * fup; cb = lvn * irn * fup; cc = lun * irn * fup;
        }
        a0[qq] = l0; a1[qq] = l1; a2[qq] = l2;
        if (e < E) { cA[e] = ca; cB[e] = cb; cC[e] = cc; }
    }
    __syncthreads();
    // reverse of the fluxes: cell m is the upwind cell of exactly one
    // interface (m + TR when speed > 0, m otherwise), reversed by m's own
    // thread; the parabola adjoints of m stay in its registers
    real lsl[NT][3], lsr[NT][3];
#pragma unroll
    for (int qq = 0; qq < NT; ++qq) {
        const int m = t + qq * B;
        const int e = m - up;   // the interface m feeds
        for (int f = 0; f < 3; ++f) { lsl[qq][f] = real(0); lsr[qq][f] = real(0); }
        if (m >= E || e < NPAD * TR || e >= (NK - 2) * TR) continue;
        const real lF0 = (cA[e] - cA[e - TR]) * speed;
        const real lF1 = (cB[e] - cB[e - TR]) * speed;
        const real lF2 = (cC[e] - cC[e - TR]) * speed;
        real A0, A1, A2;
        if (speed > real(0)) {
            A0 = aveR(q, sl[0][m], c0[m], sr[0][m]);
            A1 = aveR(q, sl[1][m], c1[m], sr[1][m]);
            A2 = aveR(q, sl[2][m], c2[m], sr[2][m]);
        } else {
            A0 = aveL(q, sl[0][m], c0[m], sr[0][m]);
            A1 = aveL(q, sl[1][m], c1[m], sr[1][m]);
            A2 = aveL(q, sl[2][m], c2[m], sr[2][m]);
        }
        // F0 = A0 s, F1 = A1 A0 s, F2 = A2 A0 s
        const real lA[3] = {lF0 + lF1 * A1 + lF2 * A2, lF1 * A0, lF2 * A0};
        real lq = 0, la[3] = {0, 0, 0};
        for (int f = 0; f < 3; ++f) {
            const real* a = cf[f];
            if (speed > real(0)) aveR_adj(q, sl[f][m], a[m], sr[f][m], lA[f], lq, lsl[qq][f], la[f], lsr[qq][f]);
            else aveL_adj(q, sl[f][m], a[m], sr[f][m], lA[f], lq, lsl[qq][f], la[f], lsr[qq][f]);
        }
        a0[qq] += la[0]; a1[qq] += la[1]; a2[qq] += la[2];
    }
    __syncthreads();   // W's parabolas, fluxes and update adjoints are dead from here
    real (*ld)[4][E] = reinterpret_cast<real (*)[4][E]>(&W[0][0]);
#pragma unroll
    for (int qq = 0; qq < NT; ++qq) {
        const int e = t + qq * B;
        if (e >= E) break;
        const int k = e / TR;
        real ldf[3][4];
        for (int f = 0; f < 3; ++f)
            for (int m = 0; m < 4; ++m) ldf[f][m] = real(0);
        if (k >= 2 && k < NK - 2) {
            real lflat = 0;
            for (int f = 0; f < 3; ++f) {
                if (lsl[qq][f] == real(0) && lsr[qq][f] == real(0)) continue;
                const real* a = cf[f];
                slopes_adj(a[e - 2 * TR], a[e - TR], a[e], a[e + TR], a[e + 2 * TR], wy[0], wy[1], wy[2], wy[3], wy[4],
                           wy[5], real(0), lsl[qq][f], lsr[qq][f], ldf[f], lflat);
            }
        }
        for (int f = 0; f < 3; ++f)
            for (int m = 0; m < 4; ++m) ld[f][m][e] = ldf[f][m];
    }
    __syncthreads();
    const int rows = min(TY, ny - y0);
#pragma unroll
    for (int qq = 0; qq < NT; ++qq) {
        const int e = t + qq * B;
        const int k = e / TR;
        const int y = k - 2 * NPAD;
        if (e >= E || !live || y < 0 || y >= rows) continue;
        real add[3];
        for (int f = 0; f < 3; ++f) {
            add[f] = -ld[f][0][e + 2 * TR] + ld[f][0][e + TR] - ld[f][1][e + TR] + ld[f][1][e] - ld[f][2][e] +
                     ld[f][2][e - TR] - ld[f][3][e - TR] + ld[f][3][e - 2 * TR];
        }
        const size_t o = (size_t)(y0 + y) * nx + i;
        put(A.li, o, a0[qq] + add[0], A.accumulate);
        put(A.li, N + o, a2[qq] + add[2], A.accumulate);
        put(A.li, 2 * N + o, a1[qq] + add[1], A.accumulate);
    }
}

// ------------------------------------------------------ viscosity adjoint
// Forward: k_viscosity / vis_from_tile, a 9-point stencil of (rho nu, u, vp).
// Per output cell, the adjoints it sends to each input of its stencil:
// lu[n], lv[n] for the u and vp of neighbour n, ln[n] for its rho nu
// (nonzero for n = 0, 1, 2, 4, 7); neighbour n = 3b + a at radius i-1+a,
// row j, j-1, j+1 for b = 0, 1, 2.  Returns X fx + Y fy (the forces are
// linear in nu, hence in alpha).
template <typename real>
__device__ __forceinline__ double vis_adj_cell(const DiskArgs<real>& d, int ip, const real* nr, const real* u,
                                               const real* v, real X, real Y, real* lu, real* lv, real* ln) {
    const real i20 = TAB(T_I20, ip), i21 = TAB(T_I21, ip), i10 = TAB(T_I10, ip);
    const real ir = TAB(T_IR, ip), idy = d.idy, hdy = real(0.5) * idy;
    const real dudx = (u[2] - u[0]) * i20;
    const real d2udx2 = ((u[2] - u[1]) * i21 - (u[1] - u[0]) * i10) * (real(2) * i20);
    const real dvdx = TAB(T_DVBG, ip) + (v[2] - v[0]) * i20;
    const real d2vdx2 = TAB(T_D2VBG, ip) + ((v[2] - v[1]) * i21 - (v[1] - v[0]) * i10) * (real(2) * i20);
    const real dndx = (nr[2] - nr[0]) * i20;
    const real ky = hdy * ir, kyy = (idy * idy) * ir * ir, kxy = i20 * hdy * ir;
    const real dvdy = (v[7] - v[4]) * ky;
    const real d2vdy2 = (v[7] - real(2) * v[1] + v[4]) * kyy;
    const real d2vdxy = (v[8] - v[6] - v[5] + v[3]) * kxy;
    const real dudy = (u[7] - u[4]) * ky;
    const real d2udy2 = (u[7] - real(2) * u[1] + u[4]) * kyy;
    const real d2udxy = (u[8] - u[6] - u[5] + u[3]) * kxy;
    const real dndy = (nr[7] - nr[4]) * ky;
    const real v1 = TAB(T_VBG, ip) + v[1];
    const real n1 = nr[1];
    double lal;
    {   // forward values (for the alpha adjoint)
        real sxx = real(2) * n1 * dudx;
        real fx = real(2) * dndx * dudx + real(2) * n1 * d2udx2 + sxx * ir;
        real syy = real(2) * n1 * (dvdy + u[1] * ir);
        real sxy = n1 * (dudy + dvdx - v1 * ir);
        real fy = real(2) * dndy * dvdy + real(2) * n1 * d2vdy2 + real(2) * n1 * dudy * ir + real(2) * dndy * u[1] * ir;
        fx += dndy * (dudy + dvdx) + n1 * (d2udy2 + d2vdxy) - dndy * v1 * ir - n1 * dvdy * ir - syy * ir;
        fy += dndx * (dudy + dvdx) + n1 * (d2udxy + d2vdx2) - dndx * v1 * ir - n1 * dvdx * ir + n1 * v1 * ir * ir +
              real(2) * sxy * ir;
        lal = (double)X * (double)fx + (double)Y * (double)fy;
    }
    // adjoints of the derivative terms
    const real l_dndx = X * (real(2) * dudx) + Y * ((dudy + dvdx) - v1 * ir);
    const real l_dudx = X * (real(2) * dndx + real(2) * n1 * ir);
    const real l_n1 = X * (real(2) * d2udx2 + real(2) * dudx * ir + (d2udy2 + d2vdxy) - dvdy * ir -
                           real(2) * (dvdy + u[1] * ir) * ir) +
                      Y * (real(2) * d2vdy2 + real(2) * dudy * ir + (d2udxy + d2vdx2) - dvdx * ir + v1 * ir * ir +
                           real(2) * (dudy + dvdx - v1 * ir) * ir);
    const real l_d2udx2 = X * real(2) * n1;
    const real l_dndy = X * ((dudy + dvdx) - v1 * ir) + Y * (real(2) * dvdy + real(2) * u[1] * ir);
    const real l_dudy = X * dndy + Y * (real(4) * n1 * ir + dndx);
    const real l_dvdx = X * dndy + Y * (dndx + n1 * ir);
    const real l_d2udy2 = X * n1, l_d2vdxy = X * n1;
    const real l_v1 = X * (-dndy * ir) + Y * (-dndx * ir - n1 * ir * ir);
    const real l_dvdy = X * (real(-3) * n1 * ir) + Y * (real(2) * dndy);
    const real l_u1 = X * (real(-2) * n1 * ir * ir) + Y * (real(2) * dndy * ir);
    const real l_d2vdy2 = Y * real(2) * n1, l_d2udxy = Y * n1, l_d2vdx2 = Y * n1;
    for (int n = 0; n < 9; ++n) { lu[n] = real(0); lv[n] = real(0); ln[n] = real(0); }
    lu[2] += l_dudx * i20; lu[0] -= l_dudx * i20;
    {
        const real c = real(2) * i20 * l_d2udx2;
        lu[2] += c * i21; lu[1] -= c * (i21 + i10); lu[0] += c * i10;
    }
    lv[2] += l_dvdx * i20; lv[0] -= l_dvdx * i20;
    {
        const real c = real(2) * i20 * l_d2vdx2;
        lv[2] += c * i21; lv[1] -= c * (i21 + i10); lv[0] += c * i10;
    }
    ln[2] += l_dndx * i20; ln[0] -= l_dndx * i20;
    lv[7] += l_dvdy * ky; lv[4] -= l_dvdy * ky;
    lv[7] += l_d2vdy2 * kyy; lv[1] -= real(2) * l_d2vdy2 * kyy; lv[4] += l_d2vdy2 * kyy;
    lv[8] += l_d2vdxy * kxy; lv[6] -= l_d2vdxy * kxy; lv[5] -= l_d2vdxy * kxy; lv[3] += l_d2vdxy * kxy;
    lu[7] += l_dudy * ky; lu[4] -= l_dudy * ky;
    lu[7] += l_d2udy2 * kyy; lu[1] -= real(2) * l_d2udy2 * kyy; lu[4] += l_d2udy2 * kyy;
    lu[8] += l_d2udxy * kxy; lu[6] -= l_d2udxy * kxy; lu[5] -= l_d2udxy * kxy; lu[3] += l_d2udxy * kxy;
    ln[7] += l_dndy * ky; ln[4] -= l_dndy * ky;
    lv[1] += l_v1;
    ln[1] += l_n1;
    lu[1] += l_u1;
    return lal;
}

// Source-term adjoint of one cell.  Forward (k_sources / cell_sources, gas
// without dust or wave killing, h = dt/2):
//   mode 0:  u' = u + (ebal + vp (2 vbg + vp)/R + px + Vr/rho) h,
//            vp' = vp + (Vphi/rho + py) h
//   mode 1/2: the same with vp in the centrifugal term replaced by the
//            predicted W = vp + (Vphi/rho + py) h.
// In: the state (rho, vp), the forces V, the output adjoints ln*; out: the
// adjoints of rho, u, vp, V, and lq += dL/d(planet mass) * q.
template <typename real>
__device__ __forceinline__ void src_adj_cell(const DiskArgs<real>& d, int mode, int ip, int j, real h, real rho,
                                             real cvp, real Vr, real Vp, real lnr, real lnu, real lnv, real& lrho,
                                             real& lu, real& lcvp, real& lVr, real& lVp, double& lq) {
    const PlanetK& pk = mode == 0 ? d.ctl->pk0 : d.ctl->pk1;
    const real iR = TAB(T_IR, ip), vbg = TAB(T_VBG, ip);
    real px = real(0), py = real(0);
    if (d.n_planet > 0) planet_acc(pk, d.phitab + 4 * j, TAB(T_XC, ip), px, py);
    const real irho = FM<real>::rcp(rho);
    if (!d.visc) { Vr = real(0); Vp = real(0); }
    const real vy = Vp * irho;
    const real W = mode == 0 ? cvp : cvp + (vy + py) * h;
    const real lW = lnu * h * (real(2) * vbg + real(2) * W) * iR;
    const real lpx = lnu * h, lvx = lnu * h;
    real lvy = lnv * h, lpy = lnv * h;
    if (mode != 0) { lvy += lW * h; lpy += lW * h; }
    lrho = lnr;
    lu = lnu;
    lcvp = lnv + lW;
    lVr = real(0);
    lVp = real(0);
    if (d.visc) {
        lVr = lvx * irho;
        lVp = lvy * irho;
        lrho -= (lvx * Vr + lvy * Vp) * irho * irho;
    }
    lq += (double)lpx * (double)px + (double)lpy * (double)py;
}

// Viscosity adjoint on tiles of AV_X radii x TH rows, fused with the
// pointwise passes around it.  The outputs whose stencil meets the tile
// (the tile and a one-cell halo) compute their partial adjoints into LDS
// from the forward state staged with a two-cell halo; each tile cell then
// gathers what its 3x3 neighbours sent.
//   MODE 0: V = vis(S): lvis -> li (module entry point).
//   MODE 1: end of the step: corrector S4 = C(S3, V1) (lo = adjoint of S4,
//           s = S3, vis = V1), V1 = vis(T) (s2 = T), predictor
//           T = P(S3, V0) (vis2 = V0): li = adjoint of S3, lvis = of V0.
//   MODE 2: start of the step: kick S1 = K(S0, V0) (lo = adjoint of S1,
//           s = S0, vis = V0), V0 = vis(S0) with lvis holding what the
//           predictor sent to V0: li = adjoint of S0.
constexpr int AV_X = 64;

template <typename real, int MODE, int TH>
__global__ __launch_bounds__(AV_X * TH) void k_adj_vis(AdjArgs<real> A) {
    constexpr int B = AV_X * TH;
    constexpr int H1X = AV_X + 2, H1 = H1X * (TH + 2);   // outputs: tile + 1
    constexpr int H2X = AV_X + 4, H2 = H2X * (TH + 4);   // their state: tile + 2
    __shared__ real sr[H2], su[H2], sv[H2];
    __shared__ real P[23][H1];
    const DiskArgs<real>& d = *A.d;
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * AV_X, j0 = blockIdx.y * TH;
    const real h = (real)(0.5 * d.ctl->dt);
    const real* vs = MODE == 1 ? A.s2 : A.s;   // the state the viscosity read
    for (int e = t; e < H2; e += B) {
        const int lx = e % H2X, ly = e / H2X;
        const int i = i0 - 2 + lx, j = wrap(j0 - 2 + ly, ny);
        real r, u, v;
        if (i < 0 || i >= nx) {   // radial ghosts (the second halo column feeds no output)
            const int ip = min(max(i, -NPAD), nx + NPAD - 1) + NPAD;
            r = TAB(T_G_R, ip); u = TAB(T_G_U, ip); v = TAB(T_G_V, ip);
        } else {
            const size_t o = (size_t)j * nx + i;
            r = TAB(T_G_R, i + NPAD) + vs[o]; u = vs[N + o]; v = vs[2 * N + o];
        }
        sr[e] = r; su[e] = u; sv[e] = v;
    }
    __syncthreads();
    double lal = 0.0, lq = 0.0;
    for (int e = t; e < H1; e += B) {
        const int lx = e % H1X, ly = e / H1X;
        const int i = i0 - 1 + lx, jr = j0 - 1 + ly, j = wrap(jr, ny);
        real lu[9], lv[9], ln[9];
        for (int n = 0; n < 9; ++n) { lu[n] = real(0); lv[n] = real(0); ln[n] = real(0); }
        real nu[3] = {real(0), real(0), real(0)};
        if (i >= 0 && i < nx) {
            const int ip = i + NPAD;
            const size_t o = (size_t)j * nx + i;
            real X, Y;
            double lq_unused = 0.0;
            if (MODE == 0) {
                X = A.lvis[o]; Y = A.lvis[N + o];
            } else {
                // adjoint of the forces from the source pass that read them
                real a0, a1, a2;
                src_adj_cell(d, MODE == 1 ? 2 : 0, ip, j, h, TAB(T_G_R, ip) + A.s[o], A.s[2 * N + o], A.vis[o], A.vis[N + o], A.lo[o],
                             A.lo[N + o], A.lo[2 * N + o], a0, a1, a2, X, Y, lq_unused);
                if (MODE == 2) { X += A.lvis[o]; Y += A.lvis[N + o]; }
            }
            for (int n = 0; n < 3; ++n) nu[n] = TAB(T_NU, ip - 1 + n);
            real nr[9], u[9], v[9];
            const int rows[3] = {ly + 1, ly, ly + 2};   // H2 rows of j, j-1, j+1
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int n = 3 * b + a, s = rows[b] * H2X + lx + a;
                    nr[n] = sr[s] * nu[a]; u[n] = su[s]; v[n] = sv[s];
                }
            const double l = vis_adj_cell(d, ip, nr, u, v, X, Y, lu, lv, ln);
            if (lx >= 1 && lx <= AV_X && ly >= 1 && ly <= TH && jr < ny) lal += l;   // each output once
        }
        for (int n = 0; n < 9; ++n) { P[n][e] = lu[n]; P[9 + n][e] = lv[n]; }
        P[18][e] = ln[0] * nu[0]; P[19][e] = ln[1] * nu[1]; P[20][e] = ln[2] * nu[2];
        P[21][e] = ln[4] * nu[1]; P[22][e] = ln[7] * nu[1];
    }
    __syncthreads();
    // what the outputs sent to the cell at H1 slot (cx, cy): output
    // (cx + 1 - a, cy - db) holds it as its neighbour n = 3b + a
    // (slots outside the H1 columns hold no output: skipped for the ghost
    // columns, whose outer neighbours are not radii of the disk)
    auto gather = [&](int cx, int cy, real& gr, real& gu, real& gv) {
        gr = gu = gv = real(0);
        const int db[3] = {0, -1, 1};
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int ox = cx + 1 - a;
                if (ox < 0 || ox >= H1X) continue;
                const int n = 3 * b + a, e = (cy - db[b]) * H1X + ox;
                gu += P[n][e];
                gv += P[9 + n][e];
                if (n == 0) gr += P[18][e];
                else if (n == 1) gr += P[19][e];
                else if (n == 2) gr += P[20][e];
                else if (n == 4) gr += P[21][e];
                else if (n == 7) gr += P[22][e];
            }
    };
    const int cx = t % AV_X + 1, cy = t / AV_X + 1;
    const int i = i0 + cx - 1, j = j0 + cy - 1;
    if (j < ny) {
        // radial ghost columns next to this cell: their u = -1.5 nu/R is
        // proportional to alpha
        if (i == 0) {
            real gr, gu, gv;
            gather(cx - 1, cy, gr, gu, gv);
            lal += (double)gu * (double)TAB(T_G_U, -1 + NPAD);
        }
        if (i == nx - 1) {
            real gr, gu, gv;
            gather(cx + 1, cy, gr, gu, gv);
            lal += (double)gu * (double)TAB(T_G_U, nx + NPAD);
        }
        if (i < nx) {
            const int ip = i + NPAD;
            const size_t o = (size_t)j * nx + i;
            real gr, gu, gv;
            gather(cx, cy, gr, gu, gv);
            if (MODE == 0) {
                put(A.li, o, gr, A.accumulate & 1);
                put(A.li, N + o, gu, A.accumulate & 1);
                put(A.li, 2 * N + o, gv, A.accumulate & 1);
            } else if (MODE == 1) {
                real cr, cu, cv, cX, cY, pr, pu, pv, pX, pY;
                src_adj_cell(d, 2, ip, j, h, TAB(T_G_R, ip) + A.s[o], A.s[2 * N + o], A.vis[o], A.vis[N + o], A.lo[o], A.lo[N + o],
                             A.lo[2 * N + o], cr, cu, cv, cX, cY, lq);
                src_adj_cell(d, 1, ip, j, h, TAB(T_G_R, ip) + A.s[o], A.s[2 * N + o], A.vis2[o], A.vis2[N + o], gr, gu, gv, pr, pu, pv,
                             pX, pY, lq);
                A.li[o] = cr + pr; A.li[N + o] = cu + pu; A.li[2 * N + o] = cv + pv;
                A.lvis[o] = pX; A.lvis[N + o] = pY;
            } else {
                real kr, ku, kv, kX, kY;
                src_adj_cell(d, 0, ip, j, h, TAB(T_G_R, ip) + A.s[o], A.s[2 * N + o], A.vis[o], A.vis[N + o], A.lo[o], A.lo[N + o],
                             A.lo[2 * N + o], kr, ku, kv, kX, kY, lq);
                A.li[o] = kr + gr; A.li[N + o] = ku + gu; A.li[2 * N + o] = kv + gv;
            }
        }
    }
    if (MODE != 0 && A.q > 0.0) block_sum_add(A.lparam, lq / A.q);
    if (A.alpha > 0.0) block_sum_add(A.lparam + A.nslot, lal / A.alpha);
}

// ---------------------------------------------------- source-term adjoint
// k_sources without viscosity fused around it (module entry points, and
// steps without viscosity).  accumulate: bit 0 -> li +=, bit 1 -> lvis +=.
template <typename real>
__global__ __launch_bounds__(256) void k_adj_src(AdjArgs<real> A) {
    const DiskArgs<real>& d = *A.d;
    const int nx = d.nx, ny = d.ny;
    const size_t N = (size_t)nx * ny;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    double lq = 0.0;
    if (i < nx) {
        const int ip = i + NPAD;
        const size_t o = (size_t)j * nx + i;
        const real h = (real)(0.5 * d.ctl->dt);
        real Vr = 0, Vp = 0;
        if (d.visc) { Vr = A.vis[o]; Vp = A.vis[N + o]; }
        real lrho, lu, lcvp, lVr, lVp;
        src_adj_cell(d, A.mode, ip, j, h, TAB(T_G_R, ip) + A.s[o], A.s[2 * N + o], Vr, Vp, A.lo[o], A.lo[N + o], A.lo[2 * N + o], lrho,
                     lu, lcvp, lVr, lVp, lq);
        put(A.li, o, lrho, A.accumulate & 1);
        put(A.li, N + o, lu, A.accumulate & 1);
        put(A.li, 2 * N + o, lcvp, A.accumulate & 1);
        if (d.visc) {
            put(A.lvis, o, lVr, A.accumulate & 2);
            put(A.lvis, N + o, lVp, A.accumulate & 2);
        }
    }
    if (A.q > 0.0) block_sum_add(A.lparam, lq / A.q);
}

// ------------------------------------------------------------- launchers
template <typename real>
struct AdjLaunch {
    // 116-cell tiles of two waves (ADJ_TX): a 512-radius row takes four and
    // a 48-radius one, 9 wave passes with one idle wave, where 244-cell
    // tiles of four waves took 12 with three idle; and a CU holds more of
    // the smaller workgroups
    static constexpr int TXA = ADJ_TX;
    static hipError_t x(const AdjArgs<real>& A, int nx, int ny, hipStream_t s) {
        hipLaunchKernelGGL((k_adj_x<real, TXA>), dim3((nx + TXA - 1) / TXA, ny), dim3(AdjXBlock<TXA>::B), 0, s, A);
        return hipGetLastError();
    }
    static hipError_t phi(const AdjArgs<real>& A, int nx, int ny, hipStream_t s) {
        constexpr int TR = AdjPhiTile<real>::TR, TY = AdjPhiTile<real>::TY;
        hipLaunchKernelGGL(k_adj_phi<real>, dim3((nx + TR - 1) / TR, (ny + TY - 1) / TY), dim3(TR * TY), 0, s, A);
        return hipGetLastError();
    }
    static hipError_t adv(const AdjArgs<real>& A, int nx, int ny, hipStream_t s) {
        constexpr int TR = AdjPhiTile<real>::TR, TY = AdjPhiTile<real>::TY;
        hipLaunchKernelGGL(k_adj_adv<real>, dim3((nx + TR - 1) / TR, (ny + TY - 1) / TY), dim3(TR * TY), 0, s, A);
        return hipGetLastError();
    }
    // viscosity tiles: 64 x 8 (two blocks per CU in FP32)
    static constexpr int THV = 8;
    // mode 0: viscosity alone; 1: corrector + viscosity + predictor;
    // 2: kick + viscosity (AdjArgs field use: k_adj_vis)
    static hipError_t vis(const AdjArgs<real>& A, int mode, int nx, int ny, hipStream_t s) {
        const dim3 grid((nx + AV_X - 1) / AV_X, (ny + THV - 1) / THV), block(AV_X * THV);
        if (mode == 1) hipLaunchKernelGGL((k_adj_vis<real, 1, THV>), grid, block, 0, s, A);
        else if (mode == 2) hipLaunchKernelGGL((k_adj_vis<real, 2, THV>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((k_adj_vis<real, 0, THV>), grid, block, 0, s, A);
        return hipGetLastError();
    }
    static hipError_t src(const AdjArgs<real>& A, int nx, int ny, hipStream_t s) {
        hipLaunchKernelGGL(k_adj_src<real>, dim3((nx + 255) / 256, ny), dim3(256), 0, s, A);
        return hipGetLastError();
    }
};

template struct AdjLaunch<float>;
template struct AdjLaunch<double>;

}  // namespace pg
