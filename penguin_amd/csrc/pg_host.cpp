// pg_host.cpp -- host module code and C ABI of the MI355X Godunov-sweep
// library (include/penguin_amd.h).
//
// Host-side responsibilities, by reference module:
//   geom / init   axis edges and volumes (geom.cu:3-33, init.cu:472-500),
//                 analytic disk profiles (init.cu:7-454) -> per-radius tables
//                 and the initial state (fill_grid, init.cu:502-519)
//   planet        initial orbit and softening (planet.cu:308-403)
//   orbit         background rotation rot(R) (orbit.cu:51-72)
//   solver        the DS split order of one step (solve.cu:212-292) as a
//                 sequence of kernel launches on the handle's stream
//   timestep      CFL dt with the 1.1x growth cap (timestep.cu:113-140)
//   main loop     device-resident time loop (main.cu:320-359)
// All arithmetic here is FP64; tables are cast to the storage precision on
// upload.  The device kernels are in pg_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/penguin_amd.h"
#include "pg_internal.h"
#include "pg_launch.h"

using namespace pg;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(PG_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),       \
                        __FILE__, __LINE__);                                                   \
    } while (0)

const double kPi = 3.1415926535897932384626434;

// ------------------------------------------------------------- profiles
struct Profiles {
    const pg_params& P;
    double cs2(double x) const {  // get_cs2 (init.cu:7-17), geomx 1
        return P.sc_h * P.sc_h / P.planet_radius * std::pow(x / P.planet_radius, -P.p_beta);
    }
    double h(double x) const { return std::sqrt(cs2(x) * x * x * x); }  // get_h (init.cu:20-29)
    double nu(double x) const {  // get_nu (init.cu:31-50)
        if (!P.visc_flag) return 0.0;
        double rp = P.planet_radius;
        return P.ss_alpha * std::sqrt(cs2(rp)) * h(rp) * std::pow(x / rp, P.p_alpha);
    }
    double rho(double x) const { return std::pow(x / P.planet_radius, -P.p_alpha); }  // init_flag 2
    double p(double x) const { return cs2(x) * rho(x); }
    double u(double x) const { return -1.5 * nu(x) / x; }
    double v(double x) const {  // pressure-supported rotation, init.cu:248-252
        const double dr = 1.1514178e-7;
        double dpdr = (p(x + dr) - p(x - dr)) / (2.0 * dr);
        return std::sqrt(1.0 / x + (x / rho(x)) * dpdr);
    }
    double rho_d(double x) const { return P.D_G_ratio * std::pow(x / P.planet_radius, -P.p_alpha); }
    double v_d(double x) const { return std::sqrt(1.0 / x); }
    double u_d(double x) const {  // steady radial drift, init.cu:365-380
        const double St = P.Stokes;
        double fg = v(x) / std::sqrt(1.0 / x);
        double a = St * St * fg;
        double L = ((1.0 + a) / a) * (1.0 - std::sqrt(1.0 - 2.0 * a * (1.0 - fg) / (1.0 + a) / (1.0 + a)));
        double u0 = St * (L * L - 2.0 * L);
        double w = (1.0 - std::sqrt(1.0 - 2.0 * St * u0)) / St;
        L = (1.0 - fg + St * w / 2.0) / (1.0 + St * w / 2.0);
        u0 = St * (L * L - 2.0 * L);
        w = (1.0 - std::sqrt(1.0 - 2.0 * St * u0)) / St;
        return w * std::sqrt(1.0 / x);
    }
};

// axis edges incl. ghosts and cell volumes (make_grid, init.cu:472-500)
void make_axis(std::vector<double>& a, std::vector<double>& v, double amin, double amax, int res,
               int pad, int geom, int grid) {
    a.assign(res + 2 * pad + 1, 0.0);
    v.assign(res + 2 * pad, 0.0);
    double lr = std::log(amax / amin);
    for (int i = 0; i <= res; ++i)
        a[pad + i] = grid == 1 ? amin * std::exp(lr * (double)i / (double)res)
                               : amin + (amax - amin) * (double)i / (double)res;
    for (int i = 0; i < pad; ++i) {
        if (grid == 1) {
            a[i] = amin * std::exp(lr * (double)(i - pad) / (double)res);
            a[res + pad + 1 + i] = amin * std::exp(lr * (double)(res + 1 + i) / (double)res);
        } else {
            a[i] = a[pad] - (a[2 * pad - i] - a[pad]);
            a[res + pad + 1 + i] = a[res + pad] + (a[res + pad] - a[res + pad - 1 - i]);
        }
    }
    for (int i = 0; i < res + 2 * pad; ++i) {
        double ra = a[i], dr = a[i + 1] - a[i];
        double dvdr = geom == 2 ? ra * ra + ra * dr + dr * dr / 3.0 : geom == 1 ? ra + dr / 2.0 : 1.0;
        v[i] = dr * dvdr;
    }
}

// Weights of the linear map from the differences of five cell means to the
// left and right slopes of the middle cell: the cubic with those cell means
// (means of x, x^2, x^3 under the geometric weight 1, x, x^2 for geom
// 0/3, 1, 2; recon.cu:242-357).  x6 = edges of the five cells.
void slope_weights(int geom, const double* x6, double wl[3], double wr[3]) {
    double M[5][3];
    for (int c = 0; c < 5; ++c) {
        double xl = x6[c], xr = x6[c + 1];
        if (geom == 2) {
            double d = xr * xr * xr - xl * xl * xl;
            M[c][0] = (std::pow(xr, 4) - std::pow(xl, 4)) / d * 0.75;
            M[c][1] = (std::pow(xr, 5) - std::pow(xl, 5)) / d * 0.6;
            M[c][2] = (std::pow(xr, 6) - std::pow(xl, 6)) / d * 0.5;
        } else if (geom == 1) {
            double d = xr * xr - xl * xl;
            M[c][0] = (std::pow(xr, 3) - std::pow(xl, 3)) / d / 1.5;
            M[c][1] = (std::pow(xr, 4) - std::pow(xl, 4)) / d / 2.0;
            M[c][2] = (std::pow(xr, 5) - std::pow(xl, 5)) / d / 2.5;
        } else {
            double d = xr - xl;
            M[c][0] = (xr * xr - xl * xl) / d / 2.0;
            M[c][1] = (xr * xr * xr - xl * xl * xl) / d / 3.0;
            M[c][2] = (std::pow(xr, 4) - std::pow(xl, 4)) / d / 4.0;
        }
    }
    // cubic c1 x + c2 x^2 + c3 x^3 (+c0) whose mean differences over cells
    // m..m+3 are d[0..2]: solve the 3x3 system by Cramer's rule
    auto solve = [&](int m, const double d[3], double c[3]) {
        double A[3][3];
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) A[r][k] = M[m + r + 1][k] - M[m + r][k];
        auto det3 = [](double a[3][3]) {
            return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                   a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                   a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
        };
        double D = det3(A);
        for (int k = 0; k < 3; ++k) {
            double B[3][3];
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) B[r][q] = q == k ? d[r] : A[r][q];
            c[k] = det3(B) / D;
        }
    };
    const double xL = x6[2], xR = x6[3];
    for (int k = 0; k < 3; ++k) {
        double e[3] = {0, 0, 0};
        e[k] = 1.0;
        double c[3];
        solve(0, e, c);  // cells 0..3 -> sL = <f>_2 - f(xL)
        wl[k] = c[0] * (M[2][0] - xL) + c[1] * (M[2][1] - xL * xL) + c[2] * (M[2][2] - xL * xL * xL);
        solve(1, e, c);  // cells 1..4 -> sR = f(xR) - <f>_2
        wr[k] = c[0] * (xR - M[2][0]) + c[1] * (xR * xR - M[2][1]) + c[2] * (xR * xR * xR - M[2][2]);
    }
}

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
// U in [0,1) keyed by (seed, field, interior i, interior j)
double keyed_uniform(uint64_t seed, int field, int i, int j) {
    uint64_t key = splitmix64(seed) ^ splitmix64(((uint64_t)field << 48) ^ ((uint64_t)j << 24) ^ (uint64_t)i);
    return (double)(splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

struct KernelTimer {
    double total_ms = 0.0;
    int count = 0;
};

}  // namespace

struct pg_handle {
    pg_params P{};
    int prec = 4;
    int nx = 0, ny = 0;
    size_t N = 0;
    hipStream_t stream = nullptr;
    // device memory
    void* gas = nullptr;
    void* gasT = nullptr;
    void* dust = nullptr;
    void* dustT = nullptr;
    void* vis = nullptr;
    void* tab = nullptr;
    void* phitab = nullptr;
    void* res = nullptr;
    int* inc = nullptr;
    double* rot = nullptr;
    double* dts = nullptr;
    int dts_cap = 0;
    double* rate_part = nullptr;
    int rate_nblk = 0;
    int tile_x = 256;   // x-sweep tile width (PG_TILE_X overrides)
    StepCtl* ctl = nullptr;
    StepCtl* ctl_host = nullptr;  // pinned mirror
    size_t bytes = 0;
    // host copies of what the state conversion needs
    std::vector<double> tabh;  // per-radius tables, field-major
    std::vector<double> phih;  // 4*ny azimuthal sine/cosine table
    int tab_stride = 0;
    double dy = 0.0;
    double wly[3]{}, wry[3]{};
    Body planet_init{};
    bool rate_valid = false;
    // profiling: event pairs recorded around launches without a host wait,
    // resolved after the stream drains
    bool profile = false;
    std::map<std::string, KernelTimer> timers;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kernel;   // kernel id of pair k (events 2k, 2k+1)
    size_t ev_used = 0;
    // graph of one device-loop step, and the buffers it was captured with
    hipGraphExec_t graph = nullptr;
    void* graph_gas = nullptr;
    void* graph_dust = nullptr;
    double* graph_dts = nullptr;

    template <typename real>
    DiskArgs<real> args() const {
        DiskArgs<real> d{};
        d.nx = nx;
        d.ny = ny;
        d.has_dust = P.dust_flag;
        d.visc = P.visc_flag;
        d.orbadv = P.OrbAdv_flag;
        d.kill = P.kill_flag;
        d.n_planet = P.n_planet;
        d.tile_x = tile_x;
        d.sweep_on = 1;
        d.gas = (real*)gas;
        d.gasT = (real*)gasT;
        d.dust = (real*)dust;
        d.dustT = (real*)dustT;
        d.vis = (real*)vis;
        d.tab = (const real*)tab;
        d.tab_stride = tab_stride;
        d.dy = (real)dy;
        d.idy = (real)(1.0 / dy);
        d.phitab = (const real*)phitab;
        for (int k = 0; k < 3; ++k) { d.wLy[k] = (real)wly[k]; d.wRy[k] = (real)wry[k]; }
        d.gam = (real)P.gam;
        d.sc_h = (real)P.sc_h;
        d.Stokes = (real)P.Stokes;
        d.D_G_ratio = (real)P.D_G_ratio;
        d.frame_omega = (real)P.frame_omega;
        d.CFL = P.CFL;
        d.dy_d = dy;
        d.rot = rot;
        d.inc = inc;
        d.res = (real*)res;
        d.ctl = ctl;
        d.ramp_time = P.ramp_time;
        d.planet_mass = P.planet_mass;
        d.planet_ecc = P.planet_ecc;
        d.planet_radius = P.planet_radius;
        d.rs = planet_init.rs;
        d.rate_part = rate_part;
        d.rate_nblk = rate_nblk;
        return d;
    }
};

namespace {

int validate(const pg_params& P) {
    if (P.ndim != 2) return fail(PG_E_INVALID, "ndim=%d: only 2D cylindrical disks are supported", P.ndim);
    if (P.geomx != 1 || P.geomy != 3)
        return fail(PG_E_INVALID, "geomx=%d geomy=%d: supported geometry is geomx=1 (cylindrical R), geomy=3 (azimuth)",
                    P.geomx, P.geomy);
    if (P.gridy != 0) return fail(PG_E_INVALID, "gridy=%d: the azimuthal grid must be uniform (0)", P.gridy);
    if (P.gridx != 0 && P.gridx != 1) return fail(PG_E_INVALID, "gridx=%d: supported radial grids are 0 and 1", P.gridx);
    if (P.recon_flag != 6) return fail(PG_E_INVALID, "recon_flag=%d: supported reconstruction is 6 (PPM4)", P.recon_flag);
    if (P.EOS_flag != 0) return fail(PG_E_INVALID, "EOS_flag=%d: supported EOS is 0 (locally isothermal)", P.EOS_flag);
    if (P.gam != 1.0) return fail(PG_E_INVALID, "gam=%g: EOS_flag 0 fixes gam = 1 (parameters.h:118-119)", P.gam);
    if (!P.OrbAdv_flag) return fail(PG_E_INVALID, "OrbAdv_flag=0 is not supported");
    if (P.bound_lft != 0 || P.bound_rgh != 0)
        return fail(PG_E_INVALID, "radial boundary types %d/%d: supported type is 0 (analytic)", P.bound_lft, P.bound_rgh);
    if (P.bound_bak != 3 || P.bound_frn != 3)
        return fail(PG_E_INVALID, "azimuthal boundary types %d/%d: supported type is 3 (periodic)", P.bound_bak, P.bound_frn);
    if (P.precision != 4 && P.precision != 8) return fail(PG_E_INVALID, "precision=%d: use 4 or 8", P.precision);
    if (P.xres < 8 || P.yres < 8) return fail(PG_E_INVALID, "resolution %dx%d too small (min 8x8)", P.xres, P.yres);
    if (P.n_planet != 0 && P.n_planet != 1) return fail(PG_E_INVALID, "n_planet=%d: 0 or 1", P.n_planet);
    if (!(P.xmin > 0.0 && P.xmax > P.xmin)) return fail(PG_E_INVALID, "bad radial domain");
    size_t esz = P.precision;
    size_t lds = (3 * (size_t)P.yres + 14 * (RING_THREADS + 2 * NPAD)) * esz;
    if (lds > 160 * 1024)
        return fail(PG_E_INVALID, "yres=%d at precision %d needs %zu B of LDS for the azimuthal ring (max 163840)",
                    P.yres, P.precision, lds);
    return PG_OK;
}

template <typename T>
int dev_alloc(pg_handle* h, T** p, size_t bytes) {
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) return fail(PG_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    h->bytes += bytes;
    return PG_OK;
}

template <typename real>
int upload_cast(const std::vector<double>& src, void* dst, hipStream_t s) {
    std::vector<real> tmp(src.size());
    for (size_t k = 0; k < src.size(); ++k) tmp[k] = (real)src[k];
    HIP_TRY(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(real), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return PG_OK;
}

int upload(pg_handle* h, const std::vector<double>& src, void* dst) {
    return h->prec == 8 ? upload_cast<double>(src, dst, h->stream) : upload_cast<float>(src, dst, h->stream);
}

template <typename real>
int download_cast(pg_handle* h, const void* src, std::vector<double>& dst, size_t n) {
    std::vector<real> tmp(n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), src, n * sizeof(real), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    dst.resize(n);
    for (size_t k = 0; k < n; ++k) dst[k] = (double)tmp[k];
    return PG_OK;
}

int download(pg_handle* h, const void* src, std::vector<double>& dst, size_t n) {
    return h->prec == 8 ? download_cast<double>(h, src, dst, n) : download_cast<float>(h, src, dst, n);
}

double tabv(const pg_handle* h, int f, int ip) { return h->tabh[(size_t)f * h->tab_stride + ip]; }

// Build the per-radius tables and the FP64 initial state (interior cells,
// device field order: gas rho,u,vp; dust rho,u,vp).
void build_tables_and_state(pg_handle* h, std::vector<double>& gas, std::vector<double>& dust,
                            std::vector<double>& rot) {
    const pg_params& P = h->P;
    Profiles pr{P};
    const int nx = h->nx, ny = h->ny;
    std::vector<double> xa, xv, ya, yv;
    make_axis(xa, xv, P.xmin, P.xmax, nx, NPAD, P.geomx, P.gridx);
    make_axis(ya, yv, P.ymin, P.ymax, ny, NPAD, P.geomy, P.gridy);
    const int S = nx + 2 * NPAD + 1;
    h->tab_stride = S;
    h->tabh.assign((size_t)T_NFIELDS * S, 0.0);
    auto T = [&](int f, int ip) -> double& { return h->tabh[(size_t)f * S + ip]; };
    for (int ip = 0; ip <= nx + 2 * NPAD; ++ip) T(T_XA, ip) = xa[ip];
    const int np = nx + 2 * NPAD;
    for (int ip = 0; ip < np; ++ip) {
        double xc = 0.5 * (xa[ip + 1] + xa[ip]);
        T(T_XV, ip) = xv[ip];
        T(T_XC, ip) = xc;
        T(T_IDX, ip) = 1.0 / (xa[ip + 1] - xa[ip]);
        T(T_CS2, ip) = pr.cs2(xc);
        double vb = pr.v(xc);
        T(T_VBG, ip) = vb;
        T(T_EBAL, ip) = vb * vb / xc - 1.0 / (xc * xc);
        T(T_NU, ip) = pr.nu(xc);
        T(T_G_R, ip) = pr.rho(xc);
        T(T_G_U, ip) = pr.u(xc);
        T(T_G_V, ip) = 0.0;
        if (P.dust_flag) {
            T(T_D_R, ip) = pr.rho_d(xc);
            T(T_D_U, ip) = pr.u_d(xc);
            T(T_D_V, ip) = pr.v_d(xc) - vb;
        }
    }
    for (int ip = 0; ip < np; ++ip) {
        T(T_C0, ip) = std::sqrt(P.gam * T(T_CS2, ip));
        T(T_IC0, ip) = 1.0 / T(T_C0, ip);
        T(T_IR, ip) = 1.0 / T(T_XC, ip);
        T(T_IXV, ip) = 1.0 / T(T_XV, ip);
        T(T_TS, ip) = P.Stokes / (T(T_C0, ip) / P.sc_h);
        T(T_ITS, ip) = 1.0 / T(T_TS, ip);
    }
    for (int ip = 1; ip < np - 1; ++ip) {
        double xm = T(T_XC, ip - 1), x0 = T(T_XC, ip), xp = T(T_XC, ip + 1);
        T(T_I20, ip) = 1.0 / (xp - xm);
        T(T_I21, ip) = 1.0 / (xp - x0);
        T(T_I10, ip) = 1.0 / (x0 - xm);
        double vm = T(T_VBG, ip - 1), v0 = T(T_VBG, ip), vpp = T(T_VBG, ip + 1);
        T(T_DVBG, ip) = (vpp - vm) / (xp - xm);
        T(T_D2VBG, ip) = ((vpp - v0) / (xp - x0) - (v0 - vm) / (x0 - xm)) / (0.5 * (xp - xm));
    }
    for (int ip = 2; ip < np - 2; ++ip) {
        double wl[3], wr[3];
        slope_weights(P.geomx, &xa[ip - 2], wl, wr);
        T(T_WL0, ip) = wl[0]; T(T_WL1, ip) = wl[1]; T(T_WL2, ip) = wl[2];
        T(T_WR0, ip) = wr[0]; T(T_WR1, ip) = wr[1]; T(T_WR2, ip) = wr[2];
    }
    slope_weights(P.geomy, &ya[NPAD - 2], h->wly, h->wry);
    h->dy = (P.ymax - P.ymin) / (double)ny;
    // azimuthal cell centres (0.5*(ya[j]+ya[j+1]), init.cu via Grid::get_yc)
    h->phih.assign(4 * (size_t)ny, 0.0);
    for (int j = 0; j < ny; ++j) {
        double phi = 0.5 * (ya[j + NPAD] + ya[j + NPAD + 1]);
        h->phih[4 * j + 0] = std::sin(phi);
        h->phih[4 * j + 1] = std::cos(phi);
        h->phih[4 * j + 2] = std::sin(0.5 * phi);
        h->phih[4 * j + 3] = std::cos(0.5 * phi);
    }
    // wave-killing rate (killwave.cu:8-103): sin^2(pi/2 d) / R^1.5, with the
    // single-precision sine of the reference (killwave.cu:11)
    if (P.kill_flag) {
        double d_in = pr.h(P.xmin) * P.kill_width, d_out = pr.h(P.xmax) * P.kill_width;
        double inner = P.xmin + d_in, outer = P.xmax - d_out;
        for (int ip = NPAD; ip < NPAD + nx; ++ip) {
            double rad = T(T_XC, ip), k = 0.0;
            if (rad < inner) {
                double s = (double)sinf((float)(0.5 * kPi * ((inner - rad) / d_in)));
                k = s * s / std::pow(rad, 1.5);
            }
            if (rad > outer) {
                double s = (double)sinf((float)(0.5 * kPi * ((rad - outer) / d_out)));
                k = s * s / std::pow(rad, 1.5);
            }
            T(T_KILL, ip) = k;
        }
    }
    // orbital advection background rotation (orbit.cu:51-72) + R, FP64
    rot.assign(2 * (size_t)nx, 0.0);
    for (int i = 0; i < nx; ++i) {
        double xc = T(T_XC, i + NPAD);
        rot[i] = P.OrbAdv_flag ? pr.v(xc) - xc * P.frame_omega : 0.0;
        rot[nx + i] = xc;
    }
    // initial state (fill_grid, init.cu:502-519) + seeded perturbation
    const size_t N = (size_t)nx * ny;
    gas.assign(3 * N, 0.0);
    dust.assign(P.dust_flag ? 3 * N : 0, 0.0);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t o = (size_t)j * nx + i;
            int ip = i + NPAD;
            double xc = T(T_XC, ip);
            double rho = T(T_G_R, ip), u = T(T_G_U, ip);
            if (P.noise_rho != 0.0) rho *= 1.0 + P.noise_rho * (2.0 * keyed_uniform(P.seed, 0, i, j) - 1.0);
            if (P.noise_vr != 0.0)
                u += P.noise_vr * std::sqrt(P.gam * pr.cs2(xc)) * (2.0 * keyed_uniform(P.seed, 1, i, j) - 1.0);
            gas[o] = rho;
            gas[N + o] = u;
            gas[2 * N + o] = 0.0;
            if (P.dust_flag) {
                dust[o] = T(T_D_R, ip);
                dust[N + o] = T(T_D_U, ip);
                dust[2 * N + o] = T(T_D_V, ip);
            }
        }
    // planet (init_planet, planet.cu:319-403; 2D softening planet.cu:308-317)
    h->planet_init = Body{};
    h->planet_init.a = P.planet_radius;
    h->planet_init.e = P.planet_ecc;
    h->planet_init.rs = P.rs_fac * P.sc_h * P.planet_radius;
}

// ------------------------------------------------------------ launching
const char* kname[] = {"prologue", "viscosity", "sources_kick", "sweep_x_gas", "ring_gas", "sweep_x_dust",
                       "ring_dust", "sources_predictor", "sources_corrector", "cfl"};
enum KId { K_PRO, K_VIS, K_KICK, K_SXG, K_RG, K_SXD, K_RD, K_PRED, K_CORR, K_CFL };

int timed_begin(pg_handle* h) {
    if (!h->profile) return PG_OK;
    size_t need = 2 * (h->ev_used + 1);
    while (h->ev_pool.size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        h->ev_pool.push_back(e);
    }
    HIP_TRY(hipEventRecord(h->ev_pool[2 * h->ev_used], h->stream));
    return PG_OK;
}
int timed_end(pg_handle* h, int k) {
    if (!h->profile) return PG_OK;
    HIP_TRY(hipEventRecord(h->ev_pool[2 * h->ev_used + 1], h->stream));
    if (h->ev_kernel.size() <= h->ev_used) h->ev_kernel.resize(h->ev_used + 1);
    h->ev_kernel[h->ev_used] = k;
    h->ev_used += 1;
    return PG_OK;
}
// After the stream has drained: fold the recorded pairs into the timers.
int timers_resolve(pg_handle* h) {
    if (!h->profile) return PG_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < h->ev_used; ++k) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, h->ev_pool[2 * k], h->ev_pool[2 * k + 1]));
        KernelTimer& t = h->timers[kname[h->ev_kernel[k]]];
        t.total_ms += ms;
        t.count += 1;
    }
    h->ev_used = 0;
    return PG_OK;
}

#define LAUNCH(h, k, expr)                                                                       \
    do {                                                                                         \
        int rc_ = timed_begin(h);                                                                \
        if (rc_) return rc_;                                                                     \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(PG_E_HIP, "launch %s failed: %s", kname[k], hipGetErrorString(e_)); \
        rc_ = timed_end(h, k);                                                                   \
        if (rc_) return rc_;                                                                     \
    } while (0)

// One step (DS order, solve.cu:212-292) after the prologue; dt, t and the
// planets are in the device control block.
template <typename real>
int step_passes(pg_handle* h) {
    const pg_params& P = h->P;
    hipStream_t s = h->stream;
    if (P.visc_flag) {
        DiskArgs<real> d = h->args<real>();
        LAUNCH(h, K_VIS, Launch<real>::viscosity(d, d.gas, s));
    }
    {
        DiskArgs<real> d = h->args<real>();
        LAUNCH(h, K_KICK, Launch<real>::sources(d, 0, 0, 0, s));
    }
    {
        DiskArgs<real> d = h->args<real>();
        LAUNCH(h, K_SXG, Launch<real>::sweep_x_gas(d, s));
        std::swap(h->gas, h->gasT);
    }
    {
        DiskArgs<real> d = h->args<real>();
        LAUNCH(h, K_RG, Launch<real>::ring_gas(d, s));
        std::swap(h->gas, h->gasT);
    }
    if (P.dust_flag) {
        DiskArgs<real> d = h->args<real>();
        LAUNCH(h, K_SXD, Launch<real>::sweep_x_dust(d, s));
        std::swap(h->dust, h->dustT);
        d = h->args<real>();
        LAUNCH(h, K_RD, Launch<real>::ring_dust(d, s));
        std::swap(h->dust, h->dustT);
    }
    if (P.visc_flag) {
        DiskArgs<real> d = h->args<real>();
        LAUNCH(h, K_PRED, Launch<real>::sources(d, 1, 0, 0, s));
        LAUNCH(h, K_VIS, Launch<real>::viscosity(d, d.gasT, s));
    }
    {
        DiskArgs<real> d = h->args<real>();
        LAUNCH(h, K_CORR, Launch<real>::sources(d, 2, P.kill_flag, 1, s));
    }
    h->rate_valid = true;
    return PG_OK;
}

int passes(pg_handle* h) { return h->prec == 8 ? step_passes<double>(h) : step_passes<float>(h); }

int prologue(pg_handle* h, int mode, double* dts) {
    hipError_t e = h->prec == 8 ? launch_prologue<double>(h->args<double>(), mode, dts, h->stream)
                                : launch_prologue<float>(h->args<float>(), mode, dts, h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch prologue failed: %s", hipGetErrorString(e));
    return PG_OK;
}

int push_ctl(pg_handle* h) {
    HIP_TRY(hipMemcpyAsync(h->ctl, h->ctl_host, sizeof(StepCtl), hipMemcpyHostToDevice, h->stream));
    return PG_OK;
}
int pull_ctl(pg_handle* h) {
    HIP_TRY(hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

int cfl_rate(pg_handle* h) {
    hipError_t e = h->prec == 8 ? Launch<double>::cfl(h->args<double>(), h->stream)
                                : Launch<float>::cfl(h->args<float>(), h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch cfl failed: %s", hipGetErrorString(e));
    e = h->prec == 8 ? Launch<double>::rate_reduce(h->args<double>(), h->stream)
                     : Launch<float>::rate_reduce(h->args<float>(), h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch rate_reduce failed: %s", hipGetErrorString(e));
    h->rate_valid = true;
    return PG_OK;
}

#define CHECK_H(h) \
    if (!(h)) return fail(PG_E_INVALID, "null handle")

}  // namespace

// =================================================================== C ABI
extern "C" {

const char* pg_last_error(void) { return g_err.c_str(); }
const char* pg_version(void) { return "penguin_amd 0.1 (gfx950)"; }

int pg_create(const pg_params* params, pg_handle** out) {
    g_err.clear();
    if (!params || !out) return fail(PG_E_INVALID, "null argument");
    *out = nullptr;
    int rc = validate(*params);
    if (rc) return rc;
    pg_handle* h = new pg_handle();
    h->P = *params;
    h->prec = params->precision;
    h->nx = params->xres;
    h->ny = params->yres;
    h->tile_x = (h->nx % 512 == 0) ? 512 : (h->nx == 384 ? 384 : (h->nx == 128 ? 128 : 256));
    if (const char* e = std::getenv("PG_TILE_X")) {
        int v = std::atoi(e);
        if (v == 128 || v == 256 || v == 384 || v == 512 || v == 1024) h->tile_x = v;
    }
    h->N = (size_t)h->nx * h->ny;
    auto cleanup = [&](int code) {
        pg_destroy(h);
        return code;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(PG_E_HIP, "hipStreamCreate failed"));
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess)
        return cleanup(fail(PG_E_HIP, "hipEventCreate failed"));
    const size_t es = h->prec;
    const size_t N = h->N;
    std::vector<double> gas, dust, rot;
    build_tables_and_state(h, gas, dust, rot);
    if ((rc = dev_alloc(h, &h->gas, 3 * N * es)) || (rc = dev_alloc(h, &h->gasT, 3 * N * es)) ||
        (rc = dev_alloc(h, &h->vis, 2 * N * es)) ||
        (rc = dev_alloc(h, &h->tab, h->tabh.size() * es)) || (rc = dev_alloc(h, &h->phitab, h->phih.size() * es)) ||
        (rc = dev_alloc(h, &h->res, h->nx * es)) ||
        (rc = dev_alloc(h, &h->inc, h->nx * sizeof(int))) || (rc = dev_alloc(h, &h->rot, rot.size() * sizeof(double))) ||
        (rc = dev_alloc(h, &h->ctl, sizeof(StepCtl))))
        return cleanup(rc);
    h->rate_nblk = ((h->nx + 255) / 256) * h->ny;
    if ((rc = dev_alloc(h, &h->rate_part, sizeof(double) * h->rate_nblk))) return cleanup(rc);
    if (h->P.dust_flag) {
        if ((rc = dev_alloc(h, &h->dust, 3 * N * es)) || (rc = dev_alloc(h, &h->dustT, 3 * N * es))) return cleanup(rc);
    }
    if (hipHostMalloc((void**)&h->ctl_host, sizeof(StepCtl), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(PG_E_NOMEM, "hipHostMalloc failed"));
    std::memset(h->ctl_host, 0, sizeof(StepCtl));
    if ((rc = upload(h, h->tabh, h->tab)) || (rc = upload(h, h->phih, h->phitab)) || (rc = upload(h, gas, h->gas)) ||
        (rc = upload(h, gas, h->gasT)))
        return cleanup(rc);
    if (h->P.dust_flag && ((rc = upload(h, dust, h->dust)) || (rc = upload(h, dust, h->dustT)))) return cleanup(rc);
    std::vector<double> zero(2 * N, 0.0);
    if ((rc = upload(h, zero, h->vis))) return cleanup(rc);
    if (hipMemcpy(h->rot, rot.data(), rot.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(PG_E_HIP, "rot upload failed"));
    if (hipMemset(h->inc, 0, h->nx * sizeof(int)) != hipSuccess) return cleanup(fail(PG_E_HIP, "memset failed"));
    StepCtl& c = *h->ctl_host;
    c.planet0 = h->planet_init;
    c.planet1 = h->planet_init;
    c.end_time = h->P.end_time;
    if ((rc = push_ctl(h))) return cleanup(rc);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup(fail(PG_E_HIP, "sync failed"));
    *out = h;
    return PG_OK;
}

int pg_destroy(pg_handle* h) {
    if (!h) return PG_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->graph) hipGraphExecDestroy(h->graph);
    void* bufs[] = {h->gas, h->gasT, h->dust, h->dustT, h->vis, h->tab, h->phitab, h->res, h->inc, h->rot, h->dts, h->ctl, h->rate_part};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (h->ctl_host) hipHostFree(h->ctl_host);
    for (hipEvent_t e : h->ev_pool) hipEventDestroy(e);
    if (h->ev0) hipEventDestroy(h->ev0);
    if (h->ev1) hipEventDestroy(h->ev1);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return PG_OK;
}

int pg_get_state(pg_handle* h, double* gas, double* dust) {
    CHECK_H(h);
    const size_t N = h->N;
    const int nx = h->nx;
    std::vector<double> g, dd;
    int rc;
    if (gas) {
        if ((rc = download(h, h->gas, g, 3 * N))) return rc;
        for (size_t o = 0; o < N; ++o) {
            int ip = (int)(o % nx) + NPAD;
            double rho = g[o];
            gas[o] = rho;
            gas[N + o] = tabv(h, T_CS2, ip) * rho;
            gas[2 * N + o] = g[N + o];
            gas[3 * N + o] = tabv(h, T_VBG, ip) + g[2 * N + o];
            gas[4 * N + o] = 0.0;
        }
    }
    if (dust) {
        if (!h->P.dust_flag) return fail(PG_E_STATE, "the disk has no dust");
        if ((rc = download(h, h->dust, dd, 3 * N))) return rc;
        for (size_t o = 0; o < N; ++o) {
            int ip = (int)(o % nx) + NPAD;
            dust[o] = dd[o];
            dust[N + o] = dd[N + o];
            dust[2 * N + o] = tabv(h, T_VBG, ip) + dd[2 * N + o];
            dust[3 * N + o] = 0.0;
        }
    }
    return PG_OK;
}

int pg_set_state(pg_handle* h, const double* gas, const double* dust) {
    CHECK_H(h);
    const size_t N = h->N;
    const int nx = h->nx;
    int rc;
    if (gas) {
        std::vector<double> g(3 * N);
        for (size_t o = 0; o < N; ++o) {
            int ip = (int)(o % nx) + NPAD;
            g[o] = gas[o];
            g[N + o] = gas[2 * N + o];
            g[2 * N + o] = gas[3 * N + o] - tabv(h, T_VBG, ip);
        }
        if ((rc = upload(h, g, h->gas))) return rc;
    }
    if (dust) {
        if (!h->P.dust_flag) return fail(PG_E_STATE, "the disk has no dust");
        std::vector<double> d(3 * N);
        for (size_t o = 0; o < N; ++o) {
            int ip = (int)(o % nx) + NPAD;
            d[o] = dust[o];
            d[N + o] = dust[N + o];
            d[2 * N + o] = dust[2 * N + o] - tabv(h, T_VBG, ip);
        }
        if ((rc = upload(h, d, h->dust))) return rc;
    }
    h->rate_valid = false;
    return PG_OK;
}

int pg_global_dt(pg_handle* h, double dt_prev, double* dt_out) {
    CHECK_H(h);
    if (!dt_out) return fail(PG_E_INVALID, "null dt_out");
    int rc;
    if ((rc = cfl_rate(h))) return rc;
    if ((rc = pull_ctl(h))) return rc;
    double rate = 0.0;
    unsigned long long bits = h->ctl_host->rate_bits;
    std::memcpy(&rate, &bits, sizeof(double));
    if (!(rate > 0.0) || !std::isfinite(rate)) return fail(PG_E_STATE, "CFL rate is %g (state not finite?)", rate);
    double dt = h->P.CFL / rate;
    double cap = dt_prev * 1.1;
    *dt_out = dt < cap ? dt : cap;
    return PG_OK;
}

int pg_solve(pg_handle* h, double t, double dt) {
    CHECK_H(h);
    if (!(dt > 0.0)) return fail(PG_E_INVALID, "dt must be positive");
    int rc;
    if ((rc = pull_ctl(h))) return rc;
    h->ctl_host->t = t;
    h->ctl_host->dt = dt;
    h->ctl_host->mode = 0;
    h->ctl_host->rate_bits = 0ull;
    if ((rc = push_ctl(h))) return rc;
    if ((rc = prologue(h, 0, nullptr))) return rc;
    if ((rc = passes(h))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return timers_resolve(h);
}

int pg_run(pg_handle* h, int nsteps, double* t_io, double* dt_io, double* dts, int use_graph) {
    CHECK_H(h);
    if (nsteps < 0 || !t_io || !dt_io) return fail(PG_E_INVALID, "bad arguments");
    int rc;
    if (nsteps == 0) return PG_OK;
    if (h->dts_cap < nsteps) {
        if (h->dts) hipFree(h->dts);
        h->dts = nullptr;
        if ((rc = dev_alloc(h, &h->dts, sizeof(double) * nsteps))) return rc;
        h->dts_cap = nsteps;
    }
    if ((rc = pull_ctl(h))) return rc;
    StepCtl& c = *h->ctl_host;
    c.t = *t_io;
    c.dt = *dt_io;
    c.dt_prev = *dt_io;
    c.nstep = 0;
    c.mode = 1;
    c.end_time = h->P.end_time;
    c.rate_bits = 0ull;
    if ((rc = push_ctl(h))) return rc;
    // rate of the current state for the first step (later steps get it from
    // the corrector)
    hipError_t e = h->prec == 8 ? Launch<double>::cfl(h->args<double>(), h->stream)
                                : Launch<float>::cfl(h->args<float>(), h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch cfl failed: %s", hipGetErrorString(e));
    if (use_graph && !h->profile) {
        if (h->graph && (h->graph_gas != h->gas || h->graph_dust != h->dust || h->graph_dts != h->dts)) {
            hipGraphExecDestroy(h->graph);
            h->graph = nullptr;
        }
        if (!h->graph) {
            h->graph_gas = h->gas;
            h->graph_dust = h->dust;
            h->graph_dts = h->dts;
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            int rc1 = prologue(h, 1, h->dts);
            int rc2 = rc1 ? rc1 : passes(h);
            hipError_t ec = hipStreamEndCapture(h->stream, &g);
            if (rc2) return rc2;
            if (ec != hipSuccess) return fail(PG_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
            ec = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (ec != hipSuccess) return fail(PG_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ec));
        }
        for (int n = 0; n < nsteps; ++n) HIP_TRY(hipGraphLaunch(h->graph, h->stream));
    } else {
        for (int n = 0; n < nsteps; ++n) {
            if ((rc = timed_begin(h))) return rc;
            if ((rc = prologue(h, 1, h->dts))) return rc;
            if ((rc = timed_end(h, K_PRO))) return rc;
            if ((rc = passes(h))) return rc;
        }
    }
    if ((rc = pull_ctl(h))) return rc;
    if ((rc = timers_resolve(h))) return rc;
    if (dts) HIP_TRY(hipMemcpy(dts, h->dts, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
    *t_io = c.t + c.dt;
    *dt_io = c.dt;
    return PG_OK;
}

int pg_viscosity_eval(pg_handle* h, int on_T, double* vis_out) {
    CHECK_H(h);
    hipError_t e = h->prec == 8
                       ? Launch<double>::viscosity(h->args<double>(), on_T ? (const double*)h->gasT : (const double*)h->gas, h->stream)
                       : Launch<float>::viscosity(h->args<float>(), on_T ? (const float*)h->gasT : (const float*)h->gas, h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch viscosity failed: %s", hipGetErrorString(e));
    if (vis_out) {
        std::vector<double> v;
        int rc = download(h, h->vis, v, 2 * h->N);
        if (rc) return rc;
        std::memcpy(vis_out, v.data(), 2 * h->N * sizeof(double));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

int pg_planet(pg_handle* h, double t, double* body10) {
    CHECK_H(h);
    int rc;
    if ((rc = pull_ctl(h))) return rc;
    h->ctl_host->t = t;
    h->ctl_host->dt = 0.0;
    if ((rc = push_ctl(h))) return rc;
    // mode 0 with dt = 0 would divide by zero in the orbital shift: evaluate
    // only the planet (host copy of the same closed form would do as well)
    DiskArgs<double> dd = h->args<double>();
    DiskArgs<float> df = h->args<float>();
    dd.orbadv = 0;
    df.orbadv = 0;
    hipError_t e = h->prec == 8 ? launch_prologue<double>(dd, 0, nullptr, h->stream)
                                : launch_prologue<float>(df, 0, nullptr, h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch prologue failed: %s", hipGetErrorString(e));
    if ((rc = pull_ctl(h))) return rc;
    if (body10) std::memcpy(body10, &h->ctl_host->planet0, sizeof(Body));
    return PG_OK;
}

int pg_sources(pg_handle* h, double t_planet, double mdt, double dt, int inplace) {
    CHECK_H(h);
    // the kernel takes its half step from the control block: dt_block = 2*dt
    int rc;
    if ((rc = pull_ctl(h))) return rc;
    h->ctl_host->t = t_planet;
    h->ctl_host->dt = 2.0 * dt;
    if ((rc = push_ctl(h))) return rc;
    DiskArgs<double> dd = h->args<double>();
    DiskArgs<float> df = h->args<float>();
    dd.orbadv = 0;
    df.orbadv = 0;
    hipError_t e = h->prec == 8 ? launch_prologue<double>(dd, 0, nullptr, h->stream)
                                : launch_prologue<float>(df, 0, nullptr, h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch prologue failed: %s", hipGetErrorString(e));
    // planet0 holds the planet at t_planet; modes 1/2 read planet1, so copy
    if ((rc = pull_ctl(h))) return rc;
    h->ctl_host->planet1 = h->ctl_host->planet0;
    h->ctl_host->pk1 = h->ctl_host->pk0;
    if ((rc = push_ctl(h))) return rc;
    int mode;
    if (mdt == 0.0) {
        if (!inplace) return fail(PG_E_INVALID, "mdt == 0 is only supported in place");
        mode = 0;
    } else {
        if (std::fabs(mdt - dt) > 1e-15 * std::fabs(dt))
            return fail(PG_E_INVALID, "mdt must be 0 or equal to dt (the two uses in DS)");
        mode = inplace ? 2 : 1;
    }
    e = h->prec == 8 ? Launch<double>::sources(h->args<double>(), mode, 0, 0, h->stream)
                     : Launch<float>::sources(h->args<float>(), mode, 0, 0, h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch sources failed: %s", hipGetErrorString(e));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

static int set_dt_and_prologue(pg_handle* h, double dt, int orbadv) {
    int rc;
    if ((rc = pull_ctl(h))) return rc;
    h->ctl_host->dt = dt;
    if ((rc = push_ctl(h))) return rc;
    DiskArgs<double> dd = h->args<double>();
    DiskArgs<float> df = h->args<float>();
    dd.orbadv = orbadv;
    df.orbadv = orbadv;
    hipError_t e = h->prec == 8 ? launch_prologue<double>(dd, 0, nullptr, h->stream)
                                : launch_prologue<float>(df, 0, nullptr, h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch prologue failed: %s", hipGetErrorString(e));
    return PG_OK;
}

int pg_sweep_x(pg_handle* h, double dt) {
    CHECK_H(h);
    int rc = set_dt_and_prologue(h, dt, 0);
    if (rc) return rc;
    hipError_t e = h->prec == 8 ? Launch<double>::sweep_x_gas(h->args<double>(), h->stream)
                                : Launch<float>::sweep_x_gas(h->args<float>(), h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch sweep_x_gas failed: %s", hipGetErrorString(e));
    std::swap(h->gas, h->gasT);
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

int pg_sweep_x_dust(pg_handle* h, double dt) {
    CHECK_H(h);
    if (!h->P.dust_flag) return fail(PG_E_STATE, "the disk has no dust");
    int rc = set_dt_and_prologue(h, dt, 0);
    if (rc) return rc;
    hipError_t e = h->prec == 8 ? Launch<double>::sweep_x_dust(h->args<double>(), h->stream)
                                : Launch<float>::sweep_x_dust(h->args<float>(), h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch sweep_x_dust failed: %s", hipGetErrorString(e));
    std::swap(h->dust, h->dustT);
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

// azimuthal sweep alone: the ring kernels with the orbital stages switched off
int pg_sweep_y(pg_handle* h, double dt) {
    CHECK_H(h);
    int rc = set_dt_and_prologue(h, dt, 0);
    if (rc) return rc;
    hipError_t e;
    if (h->prec == 8) {
        DiskArgs<double> d = h->args<double>();
        d.orbadv = 0;
        e = Launch<double>::ring_gas(d, h->stream);
    } else {
        DiskArgs<float> d = h->args<float>();
        d.orbadv = 0;
        e = Launch<float>::ring_gas(d, h->stream);
    }
    if (e != hipSuccess) return fail(PG_E_HIP, "launch ring_gas failed: %s", hipGetErrorString(e));
    std::swap(h->gas, h->gasT);
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

int pg_sweep_y_dust(pg_handle* h, double dt) {
    CHECK_H(h);
    if (!h->P.dust_flag) return fail(PG_E_STATE, "the disk has no dust");
    int rc = set_dt_and_prologue(h, dt, 0);
    if (rc) return rc;
    hipError_t e;
    if (h->prec == 8) {
        DiskArgs<double> d = h->args<double>();
        d.orbadv = 0;
        e = Launch<double>::ring_dust(d, h->stream);
    } else {
        DiskArgs<float> d = h->args<float>();
        d.orbadv = 0;
        e = Launch<float>::ring_dust(d, h->stream);
    }
    if (e != hipSuccess) return fail(PG_E_HIP, "launch ring_dust failed: %s", hipGetErrorString(e));
    std::swap(h->dust, h->dustT);
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

// set_OrbAdv + shift_OrbAdv + advecty(_dust) alone: the ring kernels with
// their sweep stage switched off (each cell is carried unchanged to its
// shifted position, then advected by the residual speed).
int pg_orbital_advection(pg_handle* h, double dt, int32_t* inc, double* res) {
    CHECK_H(h);
    if (!(dt > 0.0)) return fail(PG_E_INVALID, "dt must be positive");
    int rc = set_dt_and_prologue(h, dt, 1);
    if (rc) return rc;
    hipError_t e = hipSuccess;
    for (int fluid = 0; fluid < (h->P.dust_flag ? 2 : 1); ++fluid) {
        if (h->prec == 8) {
            DiskArgs<double> d = h->args<double>();
            d.sweep_on = 0;
            e = fluid ? Launch<double>::ring_dust(d, h->stream) : Launch<double>::ring_gas(d, h->stream);
        } else {
            DiskArgs<float> d = h->args<float>();
            d.sweep_on = 0;
            e = fluid ? Launch<float>::ring_dust(d, h->stream) : Launch<float>::ring_gas(d, h->stream);
        }
        if (e != hipSuccess) return fail(PG_E_HIP, "launch ring failed: %s", hipGetErrorString(e));
        if (fluid) std::swap(h->dust, h->dustT);
        else std::swap(h->gas, h->gasT);
    }
    if (inc) HIP_TRY(hipMemcpyAsync(inc, h->inc, h->nx * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    if (res) {
        std::vector<double> r;
        if ((rc = download(h, h->res, r, h->nx))) return rc;
        std::memcpy(res, r.data(), h->nx * sizeof(double));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

int pg_killwave(pg_handle* h, double dt) {
    CHECK_H(h);
    if (!h->P.kill_flag) return PG_OK;
    int rc = set_dt_and_prologue(h, dt, 0);
    if (rc) return rc;
    hipError_t e = h->prec == 8 ? Launch<double>::sources(h->args<double>(), 3, 1, 0, h->stream)
                                : Launch<float>::sources(h->args<float>(), 3, 1, 0, h->stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch killwave failed: %s", hipGetErrorString(e));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return PG_OK;
}

int pg_snapshot_label(const pg_params* P, char* buf, size_t len) {
    if (!P || !buf || len == 0) return fail(PG_E_INVALID, "null argument");
    // create_label (output.cu:14-93): each number passes through an int,
    // i.e. truncation toward zero
    auto I = [](double x) { return std::to_string((int)x); };
    const double earth = 0.000003, jupiter = 0.0009543;
    std::string s;
    if (P->ndim == 1) s = I(P->xres);
    else if (P->ndim == 2) s = I(P->xres) + "x" + I(P->yres);
    else s = I(P->xres) + "x" + I(P->yres) + "x" + I(P->zres);
    s += "_h" + I(1000.0 * P->sc_h);
    if (P->n_planet > 0) {
        if (P->planet_mass >= jupiter) s += "_" + I(P->n_planet) + "p" + I(P->planet_mass / jupiter) + "J";
        else s += "_" + I(P->n_planet) + "p" + I(P->planet_mass / earth) + "E";
        s += "_e" + I(P->planet_ecc * 1000.0);
    }
    if (P->visc_flag == 1) s += "_a" + I(std::log10(P->ss_alpha) * 10.0);
    if (P->OrbAdv_flag) s += "_OA";
    if (P->dust_flag) s += "_St" + I(std::log10(P->Stokes) * 10.0);
    static const char* rec[] = {"_VAN", "_MOC", "_PEM2", "_PEM3", "_PPM3", "_PEM4", "_PPM4"};
    if (P->recon_flag >= 0 && P->recon_flag <= 6) s += rec[P->recon_flag];
    if (s.size() + 1 > len) return fail(PG_E_INVALID, "label buffer too small (%zu needed)", s.size() + 1);
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return PG_OK;
}

int pg_write_snapshot(pg_handle* h, const char* filename, double t) {
    CHECK_H(h);
    if (!filename) return fail(PG_E_INVALID, "null filename");
    const size_t N = h->N;
    std::vector<double> gas(5 * N), dust(h->P.dust_flag ? 4 * N : 0);
    int rc = pg_get_state(h, gas.data(), h->P.dust_flag ? dust.data() : nullptr);
    if (rc) return rc;
    std::ofstream f(filename, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f) return fail(PG_E_INVALID, "cannot open %s for writing", filename);
    std::vector<double> xa(h->nx + 1), ya(h->ny + 1);
    for (int i = 0; i <= h->nx; ++i) xa[i] = tabv(h, T_XA, i + NPAD);
    for (int j = 0; j <= h->ny; ++j) ya[j] = h->P.ymin + (h->P.ymax - h->P.ymin) * (double)j / (double)h->ny;
    f.write((const char*)&t, sizeof(double));
    f.write((const char*)xa.data(), sizeof(double) * xa.size());
    f.write((const char*)ya.data(), sizeof(double) * ya.size());
    f.write((const char*)gas.data(), sizeof(double) * 4 * N);  // rho, p, u, v (ndim 2: no w)
    if (h->P.dust_flag) f.write((const char*)dust.data(), sizeof(double) * 4 * N);  // w_d too (output.cu:171)
    if (!f) return fail(PG_E_INVALID, "write to %s failed", filename);
    return PG_OK;
}

int pg_read_snapshot(pg_handle* h, const char* filename, double* t_out) {
    CHECK_H(h);
    if (!filename) return fail(PG_E_INVALID, "null filename");
    const size_t N = h->N;
    const size_t nfield = 4 + (h->P.dust_flag ? 4 : 0);
    const size_t expect = sizeof(double) * (1 + (h->nx + 1) + (h->ny + 1) + nfield * N);
    std::ifstream f(filename, std::ios::in | std::ios::binary);
    if (!f) return fail(PG_E_INVALID, "cannot open %s", filename);
    f.seekg(0, std::ios::end);
    size_t size = (size_t)f.tellg();
    if (size != expect)
        return fail(PG_E_INVALID, "%s has %zu bytes, a %dx%d%s snapshot has %zu", filename, size, h->nx, h->ny,
                    h->P.dust_flag ? " gas+dust" : "", expect);
    f.seekg(0);
    double t = 0.0;
    f.read((char*)&t, sizeof(double));
    std::vector<double> xa(h->nx + 1), ya(h->ny + 1);
    f.read((char*)xa.data(), sizeof(double) * xa.size());
    f.read((char*)ya.data(), sizeof(double) * ya.size());
    for (int i = 0; i <= h->nx; ++i)
        if (std::fabs(xa[i] - tabv(h, T_XA, i + NPAD)) > 1e-12 * std::fabs(xa[i]))
            return fail(PG_E_INVALID, "%s: radial grid does not match this disk", filename);
    std::vector<double> gas(5 * N, 0.0), dust(h->P.dust_flag ? 4 * N : 0);
    f.read((char*)gas.data(), sizeof(double) * 4 * N);
    if (h->P.dust_flag) f.read((char*)dust.data(), sizeof(double) * 4 * N);
    if (!f) return fail(PG_E_INVALID, "read from %s failed", filename);
    int rc = pg_set_state(h, gas.data(), h->P.dust_flag ? dust.data() : nullptr);
    if (rc) return rc;
    if (t_out) *t_out = t;
    return PG_OK;
}

int pg_stream(pg_handle* h, void** stream) {
    CHECK_H(h);
    if (!stream) return fail(PG_E_INVALID, "null argument");
    *stream = (void*)h->stream;
    return PG_OK;
}

int pg_profile_enable(pg_handle* h, int on) {
    CHECK_H(h);
    h->profile = on != 0;
    h->timers.clear();
    h->ev_used = 0;
    return PG_OK;
}

int pg_profile_get(pg_handle* h, const char* kernel, double* avg_us, int* count) {
    CHECK_H(h);
    if (!kernel || !avg_us || !count) return fail(PG_E_INVALID, "null argument");
    auto it = h->timers.find(kernel);
    if (it == h->timers.end()) {
        *avg_us = 0.0;
        *count = 0;
        return PG_OK;
    }
    *count = it->second.count;
    *avg_us = it->second.count ? 1000.0 * it->second.total_ms / it->second.count : 0.0;
    return PG_OK;
}

int pg_device_bytes(pg_handle* h, size_t* bytes) {
    CHECK_H(h);
    if (!bytes) return fail(PG_E_INVALID, "null argument");
    *bytes = h->bytes;
    return PG_OK;
}

}  // extern "C"
