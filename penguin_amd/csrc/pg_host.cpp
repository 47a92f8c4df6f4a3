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
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/penguin_amd.h"
#include "pg_device.h"
#include "pg_internal.h"
#include "pg_launch.h"
#include "pg_sph.h"

using namespace pg;

namespace {

thread_local std::string g_err;

// f(lo, hi) over [0, n) in contiguous chunks on up to 8 host threads (the
// conversions between the caller's FP64 fields and device staging run at
// host memory bandwidth, which one thread does not reach)
template <typename F>
void host_parallel(size_t n, F f) {
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nt = std::min<size_t>(std::max(1u, std::min(hw, 8u)), std::max<size_t>(1, n / 65536));
    if (nt <= 1) return f((size_t)0, n);
    std::vector<std::thread> th;
    const size_t chunk = (n + nt - 1) / nt;
    for (size_t k = 1; k < nt; ++k) {
        const size_t lo = std::min(n, k * chunk), hi = std::min(n, lo + chunk);
        th.emplace_back([=, &f] { f(lo, hi); });
    }
    f((size_t)0, std::min(n, chunk));
    for (auto& t : th) t.join();
}

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
    for (int i = 0; i < res + 2 * pad; ++i) {   // get_volume (geom.cu:3-20)
        double ra = a[i], dr = a[i + 1] - a[i];
        double dvdr = geom == 5   ? (dr < 1.0e-4 ? std::sin(ra + 0.5 * dr) : (std::cos(ra) - std::cos(ra + dr)) / dr)
                      : geom == 2 ? ra * ra + ra * dr + dr * dr / 3.0
                      : geom == 1 ? ra + dr / 2.0
                                  : 1.0;
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

// One disk: its device buffers and the host copies of the tables its state
// conversion needs.  Fluid buffers come in pairs; the owning group's parity
// says which of the two holds the current state.
struct DiskDev {
    pg_params P{};
    int prec = 4;
    int nx = 0, ny = 0;
    size_t N = 0;
    void* gas[2] = {nullptr, nullptr};
    void* dust[2] = {nullptr, nullptr};
    void* vis = nullptr;
    void* vis2 = nullptr;
    double* ave = nullptr;        // time average of the gas (3N FP64), while averaging
    double* diag = nullptr;       // per-tile partial sums of a diagnostic
    void* tab = nullptr;
    void* phitab = nullptr;
    void* res = nullptr;
    int* inc = nullptr;
    double* rot = nullptr;
    double* dts = nullptr;
    int dts_cap = 0;
    double* rate_part = nullptr;
    int rate_nblk = 0;
    StepCtl* ctl = nullptr;
    StepCtl* ctl_host = nullptr;  // pinned mirror
    size_t bytes = 0;
    std::vector<double> tabh;  // per-radius tables, field-major
    std::vector<double> phih;  // 4*ny azimuthal sine/cosine table
    std::vector<double> rho_bg;  // nx: the background density the stored gas density deviates from
    int tab_stride = 0;
    double dy = 0.0;
    double wly[3]{}, wry[3]{};
    Body planet_init{};

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
        for (int k = 0; k < 2; ++k) {
            d.gas[k] = (real*)gas[k];
            d.dust[k] = (real*)dust[k];
        }
        d.vis = (real*)vis;
        d.vis2 = (real*)vis2;
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
        d.dts = dts;
        d.dts_cap = dts_cap;
        d.ave = ave;
        return d;
    }
};

}  // namespace

// A set of disks launched together: a pg_handle holds one, a pg_batch many.
// Every kernel launch covers all disks of the set through a work list of
// (disk, local block) pairs; the per-disk descriptors live in device memory,
// so a captured step reads the current descriptors when it replays.
struct Group {
    int prec = 4;
    std::vector<DiskDev*> disks;
    hipStream_t stream = nullptr;
    bool dust = false, visc = false, kill = false;
    int gpar = 0, dpar = 0;      // which buffer of each fluid pair is current
    void* d_args = nullptr;      // DiskArgs<real>[ndisk]
    int2* d_wl = nullptr;        // work lists: x sweeps | rings | pointwise kernels
    size_t wl_cap = 0;
    int n_x = 0, n_ring = 0, n_pt = 0;
    int tile_x = 256, ring_b = RING_THREADS, order = 0;
    int pt_rows = PT_YMAX;       // rows of the pointwise/stencil tiles
    bool ave = false;            // time-average the gas around each device-loop step
    int ny_max = 0;
    // profiling: event pairs recorded around launches without a host wait,
    // resolved after the stream drains
    bool profile = false;
    std::map<std::string, KernelTimer> timers;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_kernel;   // kernel id of pair k (events 2k, 2k+1)
    size_t ev_used = 0;
    hipEvent_t ev_a = nullptr, ev_b = nullptr;
    // graphs of one device-loop step, one per start parity (a step may end
    // at another parity), and the parity each ends at
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    int graph_next[4] = {0, 0, 0, 0};
    size_t bytes = 0;
    // pinned host staging for results read back after a run (the step sizes)
    double* stage = nullptr;
    size_t stage_bytes = 0;

    int par() const { return gpar | (dpar << 1); }
    int ndisk() const { return (int)disks.size(); }
    template <typename real>
    LaunchSet<real> ls() const {
        LaunchSet<real> L{};
        L.disks = (const DiskArgs<real>*)d_args;
        L.ndisk = ndisk();
        L.wl_x = d_wl;
        L.n_x = n_x;
        L.tile_x = tile_x;
        L.wl_ring = d_wl + n_x;
        L.n_ring = n_ring;
        L.ring_b = ring_b;
        L.ny_max = ny_max;
        L.wl_pt = d_wl + n_x + n_ring;
        L.n_pt = n_pt;
        L.pt_rows = pt_rows;
        return L;
    }
    void set_par(int p) {
        gpar = p & 1;
        dpar = (p >> 1) & 1;
    }
    void drop_graph() {
        for (auto& x : graph) {
            if (x) (void)hipGraphExecDestroy(x);
            x = nullptr;
        }
    }
};

struct AdjWork;
struct pg_handle {
    Group g;
    AdjWork* adj = nullptr;   // adjoint workspace, allocated by the first adjoint call
    pg::Sph* sph = nullptr;   // ndim 3: the spherical disk (pg_sph.hip) instead of g
};
struct pg_batch {
    Group g;
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
    return PG_OK;
}

template <typename T>
int dev_alloc(size_t& counter, T** p, size_t bytes) {
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(PG_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    counter += bytes;
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

int upload(int prec, hipStream_t s, const std::vector<double>& src, void* dst) {
    return prec == 8 ? upload_cast<double>(src, dst, s) : upload_cast<float>(src, dst, s);
}

template <typename real>
int download_cast(hipStream_t s, const void* src, std::vector<double>& dst, size_t n) {
    std::vector<real> tmp(n);
    HIP_TRY(hipMemcpyAsync(tmp.data(), src, n * sizeof(real), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    dst.resize(n);
    for (size_t k = 0; k < n; ++k) dst[k] = (double)tmp[k];
    return PG_OK;
}

int download(int prec, hipStream_t s, const void* src, std::vector<double>& dst, size_t n) {
    return prec == 8 ? download_cast<double>(s, src, dst, n) : download_cast<float>(s, src, dst, n);
}

double tabv(const DiskDev* h, int f, int ip) { return h->tabh[(size_t)f * h->tab_stride + ip]; }
// Build the per-radius tables and the FP64 initial state (interior cells,
// device field order: gas rho,u,vp; dust rho,u,vp).
void build_tables_and_state(DiskDev* h, std::vector<double>& gas, std::vector<double>& dust,
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
    // the background the stored density deviates from: its differences
    // across each edge and the x sweep's pressure force on it, in FP64
    for (int ip = 1; ip < np; ++ip) {
        T(T_TBR, ip) = T(T_G_R, ip) - T(T_G_R, ip - 1);
        T(T_TBP, ip) = T(T_CS2, ip) * T(T_G_R, ip) - T(T_CS2, ip - 1) * T(T_G_R, ip - 1);
    }
    for (int ip = 1; ip < np - 1; ++ip)
        T(T_PSRC, ip) = -0.25 * (T(T_XA, ip) + T(T_XA, ip + 1)) * (T(T_TBP, ip) + T(T_TBP, ip + 1));
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
    // orbital advection background rotation (orbit.cu:51-72), R and the
    // background density, FP64
    rot.assign(3 * (size_t)nx, 0.0);
    for (int i = 0; i < nx; ++i) {
        const int ip = i + NPAD;
        double xc = T(T_XC, ip);
        rot[i] = P.OrbAdv_flag ? pr.v(xc) - xc * P.frame_omega : 0.0;
        rot[nx + i] = xc;
        rot[2 * nx + i] = T(T_G_R, ip);
    }
    h->rho_bg.assign(rot.begin() + 2 * nx, rot.begin() + 3 * nx);
    // initial state (fill_grid, init.cu:502-519) + seeded perturbation
    const size_t N = (size_t)nx * ny;
    gas.assign(3 * N, 0.0);
    dust.assign(P.dust_flag ? 3 * N : 0, 0.0);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t o = (size_t)j * nx + i;
            int ip = i + NPAD;
            double xc = T(T_XC, ip);
            // stored: rho - rho_b, for the perturbed rho_b (1 + eps); the
            // difference is exact in FP64 (the two are within a factor 2),
            // so rho_b + the stored value gives back rho_b (1 + eps) itself
            double drho = 0.0, u = T(T_G_U, ip);
            if (P.noise_rho != 0.0) {
                const double rb = T(T_G_R, ip);
                drho = rb * (1.0 + P.noise_rho * (2.0 * keyed_uniform(P.seed, 0, i, j) - 1.0)) - rb;
            }
            if (P.noise_vr != 0.0)
                u += P.noise_vr * std::sqrt(P.gam * pr.cs2(xc)) * (2.0 * keyed_uniform(P.seed, 1, i, j) - 1.0);
            gas[o] = drho;
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

// ------------------------------------------------------------ disks
void disk_free(DiskDev* d) {
    if (!d) return;
    void* bufs[] = {d->gas[0], d->gas[1], d->dust[0], d->dust[1], d->vis, d->vis2, d->tab, d->phitab,
                    d->res, d->inc, d->rot, d->dts, d->ctl, d->rate_part, d->ave, d->diag};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (d->ctl_host) (void)hipHostFree(d->ctl_host);
    delete d;
}

// Allocate a disk and upload its tables and initial state (init.cu:472-532,
// planet.cu:319-403, orbit.cu:111-120).
int disk_create(const pg_params& P, hipStream_t s, DiskDev** out) {
    *out = nullptr;
    DiskDev* h = new DiskDev();
    h->P = P;
    h->prec = P.precision;
    h->nx = P.xres;
    h->ny = P.yres;
    h->N = (size_t)h->nx * h->ny;
    auto cleanup = [&](int code) {
        disk_free(h);
        return code;
    };
    const size_t es = h->prec;
    const size_t N = h->N;
    std::vector<double> gas, dust, rot;
    build_tables_and_state(h, gas, dust, rot);
    int rc;
    size_t& B = h->bytes;
    if ((rc = dev_alloc(B, &h->gas[0], 3 * N * es)) || (rc = dev_alloc(B, &h->gas[1], 3 * N * es)) ||
        (rc = dev_alloc(B, &h->vis, 2 * N * es)) || (rc = dev_alloc(B, &h->vis2, 2 * N * es)) ||
        (rc = dev_alloc(B, &h->tab, h->tabh.size() * es)) || (rc = dev_alloc(B, &h->phitab, h->phih.size() * es)) ||
        (rc = dev_alloc(B, &h->res, h->nx * es)) ||
        (rc = dev_alloc(B, &h->inc, h->nx * sizeof(int))) || (rc = dev_alloc(B, &h->rot, rot.size() * sizeof(double))) ||
        (rc = dev_alloc(B, &h->ctl, sizeof(StepCtl))))
        return cleanup(rc);
    // one rate partial per tile of the pointwise kernels
    // (sized for the shortest tiles; the schedule sets the count in use)
    h->rate_nblk = ((h->nx + PT_X - 1) / PT_X) * ((h->ny + PT_TY - 1) / PT_TY);
    if ((rc = dev_alloc(B, &h->rate_part, sizeof(double) * h->rate_nblk))) return cleanup(rc);
    if (P.dust_flag) {
        if ((rc = dev_alloc(B, &h->dust[0], 3 * N * es)) || (rc = dev_alloc(B, &h->dust[1], 3 * N * es)))
            return cleanup(rc);
    }
    if (hipHostMalloc((void**)&h->ctl_host, sizeof(StepCtl), hipHostMallocDefault) != hipSuccess) {
        h->ctl_host = nullptr;
        return cleanup(fail(PG_E_NOMEM, "hipHostMalloc failed"));
    }
    std::memset(h->ctl_host, 0, sizeof(StepCtl));
    if ((rc = upload(h->prec, s, h->tabh, h->tab)) || (rc = upload(h->prec, s, h->phih, h->phitab)) ||
        (rc = upload(h->prec, s, gas, h->gas[0])) || (rc = upload(h->prec, s, gas, h->gas[1])))
        return cleanup(rc);
    if (P.dust_flag &&
        ((rc = upload(h->prec, s, dust, h->dust[0])) || (rc = upload(h->prec, s, dust, h->dust[1]))))
        return cleanup(rc);
    std::vector<double> zero(2 * N, 0.0);
    if ((rc = upload(h->prec, s, zero, h->vis)) || (rc = upload(h->prec, s, zero, h->vis2))) return cleanup(rc);
    if (hipMemcpy(h->rot, rot.data(), rot.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(PG_E_HIP, "rot upload failed"));
    if (hipMemset(h->inc, 0, h->nx * sizeof(int)) != hipSuccess) return cleanup(fail(PG_E_HIP, "memset failed"));
    StepCtl& c = *h->ctl_host;
    // the planet at t = 0 (init_planet, planet.cu:319-403), for diagnostics
    // taken before the first step; every step's prologue recomputes it
    Body b0 = h->planet_init;
    if (P.n_planet > 0) orbit_solution(b0, 0.0, P.ramp_time, P.planet_mass, P.frame_omega);
    c.planet0 = b0;
    c.planet1 = b0;
    c.pk0 = c.pk1 = planet_kernel_view(b0);
    c.end_time = P.end_time;
    if (hipMemcpyAsync(h->ctl, h->ctl_host, sizeof(StepCtl), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return cleanup(fail(PG_E_HIP, "control block upload failed"));
    *out = h;
    return PG_OK;
}

// ------------------------------------------------------------ groups
// Work lists of one launch schedule.
//   x sweeps: every (tile, row) of every disk, disk by disk; pointwise and
//   stencil kernels: every PT_X x pt_rows tile, tile rows in
//   XCD-contiguous bands.
//   azimuthal pass: tiles of TR radii x TY rows (tile = ty * ntr + tr).
//   Workgroups are dispatched round-robin over the 8 XCDs, each with its own
//   L2.  Tiles that are neighbours in azimuth share their 6-row halos and
//   tiles that are neighbours in radius share cache lines, so with order 0
//   each disk's tile rows are cut into 8 contiguous runs and XCD x receives
//   run x (entry 8k+x of the list), radius strips innermost.  Order 1 lists
//   the tiles row by row, order 2 strip by strip.  Entries with tile -1 pad
//   the list.
void build_worklists(const Group& g, int tile_x, int ring_b, int order, int pt_rows, std::vector<int2>& wx,
                     std::vector<int2>& wr, std::vector<int2>& wp) {
    wx.clear();
    wr.clear();
    wp.clear();
    const int nd = g.ndisk();
    // pointwise kernels (the viscosity stencil reads rows j-1, j, j+1): each
    // XCD receives a contiguous band of rows of every disk, so neighbouring
    // rows are read through one L2
    std::vector<int2> qp[8];
    for (int k = 0; k < nd; ++k) {
        const DiskDev* d = g.disks[k];
        const int tiles = (d->nx + tile_x - 1) / tile_x;
        for (int b = 0; b < tiles * d->ny; ++b) wx.push_back(make_int2(k, b));
        const int ntx = (d->nx + PT_X - 1) / PT_X, nty = (d->ny + pt_rows - 1) / pt_rows;
        const int per = (nty + 7) / 8;
        for (int x = 0; x < 8; ++x)
            for (int ty = x * per; ty < std::min((x + 1) * per, nty); ++ty)
                for (int tx = 0; tx < ntx; ++tx) qp[x].push_back(make_int2(k, tx + ntx * ty));
    }
    {
        size_t len = 0;
        for (int x = 0; x < 8; ++x) len = std::max(len, qp[x].size());
        wp.assign(8 * len, make_int2(0, -1));
        for (int x = 0; x < 8; ++x)
            for (size_t k = 0; k < qp[x].size(); ++k) wp[8 * k + x] = qp[x][k];
    }
    const int TR = g.prec == 8 ? Launch<double>::phi_tile_radii(ring_b) : Launch<float>::phi_tile_radii(ring_b);
    const int TY = g.prec == 8 ? Launch<double>::phi_tile_rows(ring_b) : Launch<float>::phi_tile_rows(ring_b);
    if (order != 0) {
        for (int k = 0; k < nd; ++k) {
            const int ntr = (g.disks[k]->nx + TR - 1) / TR, nty = (g.disks[k]->ny + TY - 1) / TY;
            if (order == 1) {
                for (int b = 0; b < ntr * nty; ++b) wr.push_back(make_int2(k, b));
            } else {
                for (int tr = 0; tr < ntr; ++tr)
                    for (int ty = 0; ty < nty; ++ty) wr.push_back(make_int2(k, ty * ntr + tr));
            }
        }
        return;
    }
    std::vector<int2> q[8];
    for (int k = 0; k < nd; ++k) {
        const int ntr = (g.disks[k]->nx + TR - 1) / TR, nty = (g.disks[k]->ny + TY - 1) / TY;
        const int per = (nty + 7) / 8;
        for (int x = 0; x < 8; ++x)
            for (int ty = x * per; ty < std::min((x + 1) * per, nty); ++ty)
                for (int tr = 0; tr < ntr; ++tr) q[x].push_back(make_int2(k, ty * ntr + tr));
    }
    size_t len = 0;
    for (int x = 0; x < 8; ++x) len = std::max(len, q[x].size());
    wr.assign(8 * len, make_int2(0, -1));
    for (int x = 0; x < 8; ++x)
        for (size_t k = 0; k < q[x].size(); ++k) wr[8 * k + x] = q[x][k];
}

bool tile_ok(int t) {
    return t == 122 || t == 128 || t == 250 || t == 256 || t == 384 || t == 506 || t == 512 || t == 1024;
}
bool ring_ok(int b) { return b == 256 || b == 512 || b == 1024; }

// Install a launch schedule: rebuild and upload the work lists.
int group_upload_args(Group& g);

// Rows of the pointwise/stencil tiles when the caller leaves the choice:
// tall tiles (a smaller stencil halo) when the set has enough of them to
// give every CU several workgroups, short ones otherwise.
int auto_pt_rows(const Group& g) {
    long tall = 0;
    for (const DiskDev* d : g.disks)
        tall += (long)((d->nx + PT_X - 1) / PT_X) * ((d->ny + PT_YMAX - 1) / PT_YMAX);
    return tall >= 1024 ? PT_YMAX : PT_TY;
}

int group_schedule(Group& g, int tile_x, int ring_b, int order, int pt_rows) {
    if (!tile_ok(tile_x))
        return fail(PG_E_INVALID, "tile_x=%d: use 122, 128, 250, 256, 384, 506, 512 or 1024", tile_x);
    if (!ring_ok(ring_b)) return fail(PG_E_INVALID, "ring_b=%d: use 256, 512 or 1024", ring_b);
    if (order < 0 || order > 2) return fail(PG_E_INVALID, "order=%d: use 0, 1 or 2", order);
    if (pt_rows == 0) pt_rows = auto_pt_rows(g);
    if (pt_rows != 4 && pt_rows != 8 && pt_rows != 16) return fail(PG_E_INVALID, "pt_rows=%d: use 4, 8 or 16", pt_rows);
    std::vector<int2> wx, wr, wp;
    build_worklists(g, tile_x, ring_b, order, pt_rows, wx, wr, wp);
    const size_t need = wx.size() + wr.size() + wp.size();
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (need > g.wl_cap) {
        if (g.d_wl) (void)hipFree(g.d_wl);
        g.d_wl = nullptr;
        g.wl_cap = 0;
        int rc = dev_alloc(g.bytes, &g.d_wl, need * sizeof(int2));
        if (rc) return rc;
        g.wl_cap = need;
    }
    std::vector<int2> all;
    all.reserve(need);
    all.insert(all.end(), wx.begin(), wx.end());
    all.insert(all.end(), wr.begin(), wr.end());
    all.insert(all.end(), wp.begin(), wp.end());
    HIP_TRY(hipMemcpy(g.d_wl, all.data(), need * sizeof(int2), hipMemcpyHostToDevice));
    g.n_x = (int)wx.size();
    g.n_ring = (int)wr.size();
    g.n_pt = (int)wp.size();
    g.tile_x = tile_x;
    g.ring_b = ring_b;
    g.order = order;
    g.pt_rows = pt_rows;
    // one CFL-rate partial per pointwise tile
    for (DiskDev* d : g.disks) d->rate_nblk = ((d->nx + PT_X - 1) / PT_X) * ((d->ny + pt_rows - 1) / pt_rows);
    g.drop_graph();
    return group_upload_args(g);
}

// Upload the per-disk descriptors (after creation or a reallocation).
int group_upload_args(Group& g) {
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (g.prec == 8) {
        std::vector<DiskArgs<double>> a;
        for (DiskDev* d : g.disks) a.push_back(d->args<double>());
        HIP_TRY(hipMemcpy(g.d_args, a.data(), a.size() * sizeof(a[0]), hipMemcpyHostToDevice));
    } else {
        std::vector<DiskArgs<float>> a;
        for (DiskDev* d : g.disks) a.push_back(d->args<float>());
        HIP_TRY(hipMemcpy(g.d_args, a.data(), a.size() * sizeof(a[0]), hipMemcpyHostToDevice));
    }
    return PG_OK;
}

void group_release(Group& g) {
    if (g.stream) (void)hipStreamSynchronize(g.stream);
    g.drop_graph();
    for (DiskDev* d : g.disks) disk_free(d);
    g.disks.clear();
    if (g.d_args) (void)hipFree(g.d_args);
    if (g.d_wl) (void)hipFree(g.d_wl);
    g.d_args = nullptr;
    g.d_wl = nullptr;
    for (hipEvent_t e : g.ev_pool) (void)hipEventDestroy(e);
    g.ev_pool.clear();
    if (g.stage) (void)hipHostFree(g.stage);
    g.stage = nullptr;
    g.stage_bytes = 0;
    if (g.ev_a) (void)hipEventDestroy(g.ev_a);
    if (g.ev_b) (void)hipEventDestroy(g.ev_b);
    if (g.stream) (void)hipStreamDestroy(g.stream);
    g.stream = nullptr;
}

int env_int(const char* name, int dflt) {
    const char* e = std::getenv(name);
    return e ? std::atoi(e) : dflt;
}

// Create the disks of a group.  All disks of a group share the storage
// precision and the dust / viscosity / wave-killing switches (each kernel
// launch then runs the same passes for every disk).
int group_init(Group& g, int n, const pg_params* params) {
    for (int k = 0; k < n; ++k) {
        int rc = validate(params[k]);
        if (rc) return n > 1 ? fail(rc, "disk %d: %s", k, g_err.c_str()) : rc;
        const pg_params& a = params[0];
        const pg_params& b = params[k];
        if (b.precision != a.precision || b.dust_flag != a.dust_flag || b.visc_flag != a.visc_flag ||
            b.kill_flag != a.kill_flag)
            return fail(PG_E_INVALID,
                        "disk %d: precision, dust_flag, visc_flag and kill_flag must agree across a batch", k);
    }
    g.prec = params[0].precision;
    g.dust = params[0].dust_flag;
    g.visc = params[0].visc_flag;
    g.kill = params[0].kill_flag;
    if (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess) {
        g.stream = nullptr;
        return fail(PG_E_HIP, "hipStreamCreate failed");
    }
    if (hipEventCreate(&g.ev_a) != hipSuccess || hipEventCreate(&g.ev_b) != hipSuccess)
        return fail(PG_E_HIP, "hipEventCreate failed");
    int rc;
    for (int k = 0; k < n; ++k) {
        DiskDev* d = nullptr;
        if ((rc = disk_create(params[k], g.stream, &d))) return n > 1 ? fail(rc, "disk %d: %s", k, g_err.c_str()) : rc;
        g.disks.push_back(d);
        g.bytes += d->bytes;
        g.ny_max = std::max(g.ny_max, d->ny);
    }
    const size_t asz = g.prec == 8 ? sizeof(DiskArgs<double>) : sizeof(DiskArgs<float>);
    if ((rc = dev_alloc(g.bytes, &g.d_args, asz * n))) return rc;
    if ((rc = group_upload_args(g))) return rc;
    // default schedule: widest x tile that divides every disk's radii
    int tile = 512;
    for (DiskDev* d : g.disks)
        if (d->nx % tile) tile = 256;
    if (n == 1) {
        // 122-cell tiles: their 128 slots fill two waves, every table value
        // a thread needs is loaded with the state, and a CU holds up to 16
        // such workgroups, so one tile's load latency hides behind the
        // others' arithmetic (the 1024x3072 disk: x sweep 78.5 us against
        // 89.3 with 506-cell tiles of eight waves, 80.7 with 250)
        tile = 122;
    }
    int ringb = RING_THREADS;
    tile = env_int("PG_TILE_X", tile);
    ringb = env_int("PG_RING_B", ringb);
    if (!tile_ok(tile)) tile = 256;
    if (!ring_ok(ringb)) ringb = RING_THREADS;
    return group_schedule(g, tile, ringb, 0, env_int("PG_PT_ROWS", 0));
}

// ------------------------------------------------------------ launching
const char* kname[] = {"prologue", "viscosity", "sources_kick", "sweep_x_gas", "ring_gas", "sweep_x_dust",
                       "ring_dust", "predictor_viscosity", "corrector_viscosity", "cfl", "sweep_x", "ring"};
enum KId { K_PRO, K_VIS, K_KICK, K_SXG, K_RG, K_SXD, K_RD, K_PRED, K_CORR, K_CFL, K_SX, K_RING };

int timed_begin(Group& g) {
    if (!g.profile) return PG_OK;
    size_t need = 2 * (g.ev_used + 1);
    while (g.ev_pool.size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        g.ev_pool.push_back(e);
    }
    HIP_TRY(hipEventRecord(g.ev_pool[2 * g.ev_used], g.stream));
    return PG_OK;
}
int timed_end(Group& g, int k) {
    if (!g.profile) return PG_OK;
    HIP_TRY(hipEventRecord(g.ev_pool[2 * g.ev_used + 1], g.stream));
    if (g.ev_kernel.size() <= g.ev_used) g.ev_kernel.resize(g.ev_used + 1);
    g.ev_kernel[g.ev_used] = k;
    g.ev_used += 1;
    return PG_OK;
}
// After the stream has drained: fold the recorded pairs into the timers.
int timers_resolve(Group& g) {
    if (!g.profile) return PG_OK;
    HIP_TRY(hipStreamSynchronize(g.stream));
    for (size_t k = 0; k < g.ev_used; ++k) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, g.ev_pool[2 * k], g.ev_pool[2 * k + 1]));
        KernelTimer& t = g.timers[kname[g.ev_kernel[k]]];
        t.total_ms += ms;
        t.count += 1;
    }
    g.ev_used = 0;
    return PG_OK;
}

#define LAUNCH(g, k, expr)                                                                       \
    do {                                                                                         \
        int rc_ = timed_begin(g);                                                                \
        if (rc_) return rc_;                                                                     \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(PG_E_HIP, "launch %s failed: %s", kname[k], hipGetErrorString(e_)); \
        rc_ = timed_end(g, k);                                                                   \
        if (rc_) return rc_;                                                                     \
    } while (0)

// One step (DS order, solve.cu:212-292) of every disk of the group after the
// prologue; each disk's dt, t and planets are in its device control block.
// With viscosity a step flips the gas pair three times (x sweep, azimuthal
// pass, corrector), so consecutive steps alternate between two parities.
// With viscosity, the forces of the step start (d.vis) come from the
// previous step's corrector kernel; a step whose start state did not come
// from one (vis_ready false) computes them first.
template <typename real>
int step_passes(Group& g, bool vis_ready) {
    const LaunchSet<real> L = g.ls<real>();
    hipStream_t s = g.stream;
    if (g.visc && !vis_ready) LAUNCH(g, K_VIS, Launch<real>::viscosity(L, g.par(), 0, s));
    // first half kick fused into the x sweeps; gas and dust sweep in one
    // launch (the dust's drag reads the gas before its sweep, the buffer the
    // gas sweep reads), and their azimuthal passes in one launch
    if (g.dust) {
        LAUNCH(g, K_SX, Launch<real>::sweep_x_both(L, g.par(), 1, s));
        g.gpar ^= 1;
        g.dpar ^= 1;
        LAUNCH(g, K_RING, Launch<real>::ring_both(L, g.par(), 3, s));
        g.gpar ^= 1;
        g.dpar ^= 1;
    } else {
        LAUNCH(g, K_SXG, Launch<real>::sweep_x_gas(L, g.par(), 1, s));
        g.gpar ^= 1;
        LAUNCH(g, K_RG, Launch<real>::ring_gas(L, g.par(), 3, s));
        g.gpar ^= 1;
    }
    if (g.visc) {
        // predictor + viscosity of the predicted state; corrector + wave
        // killing + CFL rate + viscosity of the corrected state
        LAUNCH(g, K_PRED, Launch<real>::predvis(L, g.par(), s));
        LAUNCH(g, K_CORR, Launch<real>::corrvis(L, g.par(), g.kill ? 1 : 0, s));
        g.gpar ^= 1;   // the corrected gas is in the other buffer
    } else {
        LAUNCH(g, K_CORR, Launch<real>::sources(L, g.par(), 2, g.kill ? 1 : 0, 1, s));
    }
    return PG_OK;
}

int passes(Group& g, bool vis_ready) {
    return g.prec == 8 ? step_passes<double>(g, vis_ready) : step_passes<float>(g, vis_ready);
}

// viscous forces of the current state into d.vis (start of a run)
int vis_start(Group& g) {
    if (!g.visc) return PG_OK;
    hipError_t e = g.prec == 8 ? Launch<double>::viscosity(g.ls<double>(), g.par(), 0, g.stream)
                               : Launch<float>::viscosity(g.ls<float>(), g.par(), 0, g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch viscosity failed: %s", hipGetErrorString(e));
    return PG_OK;
}

int prologue(Group& g, int mode, int do_orbadv) {
    hipError_t e = g.prec == 8 ? Launch<double>::prologue(g.ls<double>(), mode, do_orbadv, g.stream)
                               : Launch<float>::prologue(g.ls<float>(), mode, do_orbadv, g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch prologue failed: %s", hipGetErrorString(e));
    return PG_OK;
}

int push_ctl(Group& g) {
    for (DiskDev* d : g.disks)
        HIP_TRY(hipMemcpyAsync(d->ctl, d->ctl_host, sizeof(StepCtl), hipMemcpyHostToDevice, g.stream));
    return PG_OK;
}
int pull_ctl(Group& g) {
    for (DiskDev* d : g.disks)
        HIP_TRY(hipMemcpyAsync(d->ctl_host, d->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PG_OK;
}

// the time average's share of the current state (before and after each
// step, main.cu:344-350), when averaging is on
int average_pass(Group& g) {
    if (!g.ave) return PG_OK;
    hipError_t e = g.prec == 8 ? Launch<double>::average(g.ls<double>(), g.par(), g.stream)
                               : Launch<float>::average(g.ls<float>(), g.par(), g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch average failed: %s", hipGetErrorString(e));
    return PG_OK;
}

int launch_cfl(Group& g) {
    hipError_t e = g.prec == 8 ? Launch<double>::cfl(g.ls<double>(), g.par(), g.stream)
                               : Launch<float>::cfl(g.ls<float>(), g.par(), g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch cfl failed: %s", hipGetErrorString(e));
    return PG_OK;
}

int cfl_rate(Group& g) {
    int rc = launch_cfl(g);
    if (rc) return rc;
    hipError_t e = g.prec == 8 ? Launch<double>::rate_reduce(g.ls<double>(), g.stream)
                               : Launch<float>::rate_reduce(g.ls<float>(), g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch rate_reduce failed: %s", hipGetErrorString(e));
    return PG_OK;
}

// Pinned staging of at least `bytes` (grown on demand, after the stream has
// drained, so no copy in flight still targets the old block).
int group_stage(Group& g, size_t bytes) {
    if (g.stage_bytes >= bytes) return PG_OK;
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (g.stage) (void)hipHostFree(g.stage);
    g.stage = nullptr;
    g.stage_bytes = 0;
    if (hipHostMalloc((void**)&g.stage, bytes, hipHostMallocDefault) != hipSuccess) {
        g.stage = nullptr;
        return fail(PG_E_NOMEM, "hipHostMalloc(%zu) failed", bytes);
    }
    g.stage_bytes = bytes;
    return PG_OK;
}

// nsteps of the reference loop (main.cu:320-359) for every disk of the
// group, each with its own dt.  t_io/dt_io hold one entry per disk; dts
// (ndisk*nsteps, disk-major, may be null) receives the step sizes;
// loop_ms (may be null) the device time of the step loop.
int group_run(Group& g, int nsteps, double* t_io, double* dt_io, double* dts, int use_graph, float* loop_ms) {
    if (nsteps < 0 || !t_io || !dt_io) return fail(PG_E_INVALID, "bad arguments");
    if (nsteps == 0) return PG_OK;
    int rc;
    bool realloc = false;
    for (DiskDev* d : g.disks) {
        if (d->dts_cap >= nsteps) continue;
        HIP_TRY(hipStreamSynchronize(g.stream));
        if (d->dts) (void)hipFree(d->dts);
        d->dts = nullptr;
        d->dts_cap = 0;
        if ((rc = dev_alloc(g.bytes, &d->dts, sizeof(double) * nsteps))) return rc;
        d->dts_cap = nsteps;
        realloc = true;
    }
    if (realloc && (rc = group_upload_args(g))) return rc;
    if ((rc = pull_ctl(g))) return rc;
    for (int k = 0; k < g.ndisk(); ++k) {
        StepCtl& c = *g.disks[k]->ctl_host;
        c.t = t_io[k];
        c.dt = dt_io[k];
        c.dt_prev = dt_io[k];
        c.nstep = 0;
        c.mode = 1;
        c.end_time = g.disks[k]->P.end_time;
        c.rate_bits = 0ull;
        c.err = 0;
        c.err_step = 0;
    }
    if ((rc = push_ctl(g))) return rc;
    // rate of the current state for the first step (later steps get it from
    // the corrector), and its viscous forces
    if ((rc = launch_cfl(g))) return rc;
    if ((rc = vis_start(g))) return rc;
    if (loop_ms) HIP_TRY(hipEventRecord(g.ev_a, g.stream));
    if (use_graph && !g.profile) {
        for (int n = 0; n < nsteps; ++n) {
            const int p = g.par();
            if (!g.graph[p]) {
                // capture records the step without running it; the host's
                // parity bookkeeping advances as if it had run
                hipGraph_t gr;
                HIP_TRY(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
                int rc1 = prologue(g, 1, 1);
                if (!rc1) rc1 = average_pass(g);
                int rc2 = rc1 ? rc1 : passes(g, true);
                if (!rc2) rc2 = average_pass(g);
                hipError_t ec = hipStreamEndCapture(g.stream, &gr);
                if (rc2) return rc2;
                if (ec != hipSuccess) return fail(PG_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
                ec = hipGraphInstantiate(&g.graph[p], gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (ec != hipSuccess) {
                    g.graph[p] = nullptr;
                    return fail(PG_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ec));
                }
                g.graph_next[p] = g.par();
            } else {
                g.set_par(g.graph_next[p]);
            }
            HIP_TRY(hipGraphLaunch(g.graph[p], g.stream));
        }
    } else {
        for (int n = 0; n < nsteps; ++n) {
            if ((rc = timed_begin(g))) return rc;
            if ((rc = prologue(g, 1, 1))) return rc;
            if ((rc = timed_end(g, K_PRO))) return rc;
            if ((rc = average_pass(g))) return rc;
            if ((rc = passes(g, true))) return rc;
            if ((rc = average_pass(g))) return rc;
        }
    }
    if (loop_ms) HIP_TRY(hipEventRecord(g.ev_b, g.stream));
    if ((rc = pull_ctl(g))) return rc;
    if (loop_ms) HIP_TRY(hipEventElapsedTime(loop_ms, g.ev_a, g.ev_b));
    if ((rc = timers_resolve(g))) return rc;
    if (dts) {
        // the step sizes, staged through pinned memory on the group's stream
        const size_t n = (size_t)g.ndisk() * nsteps;
        if ((rc = group_stage(g, n * sizeof(double)))) return rc;
        for (int k = 0; k < g.ndisk(); ++k)
            HIP_TRY(hipMemcpyAsync(g.stage + (size_t)k * nsteps, g.disks[k]->dts, sizeof(double) * nsteps,
                                   hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        std::memcpy(dts, g.stage, n * sizeof(double));
    }
    int bad = -1;
    for (int k = 0; k < g.ndisk(); ++k) {
        const DiskDev* d = g.disks[k];
        t_io[k] = d->ctl_host->t + d->ctl_host->dt;
        dt_io[k] = d->ctl_host->dt;
        if (d->ctl_host->err && bad < 0) bad = k;
    }
    if (bad >= 0) {
        const StepCtl& c = *g.disks[bad]->ctl_host;
        double rate = 0.0;
        std::memcpy(&rate, &c.rate_bits, sizeof(double));
        return fail(PG_E_STATE, "disk %d: the CFL rate before step %d is %g (the state is not finite); "
                    "the loop stopped there at t = %.17g", bad, c.err_step, rate, c.t);
    }
    return PG_OK;
}

// ------------------------------------------------------------ state I/O
int disk_get_state(Group& g, int k, double* gas, double* dust) {
    DiskDev* h = g.disks[k];
    const size_t N = h->N;
    const int nx = h->nx;
    std::vector<double> gv, dv;
    int rc;
    if (gas) {
        if ((rc = download(h->prec, g.stream, h->gas[g.gpar], gv, 3 * N))) return rc;
        for (size_t o = 0; o < N; ++o) {
            int ip = (int)(o % nx) + NPAD;
            double rho = h->rho_bg[o % nx] + gv[o];
            gas[o] = rho;
            gas[N + o] = tabv(h, T_CS2, ip) * rho;
            gas[2 * N + o] = gv[N + o];
            gas[3 * N + o] = tabv(h, T_VBG, ip) + gv[2 * N + o];
            gas[4 * N + o] = 0.0;
        }
    }
    if (dust) {
        if (!h->P.dust_flag) return fail(PG_E_STATE, "the disk has no dust");
        if ((rc = download(h->prec, g.stream, h->dust[g.dpar], dv, 3 * N))) return rc;
        for (size_t o = 0; o < N; ++o) {
            int ip = (int)(o % nx) + NPAD;
            dust[o] = dv[o];
            dust[N + o] = dv[N + o];
            dust[2 * N + o] = tabv(h, T_VBG, ip) + dv[2 * N + o];
            dust[3 * N + o] = 0.0;
        }
    }
    return PG_OK;
}

int disk_set_state(Group& g, int k, const double* gas, const double* dust) {
    DiskDev* h = g.disks[k];
    const size_t N = h->N;
    const int nx = h->nx;
    int rc;
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (dust && !h->P.dust_flag) return fail(PG_E_STATE, "the disk has no dust");
    if (gas) {
        std::vector<double> v(3 * N);
        for (size_t o = 0; o < N; ++o) {
            int ip = (int)(o % nx) + NPAD;
            v[o] = gas[o] - h->rho_bg[o % nx];
            v[N + o] = gas[2 * N + o];
            v[2 * N + o] = gas[3 * N + o] - tabv(h, T_VBG, ip);
        }
        if ((rc = upload(h->prec, g.stream, v, h->gas[g.gpar]))) return rc;
    }
    if (dust) {
        std::vector<double> v(3 * N);
        for (size_t o = 0; o < N; ++o) {
            int ip = (int)(o % nx) + NPAD;
            v[o] = dust[o];
            v[N + o] = dust[N + o];
            v[2 * N + o] = dust[2 * N + o] - tabv(h, T_VBG, ip);
        }
        if ((rc = upload(h->prec, g.stream, v, h->dust[g.dpar]))) return rc;
    }
    return PG_OK;
}

int group_profile_get(Group& g, const char* kernel, double* avg_us, int* count) {
    if (!kernel || !avg_us || !count) return fail(PG_E_INVALID, "null argument");
    auto it = g.timers.find(kernel);
    if (it == g.timers.end()) {
        *avg_us = 0.0;
        *count = 0;
        return PG_OK;
    }
    *count = it->second.count;
    *avg_us = it->second.count ? 1000.0 * it->second.total_ms / it->second.count : 0.0;
    return PG_OK;
}

// Time `steps` captured steps of every schedule candidate on the batch's own
// state and keep the fastest; the state, times and control blocks are
// restored after each trial.
int group_autotune(Group& g, int steps, double* best_ms) {
    if (steps < 1) return fail(PG_E_INVALID, "steps must be >= 1");
    int rc;
    const int nd = g.ndisk();
    // save the current state and control blocks
    std::vector<void*> sg(nd, nullptr), sd(nd, nullptr);
    auto free_saved = [&]() {
        for (void* p : sg)
            if (p) (void)hipFree(p);
        for (void* p : sd)
            if (p) (void)hipFree(p);
    };
    size_t scratch = 0;
    std::vector<StepCtl> saved_ctl(nd);
    if ((rc = pull_ctl(g))) return rc;
    for (int k = 0; k < nd; ++k) {
        DiskDev* d = g.disks[k];
        saved_ctl[k] = *d->ctl_host;
        const size_t bytes = 3 * d->N * d->prec;
        if ((rc = dev_alloc(scratch, &sg[k], bytes))) return free_saved(), rc;
        if (hipMemcpyAsync(sg[k], d->gas[g.gpar], bytes, hipMemcpyDeviceToDevice, g.stream) != hipSuccess)
            return free_saved(), fail(PG_E_HIP, "state copy failed");
        if (g.dust) {
            if ((rc = dev_alloc(scratch, &sd[k], bytes))) return free_saved(), rc;
            if (hipMemcpyAsync(sd[k], d->dust[g.dpar], bytes, hipMemcpyDeviceToDevice, g.stream) != hipSuccess)
                return free_saved(), fail(PG_E_HIP, "state copy failed");
        }
    }
    auto restore = [&]() -> int {
        for (int k = 0; k < nd; ++k) {
            DiskDev* d = g.disks[k];
            const size_t bytes = 3 * d->N * d->prec;
            HIP_TRY(hipMemcpyAsync(d->gas[g.gpar], sg[k], bytes, hipMemcpyDeviceToDevice, g.stream));
            if (g.dust) HIP_TRY(hipMemcpyAsync(d->dust[g.dpar], sd[k], bytes, hipMemcpyDeviceToDevice, g.stream));
            *d->ctl_host = saved_ctl[k];
        }
        int r = push_ctl(g);
        if (r) return r;
        HIP_TRY(hipStreamSynchronize(g.stream));
        return PG_OK;
    };
    const bool prof = g.profile;
    g.profile = false;
    const int tiles[] = {122, 128, 250, 256, 512};
    const int rings[] = {256, 512, 1024};
    const int pts[] = {PT_TY, PT_YMAX};
    std::vector<double> t(nd), dt(nd);
    // device time per step of `n` steps under one schedule (from the saved
    // state, which is restored afterwards); a negative return is an error code
    auto time_schedule = [&](int tx, int rb, int order, int pr, int n, double* per) -> int {
        int r = group_schedule(g, tx, rb, order, pr);
        if (r) return r;
        float ms = 0.f;
        for (int k = 0; k < nd; ++k) {
            t[k] = saved_ctl[k].t;
            dt[k] = saved_ctl[k].dt > 0.0 ? saved_ctl[k].dt : g.disks[k]->P.end_time;
        }
        // first call captures the graph and warms up; the second is timed
        if ((r = group_run(g, 1, t.data(), dt.data(), nullptr, 1, nullptr)) ||
            (r = group_run(g, n, t.data(), dt.data(), nullptr, 1, &ms)) || (r = restore()))
            return r;
        *per = ms / n;
        return PG_OK;
    };
    struct Cand {
        double per;
        int tx, rb, order, pr;
    };
    std::vector<Cand> cands;
    for (int tx : tiles)
        for (int rb : rings)
            for (int order = 0; order < 3; ++order)
                for (int pr : pts) {
                    double per = 0.0;
                    if ((rc = time_schedule(tx, rb, order, pr, steps, &per))) {
                        if (rc == PG_E_INVALID) continue;  // does not fit this batch
                        g.profile = prof;
                        return free_saved(), rc;
                    }
                    cands.push_back({per, tx, rb, order, pr});
                }
    // the fastest few again, over four times the steps: one short timing
    // can rank a schedule by a few percent of noise
    std::sort(cands.begin(), cands.end(), [](const Cand& x, const Cand& y) { return x.per < y.per; });
    double best = 1e300;
    int bt = g.tile_x, bb = g.ring_b, bo = g.order, bp = g.pt_rows;
    const size_t nfinal = std::min<size_t>(cands.size(), 4);
    for (size_t q = 0; q < nfinal; ++q) {
        const Cand& c = cands[q];
        double per = 0.0;
        if ((rc = time_schedule(c.tx, c.rb, c.order, c.pr, 4 * steps, &per))) {
            g.profile = prof;
            return free_saved(), rc;
        }
        if (per < best) {
            best = per;
            bt = c.tx;
            bb = c.rb;
            bo = c.order;
            bp = c.pr;
        }
    }
    g.profile = prof;
    free_saved();
    g_err.clear();
    if ((rc = group_schedule(g, bt, bb, bo, bp))) return rc;
    if (best_ms) *best_ms = best;
    return PG_OK;
}

// ------------------------------------------------------------ adjoint
// Per step the backward sweep needs every intermediate of the forward step
// (the inputs of each pass).  The forward run keeps a checkpoint every
// `ckpt` steps and the step sizes; the backward sweep, segment by segment
// from the last, recomputes the segment once from its checkpoint, storing
// each step's intermediates in a slot, then runs the adjoint passes of its
// steps in reverse order.  With ckpt 0 (the tape) every step has a slot of
// its own, written by the forward run itself, and the backward sweep runs the
// adjoint passes alone: no recompute, at SLOT_WORDS reals per cell and step
// of device memory (a 512 x 1024 FP32 disk: 46 MB per step).
constexpr int SLOT_WORDS = 22;   // S0, S1, S2, SA, S3, T (3 fields each), V0, V1 (2 each)
enum { SL_S0 = 0, SL_S1 = 3, SL_S2 = 6, SL_SA = 9, SL_S3 = 12, SL_T = 15, SL_V0 = 18, SL_V1 = 20 };
// forward passes of one slot (the fused kernels of the forward step, which
// also store the intermediates the adjoint needs), each with its own disk
// descriptor naming the buffers it reads and writes
enum { AA_V0, AA_X, AA_P, AA_PRED, AA_CORR, AA_N };

}  // namespace

struct AdjWork {
    // slots: sps slots per set, nset sets.  A checkpointed sweep of several
    // segments keeps two sets: the recompute of segment k-1 fills one on
    // its own stream while the adjoint steps of segment k read the other.
    int nsteps = 0, ckpt = 0, nseg = 0, slots = 0, sps = 0, nset = 1, prec = 4;
    size_t N = 0;
    std::vector<double> ts, dts;   // start time and size of every forward step
    void* ckpts = nullptr;         // nseg states (3N)
    void* slot = nullptr;          // slots x SLOT_WORDS*N
    void* lam = nullptr;           // lamA, lamB (3N each), lamV0 (2N)
    double* lparam = nullptr;      // per-workgroup partial sums of dL/dq, dL/dalpha
    int nslot = 0;                 // slots per parameter
    double* td = nullptr;          // (t, dt) of every step, on the device
    double* td_seg = nullptr;      // (t, dt) of the segment held in each set
    void* args = nullptr;          // slots x AA_N disk descriptors
    void* stage = nullptr;         // pinned host staging of one adjoint state (3N reals)
    // per-slot control block and orbital shift (the step's dt, planets, inc,
    // res), written by the slot's prologue replay in the recompute and read
    // by the recompute and the adjoint passes of that step
    StepCtl* sctl = nullptr;
    int* sinc = nullptr;
    void* sres = nullptr;
    bool forward_done = false;
    size_t bytes = 0;
    // two sets: the recompute's stream (default priority, as every stream
    // here) and the events that order the sets' reuse between the streams
    hipStream_t rs = nullptr;
    hipEvent_t ev_r[2] = {nullptr, nullptr};   // recompute into set k done
    hipEvent_t ev_a[2] = {nullptr, nullptr};   // adjoint steps reading set k done
    // captured graphs of one segment, by kind (0: recompute into a set,
    // 1: adjoint steps reading it), set, length and, for the adjoint steps,
    // the buffer holding the incoming adjoint
    struct SegGraph {
        int kind, set, len;
        void* lam_in;
        void* lam_out;
        hipGraphExec_t exec;
    };
    std::vector<SegGraph> graphs;
};

namespace {

// workgroups of the largest adjoint grid (x adjoint: 256-radius tiles x
// rows; source adjoints: 256-radius segments x rows; azimuthal adjoints:
// 8-radius x 16-row tiles at most; viscosity: 64 x 8 tiles at most)
int adj_slots(int nx, int ny) {
    const int a = ((nx + ADJ_TX - 1) / ADJ_TX) * ny;   // x-sweep adjoint tiles
    const int b = ((nx + 2 + 255) / 256) * ny;
    const int c = ((nx + 7) / 8) * ((ny + 15) / 16);
    const int v = ((nx + 63) / 64) * ((ny + 7) / 8);   // viscosity tiles (the smaller, FP64)
    return std::max(std::max(a, v), std::max(b, c));
}

void adj_free(AdjWork* w) {
    if (!w) return;
    if (w->rs) (void)hipStreamSynchronize(w->rs);
    for (auto& sg : w->graphs)
        if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
    void* bufs[] = {w->ckpts, w->slot, w->lam, w->lparam, w->td, w->td_seg, w->args, w->sctl, w->sinc, w->sres};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (w->stage) (void)hipHostFree(w->stage);
    for (int k = 0; k < 2; ++k) {
        if (w->ev_r[k]) (void)hipEventDestroy(w->ev_r[k]);
        if (w->ev_a[k]) (void)hipEventDestroy(w->ev_a[k]);
    }
    if (w->rs) (void)hipStreamDestroy(w->rs);
    delete w;
}

template <typename real>
int adj_alloc(pg_handle* h, int nsteps, int ckpt) {
    Group& g = h->g;
    DiskDev* d = g.disks[0];
    if (h->adj && h->adj->nsteps == nsteps && h->adj->ckpt == ckpt) return PG_OK;
    adj_free(h->adj);
    h->adj = nullptr;
    AdjWork* w = new AdjWork();
    w->nsteps = nsteps;
    w->ckpt = ckpt;
    w->nseg = ckpt > 0 ? (nsteps + ckpt - 1) / ckpt : 0;
    w->sps = (ckpt > 0 ? std::min(ckpt, nsteps) : nsteps) + 1;
    w->nset = ckpt > 0 && w->nseg > 1 ? 2 : 1;
    w->slots = w->sps * w->nset;
    w->prec = g.prec;
    w->N = d->N;
    w->nslot = adj_slots(d->nx, d->ny);
    const size_t N = w->N, es = sizeof(real);
    int rc;
    auto cleanup = [&](int code) {
        adj_free(w);
        return code;
    };
    if ((w->nseg > 0 && (rc = dev_alloc(w->bytes, &w->ckpts, (size_t)w->nseg * 3 * N * es))) ||
        (rc = dev_alloc(w->bytes, &w->slot, (size_t)w->slots * SLOT_WORDS * N * es)) ||
        (rc = dev_alloc(w->bytes, &w->lam, (size_t)8 * N * es)) ||
        (rc = dev_alloc(w->bytes, &w->lparam, 2 * (size_t)adj_slots(d->nx, d->ny) * sizeof(double))) ||
        (rc = dev_alloc(w->bytes, &w->td, 2 * (size_t)nsteps * sizeof(double))) ||
        (rc = dev_alloc(w->bytes, &w->td_seg, 2 * (size_t)w->nset * (ckpt > 0 ? ckpt : 1) * sizeof(double))) ||
        (rc = dev_alloc(w->bytes, &w->args, (size_t)w->slots * AA_N * sizeof(DiskArgs<real>))) ||
        (rc = dev_alloc(w->bytes, &w->sctl, (size_t)w->slots * sizeof(StepCtl))) ||
        (rc = dev_alloc(w->bytes, &w->sinc, (size_t)w->slots * d->nx * sizeof(int))) ||
        (rc = dev_alloc(w->bytes, &w->sres, (size_t)w->slots * d->nx * es)))
        return cleanup(rc);
    if (hipHostMalloc(&w->stage, 3 * N * es, hipHostMallocDefault) != hipSuccess) {
        w->stage = nullptr;
        return cleanup(fail(PG_E_NOMEM, "pinned staging of %zu bytes", 3 * N * es));
    }
    if (w->nset == 2) {
        if (hipStreamCreateWithFlags(&w->rs, hipStreamNonBlocking) != hipSuccess) {
            w->rs = nullptr;
            return cleanup(fail(PG_E_HIP, "hipStreamCreate failed"));
        }
        for (int k = 0; k < 2; ++k)
            if (hipEventCreateWithFlags(&w->ev_r[k], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&w->ev_a[k], hipEventDisableTiming) != hipSuccess)
                return cleanup(fail(PG_E_HIP, "hipEventCreate failed"));
    }
    // descriptors of the recompute passes: the disk's own tables, with the
    // state, viscosity, control-block and shift pointers of the slot
    std::vector<DiskArgs<real>> a((size_t)w->slots * AA_N);
    real* base = (real*)w->slot;
    for (int m = 0; m < w->slots; ++m) {
        real* sl = base + (size_t)m * SLOT_WORDS * N;
        // the next slot of the same set (a set's last slot only receives)
        const int first = m / w->sps * w->sps;
        real* nx = base + (size_t)(first + std::min(m - first + 1, w->sps - 1)) * SLOT_WORDS * N;
        const DiskArgs<real> t = d->args<real>();
        // parity 0 throughout: gas[0] is a pass's input, gas[1] its output
        auto set = [&](int k, real* g0, real* g1, real* vis, real* vis2) -> DiskArgs<real>& {
            DiskArgs<real> x = t;
            x.gas[0] = g0;
            x.gas[1] = g1;
            x.dust[0] = x.dust[1] = nullptr;
            x.vis = vis;
            x.vis2 = vis2;
            x.dts = nullptr;
            x.dts_cap = 0;
            x.ctl = w->sctl + m;
            x.inc = w->sinc + (size_t)m * d->nx;
            x.res = (real*)w->sres + (size_t)m * d->nx;
            return a[(size_t)m * AA_N + k] = x;
        };
        set(AA_V0, sl + SL_S0 * N, sl + SL_S0 * N, sl + SL_V0 * N, nullptr);
        set(AA_X, sl + SL_S0 * N, sl + SL_S2 * N, sl + SL_V0 * N, nullptr).out_kick = sl + SL_S1 * N;
        set(AA_P, sl + SL_S2 * N, sl + SL_S3 * N, nullptr, nullptr).out_swept = sl + SL_SA * N;
        set(AA_PRED, sl + SL_S3 * N, nullptr, sl + SL_V0 * N, sl + SL_V1 * N).out_pred = sl + SL_T * N;
        if (d->P.visc_flag)   // corrector + viscosity of the result: next slot's S0 and V0
            set(AA_CORR, sl + SL_S3 * N, nx + SL_S0 * N, nx + SL_V0 * N, sl + SL_V1 * N);
        else                  // corrector in place on a copy of S3
            set(AA_CORR, nx + SL_S0 * N, nx + SL_S0 * N, nullptr, nullptr);
    }
    if (hipMemcpy(w->args, a.data(), a.size() * sizeof(a[0]), hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(PG_E_HIP, "adjoint descriptor upload failed"));
    h->adj = w;
    return PG_OK;
}

// prologues of the len steps of the segment held in `set`, into its slots'
// control blocks and shifts: t, dt from the set's device copy of the
// segment's step times, then the planets and the orbital shift as in the
// forward step; one workgroup per step (they depend on the times only)
template <typename real>
int set_steps(Group& g, const AdjWork* w, int set, int len, hipStream_t s) {
    LaunchSet<real> L = g.ls<real>();
    L.disks = (const DiskArgs<real>*)w->args + (size_t)set * w->sps * AA_N + AA_X;
    L.ndisk = 1;
    hipError_t e = Launch<real>::prologue_replay_many(L, AA_N, len, w->td_seg + 2 * (size_t)set * w->ckpt, s);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch prologue failed: %s", hipGetErrorString(e));
    return PG_OK;
}

#define HIPL(expr)                                                                                   \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(PG_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));         \
    } while (0)

// Recompute a step into slot m (and its end state into the next slot's
// S0); first: the segment's first step (its start state's viscous forces
// are computed here too).  The slot's prologue has run (set_steps).
template <typename real>
int adj_recompute(pg_handle* h, int m, bool first, bool write_next, hipStream_t s) {
    Group& g = h->g;
    AdjWork* w = h->adj;
    const size_t N = w->N;
    LaunchSet<real> L = g.ls<real>();
    const DiskArgs<real>* a = (const DiskArgs<real>*)w->args + (size_t)m * AA_N;
    auto at = [&](int k) {
        LaunchSet<real> x = L;
        x.disks = a + k;
        x.ndisk = 1;
        return x;
    };
    real* sl = (real*)w->slot + (size_t)m * SLOT_WORDS * N;
    real* nxt = (real*)w->slot + (size_t)(m + 1) * SLOT_WORDS * N;
    const bool visc = g.visc;
    // the forward step's fused kernels (as group_run launches them), with
    // every intermediate landing in the slot: the viscous forces of a
    // segment's first state, then kick + x sweep, azimuthal sweep + shift +
    // residual advection, predictor + its viscous forces, corrector + the
    // viscous forces of the result (the next slot's S0 and V0)
    if (visc && first) HIPL(Launch<real>::viscosity(at(AA_V0), 0, 0, s));
    HIPL(Launch<real>::sweep_x_gas(at(AA_X), 0, 1, s));
    HIPL(Launch<real>::ring_gas(at(AA_P), 0, 3, s));
    if (visc) HIPL(Launch<real>::predvis(at(AA_PRED), 0, s));
    if (write_next) {
        if (visc) {
            HIPL(Launch<real>::corrvis(at(AA_CORR), 0, 0, s));
        } else {
            HIPL(hipMemcpyAsync(nxt + SL_S0 * N, sl + SL_S3 * N, 3 * N * sizeof(real), hipMemcpyDeviceToDevice, s));
            HIPL(Launch<real>::sources(at(AA_CORR), 0, 2, 0, 0, s));
        }
    }
    return PG_OK;
}

// Adjoint passes of the step held in slot m; lamA holds the adjoint of the
// step's end state on entry, and after the swap of the two the adjoint of
// its start state.
template <typename real>
int adj_step(pg_handle* h, int m, real*& lamA, real*& lamB, hipStream_t s) {
    Group& g = h->g;
    AdjWork* w = h->adj;
    DiskDev* d = g.disks[0];
    const size_t N = w->N;
    const int nx = d->nx, ny = d->ny;
    real* sl = (real*)w->slot + (size_t)m * SLOT_WORDS * N;
    real* lamV0 = (real*)w->lam + 6 * N;
    AdjArgs<real> A{};
    // the slot's descriptor: its control block and shift hold step m's dt,
    // planets, inc and res from the recompute
    A.d = (const DiskArgs<real>*)w->args + (size_t)m * AA_N + AA_X;
    A.lparam = w->lparam;
    A.nslot = w->nslot;
    A.q = d->P.n_planet > 0 ? d->P.planet_mass : 0.0;
    A.alpha = d->P.visc_flag ? d->P.ss_alpha : 0.0;
    const bool visc = g.visc;
    if (visc) {
        // corrector S4 = C(S3, V1), V1 = vis(T), predictor T = P(S3, V0):
        // the adjoint of S3 into lamB, of V0 into lamV0
        A.s = sl + SL_S3 * N; A.vis = sl + SL_V1 * N; A.s2 = sl + SL_T * N; A.vis2 = sl + SL_V0 * N;
        A.lo = lamA; A.li = lamB; A.lvis = lamV0; A.accumulate = 0;
        HIPL(AdjLaunch<real>::vis(A, 1, nx, ny, s));
    } else {
        // corrector S4 = C(S3)
        A.mode = 2; A.s = sl + SL_S3 * N; A.vis = nullptr; A.lo = lamA; A.li = lamB; A.accumulate = 0;
        HIPL(AdjLaunch<real>::src(A, nx, ny, s));
    }
    // S3 = shift + advection(SA)
    A.s = sl + SL_SA * N; A.lo = lamB; A.li = lamA; A.accumulate = 0;
    HIPL(AdjLaunch<real>::adv(A, nx, ny, s));
    // SA = azimuthal sweep(S2)
    A.s = sl + SL_S2 * N; A.lo = lamA; A.li = lamB; A.accumulate = 0;
    HIPL(AdjLaunch<real>::phi(A, nx, ny, s));
    // S2 = x sweep(S1)
    A.s = sl + SL_S1 * N; A.lo = lamB; A.li = lamA; A.accumulate = 0;
    HIPL(AdjLaunch<real>::x(A, nx, ny, s));
    if (visc) {
        // kick S1 = K(S0, V0) and V0 = vis(S0); lamV0 holds what the
        // predictor sent to V0
        A.s = sl + SL_S0 * N; A.vis = sl + SL_V0 * N; A.lo = lamA; A.li = lamB; A.lvis = lamV0; A.accumulate = 0;
        HIPL(AdjLaunch<real>::vis(A, 2, nx, ny, s));
    } else {
        // kick S1 = K(S0)
        A.mode = 0; A.s = sl + SL_S0 * N; A.vis = nullptr; A.lo = lamA; A.li = lamB; A.accumulate = 0;
        HIPL(AdjLaunch<real>::src(A, nx, ny, s));
    }
    std::swap(lamA, lamB);
    return PG_OK;
}

// The recompute of the segment of len steps held in `set`: every step's
// prologue, then the steps in order.
template <typename real>
int adj_recompute_segment(pg_handle* h, int set, int len, hipStream_t s) {
    AdjWork* w = h->adj;
    int rc = set_steps<real>(h->g, w, set, len, s);
    const int m0 = set * w->sps;
    for (int m = 0; m < len && !rc; ++m) rc = adj_recompute<real>(h, m0 + m, m == 0, m + 1 < len, s);
    return rc;
}

// The adjoint steps of the segment held in `set`, last to first.
template <typename real>
int adj_steps_segment(pg_handle* h, int set, int len, real*& lamA, real*& lamB, hipStream_t s) {
    const int m0 = set * h->adj->sps;
    int rc = PG_OK;
    for (int m = len - 1; m >= 0 && !rc; --m) rc = adj_step<real>(h, m0 + m, lamA, lamB, s);
    return rc;
}

int adj_check_config(const pg_params& P) {
    if (P.dust_flag || P.kill_flag)
        return fail(PG_E_INVALID, "the adjoint covers gas disks without dust and wave killing");
    return PG_OK;
}

// The tape forward: the device loop's step (prologue mode 1: the CFL step
// from the previous corrector's rate, capped at 1.1 dt_prev and clipped to
// end_time) with the recompute's kernels and descriptors, so that step n's
// intermediates, control block and orbital shift stay in slot n.  Each
// slot's control block starts as a copy of the previous one, as the disk's
// own block carries over from step to step in group_run; the arithmetic is
// the device loop's, so the states and step sizes equal pg_run's.
template <typename real>
int adj_tape_forward(pg_handle* h, int nsteps, double t0, double dt0) {
    Group& g = h->g;
    AdjWork* w = h->adj;
    DiskDev* d = g.disks[0];
    const size_t N = w->N;
    hipStream_t s = g.stream;
    int rc;
    if ((rc = pull_ctl(g))) return rc;
    StepCtl& c = *d->ctl_host;
    c.t = t0;
    c.dt = dt0;
    c.dt_prev = dt0;
    c.nstep = 0;
    c.mode = 1;
    c.end_time = d->P.end_time;
    c.rate_bits = 0ull;
    if ((rc = push_ctl(g))) return rc;
    if ((rc = launch_cfl(g))) return rc;   // rate of the start state
    real* base = (real*)w->slot;
    HIPL(hipMemcpyAsync(base + SL_S0 * N, d->gas[g.gpar], 3 * N * sizeof(real), hipMemcpyDeviceToDevice, s));
    HIPL(hipMemcpyAsync(w->sctl, d->ctl, sizeof(StepCtl), hipMemcpyDeviceToDevice, s));
    const LaunchSet<real> L = g.ls<real>();
    const bool visc = g.visc;
    for (int m = 0; m < nsteps; ++m) {
        const DiskArgs<real>* a = (const DiskArgs<real>*)w->args + (size_t)m * AA_N;
        auto at = [&](int k) {
            LaunchSet<real> x = L;
            x.disks = a + k;
            x.ndisk = 1;
            return x;
        };
        real* sl = base + (size_t)m * SLOT_WORDS * N;
        real* nxt = base + (size_t)(m + 1) * SLOT_WORDS * N;
        if (m > 0) HIPL(hipMemcpyAsync(w->sctl + m, w->sctl + m - 1, sizeof(StepCtl), hipMemcpyDeviceToDevice, s));
        HIPL(Launch<real>::prologue(at(AA_X), 1, 1, s));
        if (visc && m == 0) HIPL(Launch<real>::viscosity(at(AA_V0), 0, 0, s));
        HIPL(Launch<real>::sweep_x_gas(at(AA_X), 0, 1, s));
        HIPL(Launch<real>::ring_gas(at(AA_P), 0, 3, s));
        if (visc) {
            HIPL(Launch<real>::predvis(at(AA_PRED), 0, s));
            HIPL(Launch<real>::corrvis(at(AA_CORR), 0, 0, s));
        } else {
            HIPL(hipMemcpyAsync(nxt + SL_S0 * N, sl + SL_S3 * N, 3 * N * sizeof(real), hipMemcpyDeviceToDevice, s));
            HIPL(Launch<real>::sources(at(AA_CORR), 0, 2, 0, 1, s));
        }
    }
    // the disk ends at the last step's result, with the last step's control
    // block (t and dt of that step, as the device loop leaves them) and the
    // result's viscous forces (as the device loop's corrector leaves them)
    HIPL(hipMemcpyAsync(d->gas[g.gpar], base + (size_t)nsteps * SLOT_WORDS * N + SL_S0 * N, 3 * N * sizeof(real),
                        hipMemcpyDeviceToDevice, s));
    if (visc)
        HIPL(hipMemcpyAsync(d->vis, base + (size_t)nsteps * SLOT_WORDS * N + SL_V0 * N, 2 * N * sizeof(real),
                            hipMemcpyDeviceToDevice, s));
    HIPL(hipMemcpyAsync(d->ctl, w->sctl + nsteps - 1, sizeof(StepCtl), hipMemcpyDeviceToDevice, s));
    if ((rc = pull_ctl(g))) return rc;
    std::vector<StepCtl> cs(nsteps);
    HIPL(hipMemcpy(cs.data(), w->sctl, sizeof(StepCtl) * nsteps, hipMemcpyDeviceToHost));
    for (int n = 0; n < nsteps; ++n) {
        w->ts[n] = cs[n].t;
        w->dts[n] = cs[n].dt;
    }
    return PG_OK;
}

template <typename real>
int adj_forward(pg_handle* h, int nsteps, int ckpt, double t0, double dt0, double* dts_out) {
    Group& g = h->g;
    int rc;
    if (ckpt == 0 && g.ave) return fail(PG_E_STATE, "the tape forward runs without the time average");
    if ((rc = adj_alloc<real>(h, nsteps, ckpt))) return rc;
    AdjWork* w = h->adj;
    w->forward_done = false;
    // backward graphs hold the launch configuration of the previous run
    if (w->rs) HIPL(hipStreamSynchronize(w->rs));
    for (auto& sg : w->graphs)
        if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
    w->graphs.clear();
    const size_t bytes = 3 * w->N * sizeof(real);
    w->dts.assign(nsteps, 0.0);
    w->ts.assign(nsteps, 0.0);
    double t = t0, dt = dt0;
    for (int k = 0; k < w->nseg; ++k) {
        const int n0 = k * ckpt, len = std::min(ckpt, nsteps - n0);
        HIPL(hipMemcpyAsync((char*)w->ckpts + (size_t)k * bytes, g.disks[0]->gas[g.gpar], bytes,
                            hipMemcpyDeviceToDevice, g.stream));
        if ((rc = group_run(g, len, &t, &dt, w->dts.data() + n0, 1, nullptr))) return rc;
    }
    if (ckpt == 0) {
        if ((rc = adj_tape_forward<real>(h, nsteps, t0, dt0))) return rc;
    } else {
        // start times, summed in the order the device loop sums them
        double tt = t0;
        for (int n = 0; n < nsteps; ++n) {
            w->ts[n] = tt;
            tt += w->dts[n];
        }
    }
    std::vector<double> td(2 * (size_t)nsteps);
    for (int n = 0; n < nsteps; ++n) {
        td[2 * n] = w->ts[n];
        td[2 * n + 1] = w->dts[n];
    }
    HIPL(hipMemcpy(w->td, td.data(), td.size() * sizeof(double), hipMemcpyHostToDevice));
    if (dts_out) std::memcpy(dts_out, w->dts.data(), sizeof(double) * nsteps);
    // the slots' control blocks start as the disk's (planet elements,
    // softening); each slot's replay rewrites the step-dependent fields
    // (the tape's were written by its own forward steps)
    if (ckpt > 0)
        for (int m = 0; m < w->slots; ++m)
            HIPL(hipMemcpyAsync(w->sctl + m, g.disks[0]->ctl, sizeof(StepCtl), hipMemcpyDeviceToDevice, g.stream));
    HIPL(hipStreamSynchronize(g.stream));
    w->forward_done = true;
    return PG_OK;
}

template <typename real>
int adj_backward(pg_handle* h, const double* dL, pg_grads* out) {
    Group& g = h->g;
    AdjWork* w = h->adj;
    DiskDev* d = g.disks[0];
    const size_t N = w->N;
    const int nx = d->nx;
    int rc;
    // adjoint of the final state in the device representation:
    // rho (p = cs2 rho), u, vp (v = vbg + vp)
    // (through the workspace's pinned staging: no fresh host pages per call)
    real* st = (real*)w->stage;
    const double* cs2 = d->tabh.data() + (size_t)T_CS2 * d->tab_stride + NPAD;
    host_parallel(N / nx, [&](size_t j0, size_t j1) {
        for (size_t row = j0 * nx; row < j1 * nx; row += nx)
            for (int i = 0; i < nx; ++i) {
                const size_t o = row + i;
                st[o] = (real)(dL[o] + cs2[i] * dL[N + o]);
                st[N + o] = (real)dL[2 * N + o];
                st[2 * N + o] = (real)dL[3 * N + o];
            }
    });
    real* lamA = (real*)w->lam;
    real* lamB = (real*)w->lam + 3 * N;
    HIPL(hipMemcpyAsync(lamA, st, 3 * N * sizeof(real), hipMemcpyHostToDevice, g.stream));
    HIPL(hipMemsetAsync(w->lparam, 0, 2 * (size_t)w->nslot * sizeof(double), g.stream));
    HIPL(hipMemsetAsync((real*)w->lam + 6 * N, 0, 2 * N * sizeof(real), g.stream));
    const size_t bytes = 3 * N * sizeof(real);
    if (w->ckpt == 0) {
        // the tape: the adjoint steps alone, last to first (launched
        // directly: the host issues a step's five launches well within the
        // step's device time)
        for (int m = w->nsteps - 1; m >= 0; --m)
            if ((rc = adj_step<real>(h, m, lamA, lamB, g.stream))) return rc;
    }
    // Checkpointed: per segment, last to first, the recompute into a set of
    // slots, then the adjoint steps reading it.  With two sets the
    // recompute of segment k-1 runs on the recompute stream while the
    // adjoint steps of segment k run on the disk's stream: the adjoint
    // steps of k wait for the recompute of k (ev_r), and the recompute of
    // k-1 waits for the adjoint steps of k+1, the last reader of its set
    // (ev_a).  Each is a captured graph, replayed for every segment of the
    // same length (and, for the adjoint steps, the same incoming buffer).
    hipStream_t rs = w->nset == 2 ? w->rs : g.stream;
    auto launch = [&](int kind, int set, int len, real*& la, real*& lb, hipStream_t st) -> int {
        AdjWork::SegGraph* sg = nullptr;
        for (auto& x : w->graphs)
            if (x.kind == kind && x.set == set && x.len == len && (kind == 0 || x.lam_in == (void*)la)) sg = &x;
        if (!sg) {
            real* a = la;
            real* b = lb;
            HIPL(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int r = kind == 0 ? adj_recompute_segment<real>(h, set, len, st)
                                    : adj_steps_segment<real>(h, set, len, a, b, st);
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &graph);
            if (r) {
                if (graph) (void)hipGraphDestroy(graph);
                return r;
            }
            if (ec != hipSuccess) return fail(PG_E_HIP, "adjoint capture: %s", hipGetErrorString(ec));
            hipGraphExec_t exec = nullptr;
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) return fail(PG_E_HIP, "adjoint graph: %s", hipGetErrorString(ei));
            w->graphs.push_back({kind, set, len, (void*)la, (void*)a, exec});
            sg = &w->graphs.back();
        }
        HIPL(hipGraphLaunch(sg->exec, st));
        if (kind == 1 && sg->lam_out != (void*)la) std::swap(la, lb);
        return PG_OK;
    };
    auto recompute = [&](int k) -> int {
        const int set = w->nset == 2 ? (k & 1) : 0;
        const int n0 = k * w->ckpt, len = std::min(w->ckpt, w->nsteps - n0);
        real* s0 = (real*)w->slot + (size_t)set * w->sps * SLOT_WORDS * N + SL_S0 * N;
        HIPL(hipMemcpyAsync(s0, (char*)w->ckpts + (size_t)k * bytes, bytes, hipMemcpyDeviceToDevice, rs));
        HIPL(hipMemcpyAsync(w->td_seg + 2 * (size_t)set * w->ckpt, w->td + 2 * (size_t)n0,
                            2 * (size_t)len * sizeof(double), hipMemcpyDeviceToDevice, rs));
        real* none_a = nullptr;
        real* none_b = nullptr;
        int r = launch(0, set, len, none_a, none_b, rs);
        if (r) return r;
        if (w->nset == 2) HIPL(hipEventRecord(w->ev_r[set], rs));
        return PG_OK;
    };
    if (w->nseg > 0 && (rc = recompute(w->nseg - 1))) return rc;
    for (int k = w->nseg - 1; k >= 0; --k) {
        const int set = w->nset == 2 ? (k & 1) : 0;
        const int len = std::min(w->ckpt, w->nsteps - k * w->ckpt);
        if (k > 0) {
            if (k + 1 < w->nseg) HIPL(hipStreamWaitEvent(rs, w->ev_a[(k - 1) & 1], 0));
            if ((rc = recompute(k - 1))) return rc;
        }
        if (w->nset == 2) HIPL(hipStreamWaitEvent(g.stream, w->ev_r[set], 0));
        if ((rc = launch(1, set, len, lamA, lamB, g.stream))) return rc;
        if (w->nset == 2) HIPL(hipEventRecord(w->ev_a[set], g.stream));
    }
    HIPL(hipMemcpyAsync(st, lamA, 3 * N * sizeof(real), hipMemcpyDeviceToHost, g.stream));
    std::vector<double> slots(2 * (size_t)w->nslot);
    HIPL(hipMemcpyAsync(slots.data(), w->lparam, slots.size() * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIPL(hipStreamSynchronize(g.stream));
    const real* lam0 = st;
    double lp[2] = {0.0, 0.0};
    for (int k = 0; k < w->nslot; ++k) {
        lp[0] += slots[k];
        lp[1] += slots[(size_t)w->nslot + k];
    }
    // the initial radial velocity -1.5 nu(R)/R is proportional to alpha
    double dal = lp[1];
    if (d->P.visc_flag && d->P.ss_alpha > 0.0) {
        const double* gu = d->tabh.data() + (size_t)T_G_U * d->tab_stride + NPAD;
        double acc = 0.0;
        for (size_t row = 0; row < N; row += nx)
            for (int i = 0; i < nx; ++i) acc += (double)lam0[N + row + i] * gu[i];
        dal += acc / d->P.ss_alpha;
    }
    out->dq = lp[0];
    out->dalpha = dal;
    if (out->dstate0) {
        double* s0 = out->dstate0;
        host_parallel(N, [&](size_t o0, size_t o1) {
            for (size_t o = o0; o < o1; ++o) {
                s0[o] = (double)lam0[o];
                s0[N + o] = 0.0;
                s0[2 * N + o] = (double)lam0[N + o];
                s0[3 * N + o] = (double)lam0[2 * N + o];
                s0[4 * N + o] = 0.0;
            }
        });
    }
    return PG_OK;
}

#define CHECK_H(h) \
    if (!(h)) return fail(PG_E_INVALID, "null handle")
// entry points of the two-dimensional path only
#define CHECK_2D(h) \
    if ((h)->sph) return fail(PG_E_INVALID, "%s: not available for 3D disks", __func__)

}  // namespace

namespace pg {
int report_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
void host_make_axis(double* a, double* v, double amin, double amax, int res, int pad, int geom, int grid) {
    std::vector<double> va, vv;
    make_axis(va, vv, amin, amax, res, pad, geom, grid);
    std::copy(va.begin(), va.end(), a);
    std::copy(vv.begin(), vv.end(), v);
}
void host_slope_weights(int geom, const double* x6, double wl[3], double wr[3]) { slope_weights(geom, x6, wl, wr); }
}  // namespace pg

// =================================================================== C ABI
extern "C" {

const char* pg_last_error(void) { return g_err.c_str(); }
const char* pg_version(void) { return "penguin_amd 0.2 (gfx950)"; }

int pg_create(const pg_params* params, pg_handle** out) {
    g_err.clear();
    if (!params || !out) return fail(PG_E_INVALID, "null argument");
    *out = nullptr;
    if (params->ndim == 3) {   // the spherical disk
        pg::Sph* s = nullptr;
        const int rc = pg::sph_create(*params, &s);
        if (rc) return rc;
        pg_handle* h = new pg_handle();
        h->sph = s;
        *out = h;
        return PG_OK;
    }
    pg_handle* h = new pg_handle();
    int rc = group_init(h->g, 1, params);
    if (rc) {
        group_release(h->g);
        delete h;
        return rc;
    }
    *out = h;
    return PG_OK;
}

int pg_destroy(pg_handle* h) {
    if (!h) return PG_OK;
    if (h->sph) {
        pg::sph_destroy(h->sph);
        delete h;
        return PG_OK;
    }
    if (h->g.stream) (void)hipStreamSynchronize(h->g.stream);
    adj_free(h->adj);
    group_release(h->g);
    delete h;
    return PG_OK;
}

// Ghost fills of one direction on the host, from the handle's state: what
// the kernels read in place of stored ghost cells.  dir 0: radial ghosts,
// the analytic initial state (boundary type 0: the tables of the 2D disk,
// the profiles of the 3D one); 1: azimuthal, periodic; 2: polar, mirrored
// with w negated (type 2).
static int bound_dir(pg_handle* h, int dir, double* gas, double* dust) {
    CHECK_H(h);
    if (!gas && !dust) return fail(PG_E_INVALID, "null gas and dust");
    int nx, ny, nz;
    if (h->sph) pg::sph_dims(h->sph, &nx, &ny, &nz);
    else { nx = h->g.disks[0]->nx; ny = h->g.disks[0]->ny; nz = 1; }
    if (dir == 2 && !h->sph) return fail(PG_E_INVALID, "pg_bound_z: the disk is two-dimensional");
    if (dust && (h->sph || !h->g.disks[0]->P.dust_flag)) return fail(PG_E_STATE, "the disk has no dust");
    const size_t N = (size_t)nx * ny * nz;
    std::vector<double> g5(5 * N), d4(dust ? 4 * N : 0);
    int rc = pg_get_state(h, g5.data(), dust ? d4.data() : nullptr);
    if (rc) return rc;
    const int ni = nx + (dir == 0) * 2 * NPAD, nj = ny + (dir == 1) * 2 * NPAD, nk = nz + (dir == 2) * 2 * NPAD;
    const size_t M = (size_t)ni * nj * nk;
    for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const size_t o = (size_t)i + (size_t)ni * (j + (size_t)nj * k);
            int ic = i, jc = j, kc = k;   // the interior cell a ghost copies, if any
            double sgn_w = 1.0;
            bool analytic = false;
            if (dir == 0) { ic = i - NPAD; analytic = ic < 0 || ic >= nx; }
            if (dir == 1) jc = ((j - NPAD) % ny + ny) % ny;
            if (dir == 2) {
                kc = k - NPAD;
                if (kc < 0 || kc >= nz) { kc = kc < 0 ? -1 - kc : 2 * nz - 1 - kc; sgn_w = -1.0; }
            }
            if (analytic) {
                const int ip = i;   // padded radius
                if (h->sph) {
                    double f[5];
                    pg::sph_ghost_x(h->sph, ic < 0 ? ip : ip - nx, k, f);
                    if (gas) for (int q = 0; q < 5; ++q) gas[q * M + o] = f[q];
                } else {
                    const DiskDev* dd = h->g.disks[0];
                    const double rho = tabv(dd, T_G_R, ip);
                    if (gas) {
                        gas[o] = rho; gas[M + o] = tabv(dd, T_CS2, ip) * rho; gas[2 * M + o] = tabv(dd, T_G_U, ip);
                        gas[3 * M + o] = tabv(dd, T_VBG, ip) + tabv(dd, T_G_V, ip); gas[4 * M + o] = 0.0;
                    }
                    if (dust) {
                        dust[o] = tabv(dd, T_D_R, ip); dust[M + o] = tabv(dd, T_D_U, ip);
                        dust[2 * M + o] = tabv(dd, T_VBG, ip) + tabv(dd, T_D_V, ip); dust[3 * M + o] = 0.0;
                    }
                }
                continue;
            }
            const size_t c = (size_t)ic + (size_t)nx * (jc + (size_t)ny * kc);
            if (gas) {
                for (int q = 0; q < 4; ++q) gas[q * M + o] = g5[q * N + c];
                gas[4 * M + o] = sgn_w * g5[4 * N + c];
            }
            if (dust)
                for (int q = 0; q < 4; ++q) dust[q * M + o] = d4[q * N + c];
        }
    return PG_OK;
}

int pg_bound_x(pg_handle* h, double* gas, double* dust) { return bound_dir(h, 0, gas, dust); }
int pg_bound_y(pg_handle* h, double* gas, double* dust) { return bound_dir(h, 1, gas, dust); }
int pg_bound_z(pg_handle* h, double* gas, double* dust) { return bound_dir(h, 2, gas, dust); }

int pg_get_state(pg_handle* h, double* gas, double* dust) {
    CHECK_H(h);
    if (h->sph) return pg::sph_get_state(h->sph, gas);
    return disk_get_state(h->g, 0, gas, dust);
}

int pg_set_state(pg_handle* h, const double* gas, const double* dust) {
    CHECK_H(h);
    if (h->sph) return pg::sph_set_state(h->sph, gas);
    return disk_set_state(h->g, 0, gas, dust);
}

int pg_global_dt(pg_handle* h, double dt_prev, double* dt_out) {
    CHECK_H(h);
    if (!dt_out) return fail(PG_E_INVALID, "null dt_out");
    if (h->sph) return pg::sph_global_dt(h->sph, dt_prev, dt_out);
    Group& g = h->g;
    int rc;
    if ((rc = cfl_rate(g))) return rc;
    if ((rc = pull_ctl(g))) return rc;
    double rate = 0.0;
    unsigned long long bits = g.disks[0]->ctl_host->rate_bits;
    std::memcpy(&rate, &bits, sizeof(double));
    if (!(rate > 0.0) || !std::isfinite(rate)) return fail(PG_E_STATE, "CFL rate is %g (state not finite?)", rate);
    double dt = g.disks[0]->P.CFL / rate;
    double cap = dt_prev * 1.1;
    *dt_out = dt < cap ? dt : cap;
    return PG_OK;
}

int pg_solve(pg_handle* h, double t, double dt) {
    CHECK_H(h);
    if (!(dt > 0.0)) return fail(PG_E_INVALID, "dt must be positive");
    if (h->sph) return pg::sph_solve(h->sph, t, dt);
    Group& g = h->g;
    int rc;
    if ((rc = pull_ctl(g))) return rc;
    StepCtl& c = *g.disks[0]->ctl_host;
    c.t = t;
    c.dt = dt;
    c.mode = 0;
    c.rate_bits = 0ull;
    if ((rc = push_ctl(g))) return rc;
    if ((rc = prologue(g, 0, 1))) return rc;
    if ((rc = passes(g, false))) return rc;
    HIP_TRY(hipStreamSynchronize(g.stream));
    return timers_resolve(g);
}

int pg_run(pg_handle* h, int nsteps, double* t_io, double* dt_io, double* dts, int use_graph) {
    CHECK_H(h);
    if (h->sph) {
        if (!t_io || !dt_io || nsteps < 0) return fail(PG_E_INVALID, "pg_run: null t/dt or negative steps");
        return pg::sph_run(h->sph, nsteps, t_io, dt_io, dts);
    }
    return group_run(h->g, nsteps, t_io, dt_io, dts, use_graph, nullptr);
}

int pg_viscosity_eval(pg_handle* h, int on_T, double* vis_out) {
    CHECK_H(h);
    if (h->sph) return pg::sph_viscosity(h->sph, on_T, vis_out);
    Group& g = h->g;
    hipError_t e = g.prec == 8 ? Launch<double>::viscosity(g.ls<double>(), g.par(), on_T ? 1 : 0, g.stream)
                               : Launch<float>::viscosity(g.ls<float>(), g.par(), on_T ? 1 : 0, g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch viscosity failed: %s", hipGetErrorString(e));
    const DiskDev* d = g.disks[0];
    if (vis_out) {
        std::vector<double> v;
        int rc = download(g.prec, g.stream, d->vis, v, 2 * d->N);
        if (rc) return rc;
        std::memcpy(vis_out, v.data(), 2 * d->N * sizeof(double));
    }
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PG_OK;
}

int pg_planet(pg_handle* h, double t, double* body10) {
    CHECK_H(h);
    if (h->sph) {
        if (!body10) return PG_OK;
        return pg::sph_planet(h->sph, t, body10);
    }
    Group& g = h->g;
    int rc;
    if ((rc = pull_ctl(g))) return rc;
    StepCtl& c = *g.disks[0]->ctl_host;
    c.t = t;
    c.dt = 0.0;
    if ((rc = push_ctl(g))) return rc;
    // the orbital shift divides by dt: evaluate only the planet
    if ((rc = prologue(g, 0, 0))) return rc;
    if ((rc = pull_ctl(g))) return rc;
    if (body10) std::memcpy(body10, &c.planet0, sizeof(Body));
    return PG_OK;
}

int pg_sources(pg_handle* h, double t_planet, double mdt, double dt, int inplace) {
    CHECK_H(h);
    Group& g = h->g;
    int mode;
    if (mdt == 0.0) {
        if (!inplace) return fail(PG_E_INVALID, "mdt == 0 is only supported in place");
        mode = 0;
    } else {
        if (std::fabs(mdt - dt) > 1e-15 * std::fabs(dt))
            return fail(PG_E_INVALID, "mdt must be 0 or equal to dt (the two uses in DS)");
        mode = inplace ? 2 : 1;
    }
    if (h->sph) return pg::sph_sources(h->sph, t_planet, mdt, dt, inplace);
    // the kernel takes its half step from the control block: dt_block = 2*dt
    int rc;
    if ((rc = pull_ctl(g))) return rc;
    StepCtl& c = *g.disks[0]->ctl_host;
    c.t = t_planet;
    c.dt = 2.0 * dt;
    if ((rc = push_ctl(g))) return rc;
    if ((rc = prologue(g, 0, 0))) return rc;
    // planet0 holds the planet at t_planet; modes 1/2 read planet1, so copy
    if ((rc = pull_ctl(g))) return rc;
    c.planet1 = c.planet0;
    c.pk1 = c.pk0;
    if ((rc = push_ctl(g))) return rc;
    hipError_t e = g.prec == 8 ? Launch<double>::sources(g.ls<double>(), g.par(), mode, 0, 0, g.stream)
                               : Launch<float>::sources(g.ls<float>(), g.par(), mode, 0, 0, g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch sources failed: %s", hipGetErrorString(e));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PG_OK;
}

static int set_dt_and_prologue(Group& g, double dt, int orbadv) {
    int rc;
    if ((rc = pull_ctl(g))) return rc;
    g.disks[0]->ctl_host->dt = dt;
    if ((rc = push_ctl(g))) return rc;
    return prologue(g, 0, orbadv);
}

static int sweep_module(pg_handle* h, double dt, int which) {
    CHECK_H(h);
    Group& g = h->g;
    const bool dust = which == 1 || which == 3;
    if (dust && !g.dust) return fail(PG_E_STATE, "the disk has no dust");
    if (!(dt > 0.0)) return fail(PG_E_INVALID, "dt must be positive");
    int rc = set_dt_and_prologue(g, dt, 0);
    if (rc) return rc;
    hipError_t e;
    const int par = g.par();
    if (g.prec == 8) {
        const LaunchSet<double> L = g.ls<double>();
        e = which == 0 ? Launch<double>::sweep_x_gas(L, par, 0, g.stream)
            : which == 1 ? Launch<double>::sweep_x_dust(L, par, 0, g.stream)
            : which == 2 ? Launch<double>::ring_gas(L, par, 1, g.stream)
                         : Launch<double>::ring_dust(L, par, 1, g.stream);
    } else {
        const LaunchSet<float> L = g.ls<float>();
        e = which == 0 ? Launch<float>::sweep_x_gas(L, par, 0, g.stream)
            : which == 1 ? Launch<float>::sweep_x_dust(L, par, 0, g.stream)
            : which == 2 ? Launch<float>::ring_gas(L, par, 1, g.stream)
                         : Launch<float>::ring_dust(L, par, 1, g.stream);
    }
    if (e != hipSuccess) return fail(PG_E_HIP, "launch sweep failed: %s", hipGetErrorString(e));
    if (dust) g.dpar ^= 1;
    else g.gpar ^= 1;
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PG_OK;
}

int pg_sweep_x(pg_handle* h, double dt) {
    CHECK_H(h);
    if (h->sph) return pg::sph_sweep(h->sph, 0, dt);
    return sweep_module(h, dt, 0);
}
int pg_sweep_x_dust(pg_handle* h, double dt) {
    CHECK_H(h);
    CHECK_2D(h);
    return sweep_module(h, dt, 1);
}
// azimuthal sweep alone: the ring kernels with the orbital stages switched off
int pg_sweep_y(pg_handle* h, double dt) {
    CHECK_H(h);
    if (h->sph) return pg::sph_sweep(h->sph, 1, dt);
    return sweep_module(h, dt, 2);
}
int pg_sweep_y_dust(pg_handle* h, double dt) {
    CHECK_H(h);
    CHECK_2D(h);
    return sweep_module(h, dt, 3);
}
// polar-angle sweep (sweepz_inplace, sweeps.cu:458-553), 3D disks only
int pg_sweep_z(pg_handle* h, double dt) {
    CHECK_H(h);
    if (!h->sph) return fail(PG_E_INVALID, "pg_sweep_z: the disk is two-dimensional");
    if (!(dt > 0.0)) return fail(PG_E_INVALID, "dt must be positive");
    return pg::sph_sweep(h->sph, 2, dt);
}

// set_OrbAdv + shift_OrbAdv + advecty(_dust) alone: the ring kernels with
// their sweep stage switched off (each cell is carried unchanged to its
// shifted position, then advected by the residual speed).
int pg_orbital_advection(pg_handle* h, double dt, int32_t* inc, double* res) {
    CHECK_H(h);
    if (!(dt > 0.0)) return fail(PG_E_INVALID, "dt must be positive");
    if (h->sph) return pg::sph_orbital_advection(h->sph, dt, inc, res);
    Group& g = h->g;
    int rc = set_dt_and_prologue(g, dt, 1);
    if (rc) return rc;
    for (int fluid = 0; fluid < (g.dust ? 2 : 1); ++fluid) {
        hipError_t e;
        if (g.prec == 8) {
            const LaunchSet<double> L = g.ls<double>();
            e = fluid ? Launch<double>::ring_dust(L, g.par(), 2, g.stream) : Launch<double>::ring_gas(L, g.par(), 2, g.stream);
        } else {
            const LaunchSet<float> L = g.ls<float>();
            e = fluid ? Launch<float>::ring_dust(L, g.par(), 2, g.stream) : Launch<float>::ring_gas(L, g.par(), 2, g.stream);
        }
        if (e != hipSuccess) return fail(PG_E_HIP, "launch ring failed: %s", hipGetErrorString(e));
        if (fluid) g.dpar ^= 1;
        else g.gpar ^= 1;
    }
    const DiskDev* d = g.disks[0];
    if (inc) HIP_TRY(hipMemcpyAsync(inc, d->inc, d->nx * sizeof(int), hipMemcpyDeviceToHost, g.stream));
    if (res) {
        std::vector<double> r;
        if ((rc = download(g.prec, g.stream, d->res, r, d->nx))) return rc;
        std::memcpy(res, r.data(), d->nx * sizeof(double));
    }
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PG_OK;
}

int pg_killwave(pg_handle* h, double dt) {
    CHECK_H(h);
    CHECK_2D(h);
    Group& g = h->g;
    if (!g.kill) return PG_OK;
    int rc = set_dt_and_prologue(g, dt, 0);
    if (rc) return rc;
    hipError_t e = g.prec == 8 ? Launch<double>::sources(g.ls<double>(), g.par(), 3, 1, 0, g.stream)
                               : Launch<float>::sources(g.ls<float>(), g.par(), 3, 1, 0, g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch killwave failed: %s", hipGetErrorString(e));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PG_OK;
}

int pg_get_energy(const pg_params* P, double rho, double p, double u, double v, double w, double* e) {
    if (!P || !e) return fail(PG_E_INVALID, "null argument");
    if (!(rho > 0.0)) return fail(PG_E_INVALID, "rho must be positive");
    (void)u; (void)v; (void)w;   // kinetic terms only under internal_e_flag 0
    if (P->EOS_flag == 2) {
        if (!(P->gam > 1.0)) return fail(PG_E_INVALID, "EOS_flag 2 needs gam > 1");
        *e = p / ((P->gam - 1.0) * rho);
    } else if (P->EOS_flag == 1) {
        *e = p / std::pow(rho, P->gam);
    } else if (P->EOS_flag == 0) {
        *e = p / rho;
    } else {
        return fail(PG_E_INVALID, "EOS_flag=%d", P->EOS_flag);
    }
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

}  // extern "C"

namespace {

// write_check_point layout (output.cu:125-180): t, x edges, y edges, then
// the interior fields rho, p, u, v (2D: no w) and, with dust, its four
int write_snapshot_file(const DiskDev* d, const char* filename, double t, const std::vector<double>& gas,
                        const std::vector<double>& dust) {
    const size_t N = d->N;
    std::ofstream f(filename, std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f) return fail(PG_E_INVALID, "cannot open %s for writing", filename);
    std::vector<double> xa(d->nx + 1), ya(d->ny + 1);
    for (int i = 0; i <= d->nx; ++i) xa[i] = tabv(d, T_XA, i + NPAD);
    for (int j = 0; j <= d->ny; ++j) ya[j] = d->P.ymin + (d->P.ymax - d->P.ymin) * (double)j / (double)d->ny;
    f.write((const char*)&t, sizeof(double));
    f.write((const char*)xa.data(), sizeof(double) * xa.size());
    f.write((const char*)ya.data(), sizeof(double) * ya.size());
    f.write((const char*)gas.data(), sizeof(double) * 4 * N);  // rho, p, u, v (ndim 2: no w)
    if (d->P.dust_flag) f.write((const char*)dust.data(), sizeof(double) * 4 * N);  // w_d too (output.cu:171)
    if (!f) return fail(PG_E_INVALID, "write to %s failed", filename);
    return PG_OK;
}

// the gas time average in the pg_get_state layout (rho, p, u, v, w)
int average_get(pg_handle* h, double* gas) {
    Group& g = h->g;
    DiskDev* d = g.disks[0];
    if (!d->ave) return fail(PG_E_STATE, "no time average: call pg_average_begin first");
    const size_t N = d->N;
    const int nx = d->nx;
    std::vector<double> a(3 * N);
    HIP_TRY(hipMemcpyAsync(a.data(), d->ave, 3 * N * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    int rc = pull_ctl(g);
    if (rc) return rc;
    const double W = d->ctl_host->ave_w;
    for (size_t o = 0; o < N; ++o) {
        const int ip = (int)(o % nx) + NPAD;
        gas[o] = d->rho_bg[o % nx] * W + a[o];
        gas[N + o] = tabv(d, T_CS2, ip) * gas[o];
        gas[2 * N + o] = a[N + o];
        gas[3 * N + o] = tabv(d, T_VBG, ip) * W + a[2 * N + o];
        gas[4 * N + o] = 0.0;
    }
    return PG_OK;
}

}  // namespace

// ------------------------------------------------- allocation-free torque
// pg_torque_noalloc keeps the torque's per-tile partial sums, and their sum
// in the last slot, in this file-scope device array instead of a buffer the
// library allocates.  Each GPU holds its own copy until the process ends;
// every call on one GPU shares it (calls that overlap on a GPU race), and a
// call overwrites every slot it reads, so nothing needs clearing.
constexpr int NOALLOC_TILES = 16384;
static __device__ __align__(256) double scratch[NOALLOC_TILES + 1];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch[NOALLOC_MAX_DEVICES] = {};   // the array's address on each device

// The device address of `scratch` on the current device, found once per
// device (hipGetSymbolAddress with that device current).  The array's name
// in host code is not a device address and is never passed to a kernel.
int scratch_on_device(double** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(PG_E_INVALID, "device %d: the allocation-free entry points serve devices 0..%d", dev,
                    NOALLOC_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(PG_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev] = (double*)ptr;
    }
    *out = g_scratch[dev];
    return PG_OK;
}

// The partials summed by one thread in index order, as pg_torque sums them
// on the host, so the two entry points return the same bits.
__global__ void k_sum_in_order(const double* part, int n, double* sum) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += part[k];
    *sum = s;
}

}  // namespace

extern "C" {

int pg_write_snapshot(pg_handle* h, const char* filename, double t) {
    CHECK_H(h);
    if (!filename) return fail(PG_E_INVALID, "null filename");
    if (h->sph) return pg::sph_write_snapshot(h->sph, filename, t);
    const DiskDev* d = h->g.disks[0];
    const size_t N = d->N;
    std::vector<double> gas(5 * N), dust(d->P.dust_flag ? 4 * N : 0);
    int rc = disk_get_state(h->g, 0, gas.data(), d->P.dust_flag ? dust.data() : nullptr);
    if (rc) return rc;
    return write_snapshot_file(d, filename, t, gas, dust);
}

int pg_write_average_snapshot(pg_handle* h, const char* filename, double t) {
    CHECK_H(h);
    CHECK_2D(h);
    if (!filename) return fail(PG_E_INVALID, "null filename");
    const DiskDev* d = h->g.disks[0];
    const size_t N = d->N;
    std::vector<double> gas(5 * N), dust(d->P.dust_flag ? 4 * N : 0);
    int rc = average_get(h, gas.data());
    if (rc) return rc;
    // the dust is the current state, as cpy_grid_DevicetoHost(.., true) leaves it
    if (d->P.dust_flag && (rc = disk_get_state(h->g, 0, nullptr, dust.data()))) return rc;
    return write_snapshot_file(d, filename, t, gas, dust);
}

int pg_average_begin(pg_handle* h, double interval) {
    CHECK_H(h);
    CHECK_2D(h);
    if (!(interval > 0.0)) return fail(PG_E_INVALID, "interval must be > 0");
    Group& g = h->g;
    DiskDev* d = g.disks[0];
    int rc;
    if (!d->ave && (rc = dev_alloc(g.bytes, &d->ave, 3 * d->N * sizeof(double)))) return rc;
    HIP_TRY(hipMemsetAsync(d->ave, 0, 3 * d->N * sizeof(double), g.stream));
    if ((rc = pull_ctl(g))) return rc;
    d->ctl_host->ave_interval = interval;
    d->ctl_host->ave_w = 0.0;
    if ((rc = push_ctl(g)) || (rc = group_upload_args(g))) return rc;
    g.ave = true;
    g.drop_graph();
    return PG_OK;
}

int pg_average_end(pg_handle* h) {
    CHECK_H(h);
    CHECK_2D(h);
    h->g.ave = false;
    h->g.drop_graph();
    return PG_OK;
}

int pg_average_get(pg_handle* h, double* gas) {
    CHECK_H(h);
    CHECK_2D(h);
    if (!gas) return fail(PG_E_INVALID, "null argument");
    return average_get(h, gas);
}

int pg_torque(pg_handle* h, double* torque) {
    CHECK_H(h);
    CHECK_2D(h);
    if (!torque) return fail(PG_E_INVALID, "null argument");
    Group& g = h->g;
    DiskDev* d = g.disks[0];
    if (d->P.n_planet < 1) return fail(PG_E_STATE, "the disk has no planet");
    int rc;
    const int ntile = ((d->nx + PT_X - 1) / PT_X) * ((d->ny + PT_TY - 1) / PT_TY);   // the most tiles
    if (!d->diag && (rc = dev_alloc(g.bytes, &d->diag, ntile * sizeof(double)))) return rc;
    hipError_t e = g.prec == 8 ? Launch<double>::torque(g.ls<double>(), g.par(), d->diag, g.stream)
                               : Launch<float>::torque(g.ls<float>(), g.par(), d->diag, g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch torque failed: %s", hipGetErrorString(e));
    const int n = d->rate_nblk;   // tiles of the current schedule
    std::vector<double> part(n);
    HIP_TRY(hipMemcpyAsync(part.data(), d->diag, n * sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    double s = 0.0;
    for (double x : part) s += x;
    *torque = s;
    return PG_OK;
}

// pg_torque with the partial sums in the device array `scratch` (at most
// NOALLOC_TILES tiles of the current schedule): no device or pinned memory
// is allocated; the result comes back through the caller's pointer.
int pg_probe_grid_barrier(int blocks, int npass, double* us_per_barrier, double* us_per_node) {
    std::string msg;
    const int rc = probe_grid_barrier(blocks, npass, us_per_barrier, us_per_node, msg);
    return rc == PG_OK ? PG_OK : fail(rc, "%s", msg.c_str());
}

int pg_torque_noalloc(pg_handle* h, double* torque) {
    CHECK_H(h);
    CHECK_2D(h);
    if (!torque) return fail(PG_E_INVALID, "null argument");
    Group& g = h->g;
    DiskDev* d = g.disks[0];
    if (d->P.n_planet < 1) return fail(PG_E_STATE, "the disk has no planet");
    const int n = d->rate_nblk;   // tiles of the current schedule: one partial each
    if (n > NOALLOC_TILES)
        return fail(PG_E_INVALID, "pg_torque_noalloc: the disk has %d tiles of %d x %d cells; the allocation-free "
                    "scratch holds %d (use pg_torque)", n, PT_X, g.pt_rows, NOALLOC_TILES);
    double* ptr = nullptr;
    int rc = scratch_on_device(&ptr);
    if (rc) return rc;
    hipError_t e = g.prec == 8 ? Launch<double>::torque(g.ls<double>(), g.par(), ptr, g.stream)
                               : Launch<float>::torque(g.ls<float>(), g.par(), ptr, g.stream);
    if (e != hipSuccess) return fail(PG_E_HIP, "launch torque failed: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(k_sum_in_order, dim3(1), dim3(64), 0, g.stream, (const double*)ptr, n, ptr + NOALLOC_TILES);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(torque, ptr + NOALLOC_TILES, sizeof(double), hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return PG_OK;
}

int pg_read_snapshot(pg_handle* h, const char* filename, double* t_out) {
    CHECK_H(h);
    if (!filename) return fail(PG_E_INVALID, "null filename");
    if (h->sph) return pg::sph_read_snapshot(h->sph, filename, t_out);
    const DiskDev* d = h->g.disks[0];
    const size_t N = d->N;
    const size_t nfield = 4 + (d->P.dust_flag ? 4 : 0);
    const size_t expect = sizeof(double) * (1 + (d->nx + 1) + (d->ny + 1) + nfield * N);
    std::ifstream f(filename, std::ios::in | std::ios::binary);
    if (!f) return fail(PG_E_INVALID, "cannot open %s", filename);
    f.seekg(0, std::ios::end);
    size_t size = (size_t)f.tellg();
    if (size != expect)
        return fail(PG_E_INVALID, "%s has %zu bytes, a %dx%d%s snapshot has %zu", filename, size, d->nx, d->ny,
                    d->P.dust_flag ? " gas+dust" : "", expect);
    f.seekg(0);
    double t = 0.0;
    f.read((char*)&t, sizeof(double));
    std::vector<double> xa(d->nx + 1), ya(d->ny + 1);
    f.read((char*)xa.data(), sizeof(double) * xa.size());
    f.read((char*)ya.data(), sizeof(double) * ya.size());
    for (int i = 0; i <= d->nx; ++i)
        if (std::fabs(xa[i] - tabv(d, T_XA, i + NPAD)) > 1e-12 * std::fabs(xa[i]))
            return fail(PG_E_INVALID, "%s: radial grid does not match this disk", filename);
    std::vector<double> gas(5 * N, 0.0), dust(d->P.dust_flag ? 4 * N : 0);
    f.read((char*)gas.data(), sizeof(double) * 4 * N);
    if (d->P.dust_flag) f.read((char*)dust.data(), sizeof(double) * 4 * N);
    if (!f) return fail(PG_E_INVALID, "read from %s failed", filename);
    int rc = disk_set_state(h->g, 0, gas.data(), d->P.dust_flag ? dust.data() : nullptr);
    if (rc) return rc;
    if (t_out) *t_out = t;
    return PG_OK;
}

int pg_stream(pg_handle* h, void** stream) {
    CHECK_H(h);
    if (!stream) return fail(PG_E_INVALID, "null argument");
    *stream = h->sph ? pg::sph_stream(h->sph) : (void*)h->g.stream;
    return PG_OK;
}

int pg_profile_enable(pg_handle* h, int on) {
    CHECK_H(h);
    if (h->sph) return pg::sph_profile_enable(h->sph, on);
    h->g.profile = on != 0;
    h->g.timers.clear();
    h->g.ev_used = 0;
    return PG_OK;
}

int pg_profile_get(pg_handle* h, const char* kernel, double* avg_us, int* count) {
    CHECK_H(h);
    if (h->sph) return pg::sph_profile_get(h->sph, kernel, avg_us, count);
    return group_profile_get(h->g, kernel, avg_us, count);
}

int pg_device_bytes(pg_handle* h, size_t* bytes) {
    CHECK_H(h);
    if (!bytes) return fail(PG_E_INVALID, "null argument");
    if (h->sph) return pg::sph_device_bytes(h->sph, bytes);
    *bytes = h->g.bytes + (h->adj ? h->adj->bytes : 0);   // with the adjoint's workspace
    return PG_OK;
}

// ---- adjoint ---------------------------------------------------------------
int pg_adjoint_forward(pg_handle* h, int nsteps, int ckpt_every, double t0, double dt0, double* dts) {
    CHECK_H(h);
    CHECK_2D(h);
    if (nsteps < 1 || ckpt_every < 0) return fail(PG_E_INVALID, "nsteps must be >= 1 and ckpt_every >= 0");
    int rc = adj_check_config(h->g.disks[0]->P);
    if (rc) return rc;
    return h->g.prec == 8 ? adj_forward<double>(h, nsteps, ckpt_every, t0, dt0, dts)
                          : adj_forward<float>(h, nsteps, ckpt_every, t0, dt0, dts);
}

int pg_adjoint_backward(pg_handle* h, const double* dL_dstate, pg_grads* grads) {
    CHECK_H(h);
    CHECK_2D(h);
    if (!dL_dstate || !grads) return fail(PG_E_INVALID, "null argument");
    if (!h->adj || !h->adj->forward_done) return fail(PG_E_STATE, "pg_adjoint_forward has not run");
    return h->g.prec == 8 ? adj_backward<double>(h, dL_dstate, grads) : adj_backward<float>(h, dL_dstate, grads);
}

}  // extern "C"

template <typename real>
static int adjoint_pass(pg_handle* h, int pass, double t, double dt, const double* lam_out, double* lam_in,
                        double* dparam) {
    Group& g = h->g;
    DiskDev* d = g.disks[0];
    const size_t N = d->N;
    int rc;
    // control block as the module entry point sets it
    if ((rc = pull_ctl(g))) return rc;
    StepCtl& c = *d->ctl_host;
    const bool src = pass >= 4;
    c.t = t;
    c.dt = src ? 2.0 * dt : dt;
    if ((rc = push_ctl(g))) return rc;
    if ((rc = prologue(g, 0, pass == 2 ? 1 : 0))) return rc;
    if (src) {
        if ((rc = pull_ctl(g))) return rc;
        c.planet1 = c.planet0;
        c.pk1 = c.pk0;
        if ((rc = push_ctl(g))) return rc;
    }
    size_t scratch = 0;
    real *lo = nullptr, *li = nullptr, *lv = nullptr;
    double* lp = nullptr;
    auto cleanup = [&](int code) {
        void* b[] = {lo, li, lv, lp};
        for (void* x : b)
            if (x) (void)hipFree(x);
        return code;
    };
    if ((rc = dev_alloc(scratch, &lo, 3 * N * sizeof(real))) || (rc = dev_alloc(scratch, &li, 3 * N * sizeof(real))) ||
        (rc = dev_alloc(scratch, &lv, 2 * N * sizeof(real))) ||
        (rc = dev_alloc(scratch, &lp, 2 * (size_t)adj_slots(d->nx, d->ny) * sizeof(double))))
        return cleanup(rc);
    const int nslot = adj_slots(d->nx, d->ny);
    std::vector<double> lin(lam_out, lam_out + (pass == 3 ? 2 : 3) * N);
    if ((rc = upload(g.prec, g.stream, lin, pass == 3 ? (void*)lv : (void*)lo))) return cleanup(rc);
    if (hipMemsetAsync(lp, 0, 2 * (size_t)nslot * sizeof(double), g.stream) != hipSuccess)
        return cleanup(fail(PG_E_HIP, "memset"));
    AdjArgs<real> A{};
    A.d = (const DiskArgs<real>*)g.d_args;
    A.s = (const real*)d->gas[g.gpar];
    A.vis = (const real*)d->vis;
    A.lo = lo;
    A.li = li;
    A.lvis = lv;
    A.lparam = lp;
    A.nslot = nslot;
    A.q = d->P.n_planet > 0 ? d->P.planet_mass : 0.0;
    A.alpha = d->P.visc_flag ? d->P.ss_alpha : 0.0;
    A.accumulate = 0;
    hipError_t e = hipSuccess;
    switch (pass) {
        case 0: e = AdjLaunch<real>::x(A, d->nx, d->ny, g.stream); break;
        case 1: e = AdjLaunch<real>::phi(A, d->nx, d->ny, g.stream); break;
        case 2: e = AdjLaunch<real>::adv(A, d->nx, d->ny, g.stream); break;
        case 3: e = AdjLaunch<real>::vis(A, 0, d->nx, d->ny, g.stream); break;
        default: A.mode = pass - 4; e = AdjLaunch<real>::src(A, d->nx, d->ny, g.stream); break;
    }
    if (e != hipSuccess) return cleanup(fail(PG_E_HIP, "adjoint pass launch: %s", hipGetErrorString(e)));
    std::vector<double> out;
    if ((rc = download(g.prec, g.stream, li, out, 3 * N))) return cleanup(rc);
    std::memcpy(lam_in, out.data(), 3 * N * sizeof(double));
    if (dparam) {
        std::vector<double> sl(2 * (size_t)nslot);
        if (hipMemcpy(sl.data(), lp, sl.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            return cleanup(fail(PG_E_HIP, "download"));
        dparam[0] = dparam[1] = 0.0;
        for (int k = 0; k < nslot; ++k) {
            dparam[0] += sl[k];
            dparam[1] += sl[(size_t)nslot + k];
        }
    }
    return cleanup(PG_OK);
}

extern "C" {

int pg_adjoint_pass(pg_handle* h, int pass, double t, double dt, const double* lam_out, double* lam_in,
                    double* dparam) {
    CHECK_H(h);
    CHECK_2D(h);
    if (pass < 0 || pass > 6) return fail(PG_E_INVALID, "pass %d: use 0..6", pass);
    if (!lam_out || !lam_in) return fail(PG_E_INVALID, "null argument");
    int rc = adj_check_config(h->g.disks[0]->P);
    if (rc) return rc;
    return h->g.prec == 8 ? adjoint_pass<double>(h, pass, t, dt, lam_out, lam_in, dparam)
                          : adjoint_pass<float>(h, pass, t, dt, lam_out, lam_in, dparam);
}

// ---- batches ---------------------------------------------------------------
int pg_batch_create(int n, const pg_params* params, pg_batch** out) {
    g_err.clear();
    if (!params || !out || n < 1) return fail(PG_E_INVALID, "bad arguments");
    *out = nullptr;
    pg_batch* b = new pg_batch();
    int rc = group_init(b->g, n, params);
    if (rc) {
        group_release(b->g);
        delete b;
        return rc;
    }
    *out = b;
    return PG_OK;
}

int pg_batch_destroy(pg_batch* b) {
    if (!b) return PG_OK;
    group_release(b->g);
    delete b;
    return PG_OK;
}

int pg_batch_size(pg_batch* b, int* n) {
    CHECK_H(b);
    if (!n) return fail(PG_E_INVALID, "null argument");
    *n = b->g.ndisk();
    return PG_OK;
}

int pg_batch_get_state(pg_batch* b, int disk, double* gas, double* dust) {
    CHECK_H(b);
    if (disk < 0 || disk >= b->g.ndisk()) return fail(PG_E_INVALID, "disk %d out of range", disk);
    return disk_get_state(b->g, disk, gas, dust);
}

int pg_batch_set_state(pg_batch* b, int disk, const double* gas, const double* dust) {
    CHECK_H(b);
    if (disk < 0 || disk >= b->g.ndisk()) return fail(PG_E_INVALID, "disk %d out of range", disk);
    return disk_set_state(b->g, disk, gas, dust);
}

int pg_batch_run(pg_batch* b, int nsteps, double* t_io, double* dt_io, double* dts, int use_graph) {
    CHECK_H(b);
    return group_run(b->g, nsteps, t_io, dt_io, dts, use_graph, nullptr);
}

int pg_batch_set_schedule(pg_batch* b, int tile_x, int ring_b, int order, int pt_rows) {
    CHECK_H(b);
    return group_schedule(b->g, tile_x, ring_b, order, pt_rows);
}

int pg_batch_get_schedule(pg_batch* b, int* tile_x, int* ring_b, int* order, int* pt_rows) {
    CHECK_H(b);
    if (!tile_x || !ring_b || !order || !pt_rows) return fail(PG_E_INVALID, "null argument");
    *tile_x = b->g.tile_x;
    *ring_b = b->g.ring_b;
    *order = b->g.order;
    *pt_rows = b->g.pt_rows;
    return PG_OK;
}

int pg_batch_autotune(pg_batch* b, int steps, double* best_ms) {
    CHECK_H(b);
    return group_autotune(b->g, steps, best_ms);
}

int pg_batch_profile_enable(pg_batch* b, int on) {
    CHECK_H(b);
    b->g.profile = on != 0;
    b->g.timers.clear();
    b->g.ev_used = 0;
    return PG_OK;
}

int pg_batch_profile_get(pg_batch* b, const char* kernel, double* avg_us, int* count) {
    CHECK_H(b);
    return group_profile_get(b->g, kernel, avg_us, count);
}

int pg_batch_device_bytes(pg_batch* b, size_t* bytes) {
    CHECK_H(b);
    if (!bytes) return fail(PG_E_INVALID, "null argument");
    *bytes = b->g.bytes;
    return PG_OK;
}

}  // extern "C"
