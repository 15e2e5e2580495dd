// C-ABI of the channel-flow time-step library: a context owns the device
// state of one channel case (one HIP stream per case) and drives the RK3
// step of the reference (time_integration.f90:89-173) entirely on the GPU.
// Host precomputation follows initialization.f90 (grids, operators,
// weights); device scalars (dt, dPdx, ...) never round-trip through the host
// inside a step, so steps can be captured in HIP graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/channel_flow.h"
#include "cf_eqwm.hpp"
#include "cf_internal.hpp"
#include "cf_kernels.hpp"

namespace cf {

thread_local std::string g_err;

void set_error(const std::string &m) { g_err = m; }
Batch &launch_batch() {
    static thread_local Batch b;
    return b;
}
std::string last_error_copy() { return g_err; }

#define HCK(call)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #call);      \
            return -1;                                                                         \
        }                                                                                      \
    } while (0)

// radix 8 first when radix8 (both transforms use it: fewer LDS passes),
// then 4, one 6 for a remaining 2 x 3, 2, 3, 5 and 7; the device's compiled
// plans (cplan) factor the same way, so the pass twiddle tables match
static FFTPlan make_plan(int n, std::string &err, bool radix8) {
    FFTPlan p{};
    p.n = n;
    int m = n;
    while (radix8 && m % 8 == 0 && p.nr < 16) { p.radix[p.nr++] = 8; m /= 8; }
    while (m % 4 == 0 && p.nr < 16) { p.radix[p.nr++] = 4; m /= 4; }
    if (m % 6 == 0 && p.nr < 16) { p.radix[p.nr++] = 6; m /= 6; }  // the last 2 with a 3: one pass
    while (m % 2 == 0 && p.nr < 16) { p.radix[p.nr++] = 2; m /= 2; }
    while (m % 3 == 0 && p.nr < 16) { p.radix[p.nr++] = 3; m /= 3; }
    while (m % 5 == 0 && p.nr < 16) { p.radix[p.nr++] = 5; m /= 5; }
    while (m % 7 == 0 && p.nr < 16) { p.radix[p.nr++] = 7; m /= 7; }
    if (m != 1) err = "FFT length " + std::to_string(n) + " has a prime factor other than 2, 3, 5 or 7";
    return p;
}

// interpolation.f90:114-252 stencil search, 1-based centre index + weights
static void interp_2nd(int ny, const double *y, int n, const double *yi, bool ge, int *jjo,
                       double *w) {
    auto Y = [&](int j) { return y[j - 1]; };
    int jj = ny - 1;
    for (int t = 1; t <= n; t++) {
        const double yref = yi[t - 1];
        if (t == 1) jj = 2;
        if (t == n) jj = ny - 1;
        for (int j2 = 3; j2 <= ny; j2++) {
            const bool hit = ge ? (Y(j2) >= yref) : (Y(j2) > yref);
            if (hit) { jj = j2 - 1; break; }
        }
        if (t == 1 && ge && jj == 0) jj = ny;
        w[3 * (t - 1) + 0] = (yref - Y(jj)) * (yref - Y(jj + 1)) / ((Y(jj - 1) - Y(jj)) * (Y(jj - 1) - Y(jj + 1)));
        w[3 * (t - 1) + 1] = (yref - Y(jj - 1)) * (yref - Y(jj + 1)) / ((Y(jj) - Y(jj - 1)) * (Y(jj) - Y(jj + 1)));
        w[3 * (t - 1) + 2] = (yref - Y(jj - 1)) * (yref - Y(jj)) / ((Y(jj + 1) - Y(jj - 1)) * (Y(jj + 1) - Y(jj)));
        jjo[t - 1] = jj;
    }
}

// ---------------------------------------------------------------------------
// Host-side operators (initialization.f90:164-463)
// ---------------------------------------------------------------------------
Host::Host(const cf_params &p, const double *x_, const double *y_, const double *z_) : prm(p) {
    d.nx = p.nx; d.ny = p.ny; d.nz = p.nz;
    d.nxg = p.nx + 1; d.nyg = p.ny + 1; d.nzg = p.nz + 1;
    d.nxp = p.nx - 2; d.nzp = p.nz - 2; d.nyp = d.nyg - 2;
    d.mh = d.nxp / 2 + 1;
    x.assign(x_, x_ + p.nx);
    y.assign(y_, y_ + p.ny);
    z.assign(z_, z_ + p.nz);
    const int nx = d.nx, ny = d.ny, nz = d.nz, nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    xg.assign(nxg, 0); yg.assign(nyg, 0); zg.assign(nzg, 0);
    auto centers = [](const std::vector<double> &f, std::vector<double> &g) {
        const int n = (int)f.size(), nm = n - 1;
        std::vector<double> m(nm);
        for (int i = 0; i < nm; i++) m[i] = 0.5 * (f[i] + f[i + 1]);
        for (int i = 0; i < nm; i++) g[i + 1] = m[i];
        g[0] = m[0] - 2.0 * (m[0] - f[0]);
        g[nm + 1] = m[nm - 1] + 2.0 * (f[n - 1] - m[nm - 1]);
    };
    centers(x, xg);
    centers(y, yg);
    centers(z, zg);
    auto dmin = [](const std::vector<double> &g) {
        double m = 1e300;
        for (size_t i = 1; i < g.size(); i++) m = std::min(m, g[i] - g[i - 1]);
        return m;
    };
    dxmin = dmin(xg); dymin = dmin(yg); dzmin = dmin(zg);
    // 1-based reciprocal metrics
    rdxf.assign(nx + 2, 0); rdxc.assign(nxg + 2, 0); dxc.assign(nxg + 2, 0);
    rdyf.assign(ny + 2, 0); rdyc.assign(nyg + 2, 0); dyc.assign(nyg + 2, 0);
    rdzf.assign(nz + 2, 0); rdzc.assign(nzg + 2, 0); dzc.assign(nzg + 2, 0);
    dyg.assign(nyg + 2, 0);
    for (int i = 2; i <= nx; i++) rdxf[i] = 1.0 / (x[i - 1] - x[i - 2]);
    for (int i = 1; i <= nxg - 1; i++) { dxc[i] = xg[i] - xg[i - 1]; rdxc[i] = 1.0 / dxc[i]; }
    for (int j = 2; j <= ny; j++) rdyf[j] = 1.0 / (y[j - 1] - y[j - 2]);
    for (int j = 1; j <= nyg - 1; j++) { dyc[j] = yg[j] - yg[j - 1]; rdyc[j] = 1.0 / dyc[j]; }
    for (int k = 2; k <= nz; k++) rdzf[k] = 1.0 / (z[k - 1] - z[k - 2]);
    for (int k = 1; k <= nzg - 1; k++) { dzc[k] = zg[k] - zg[k - 1]; rdzc[k] = 1.0 / dzc[k]; }
    for (int j = 2; j <= nyg; j++) dyg[j] = yg[j - 1] - yg[j - 2];
    x1.assign(nx + 1, 0); y1.assign(ny + 1, 0); z1.assign(nz + 1, 0);
    xg1.assign(nxg + 1, 0); yg1.assign(nyg + 1, 0); zg1.assign(nzg + 1, 0);
    for (int i = 1; i <= nx; i++) x1[i] = x[i - 1];
    for (int j = 1; j <= ny; j++) y1[j] = y[j - 1];
    for (int k = 1; k <= nz; k++) z1[k] = z[k - 1];
    for (int i = 1; i <= nxg; i++) xg1[i] = xg[i - 1];
    for (int j = 1; j <= nyg; j++) yg1[j] = yg[j - 1];
    for (int k = 1; k <= nzg; k++) zg1[k] = zg[k - 1];
    // initialization.f90:386-388
    w0.assign(ny + 2, 0); w1.assign(ny + 2, 0);
    for (int j = 1; j <= ny; j++) {
        w0[j] = (yg[j] - y[j - 1]) / (yg[j] - yg[j - 1]);
        w1[j] = 1.0 - w0[j];
    }
    // initialization.f90:281-296
    const double pi = 4.0 * std::atan(1.0);
    const int nxp = d.nxp, nzp = d.nzp;
    kxx.assign(nxp, 0); kzz.assign(nzp, 0);
    const int cx = (nxp + 1) / 2, cz = (nzp + 1) / 2;
    for (int i = 0; i < nxp; i++) {
        const double a = (i <= cx) ? (double)i : (double)(-nxp + i);
        kxx[i] = 2.0 * (std::cos(2.0 * pi * a / (double)nxp) - 1.0) / (dxmin * dxmin);
    }
    for (int k = 0; k < nzp; k++) {
        const double a = (k <= cz) ? (double)k : (double)(-nzp + k);
        kzz[k] = 2.0 * (std::cos(2.0 * pi * a / (double)nzp) - 1.0) / (dzmin * dzmin);
    }
    // initialization.f90:343-373
    const int nr = d.nyp;
    dd.assign(nr, 0); dl.assign(std::max(nr - 1, 1), 0); du.assign(std::max(nr - 1, 1), 0);
    auto Yf = [&](int j) { return y[j - 1]; };
    auto YG = [&](int j) { return yg[j - 1]; };
    for (int j = 2; j <= nyg - 1; j++) {
        const double a = 1.0 / (Yf(j) - Yf(j - 1)) / (YG(j + 1) - YG(j));
        const double b = 1.0 / (Yf(j) - Yf(j - 1)) * (-1.0 / (YG(j + 1) - YG(j)) - 1.0 / (YG(j) - YG(j - 1)));
        const double c = 1.0 / (Yf(j) - Yf(j - 1)) / (YG(j) - YG(j - 1));
        dd[j - 2] = b;
        if (j < nyg - 1) du[j - 2] = a;
        if (j > 2) dl[j - 3] = c;
        if (j == 2) dd[0] = b + c;
        if (j == nyg - 1) dd[j - 2] = a + b;
    }
    // mass_flow.f90:35-44 normalisation, accumulated in the reference's order
    norm_u = 0.0;
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 3; j <= nyg - 1; j++)
            for (int i = 2; i <= nx - 1; i++) norm_u = norm_u + (1.0 + 1.0) * 0.5 * (YG(j) - YG(j - 1));
    norm_v = 0.0;
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= ny; j++)
            for (int i = 2; i <= nxg - 1; i++) norm_v = norm_v + (1.0 + 1.0) * 0.5 * (Yf(j) - Yf(j - 1));
    // time_integration.f90:208-212
    dt_vis = std::min(std::min(p.cfl * dxmin * dxmin / p.nu, p.cfl * dymin * dymin / p.nu),
                      p.cfl * dzmin * dzmin / p.nu);
    // subgrid.f90:600 filter width
    delta.assign(nyg + 2, 0);
    for (int j = 2; j <= nyg - 1; j++) delta[j] = std::pow(dxmin * dzmin * (Yf(j) - Yf(j - 1)), 1.0 / 3.0);
    // wallmodel.f90:605-610 stencils
    jjz.assign(nzg, 0); wz.assign(3 * nzg, 0); jjx.assign(nx, 0); wx.assign(3 * nx, 0);
    interp_2nd(nz, z.data(), nzg, zg.data(), false, jjz.data(), wz.data());
    interp_2nd(nxg, xg.data(), nx, x.data(), true, jjx.data(), wx.data());
}

std::string Host::validate() const {
    const cf_params &p = prm;
    if (p.nx < 6 || p.ny < 4 || p.nz < 6) return "grid too small";
    if (p.nx % 2 || p.nz % 2) return "nx and nz must be even (initialization.f90:54-55)";
    if ((int64_t)d.nxg * d.nyg * d.nzg > (int64_t)2000000000) return "grid too large";
    if (d.nxp / 2 > 1024 || d.nzp > 2048) return "FFT length too large";
    {
        std::string ferr;
        make_plan(d.nxp / 2, ferr, FFT_X_RADIX8);
        make_plan(d.nzp, ferr, true);
        if (!ferr.empty()) return ferr;
    }
    if (!thomas_supported(d)) return "too many y cells for the y solve";
    if (p.precision != 0 && p.precision != 1) return "precision must be 0 (fp64) or 1 (mixed)";
    if (p.les_model < 0 || p.les_model > 2) return "les_model must be 0, 1 or 2";
    if (p.iwall_model != 0 && p.iwall_model != 3) return "iwall_model must be 0 or 3";
    if (p.iwall_model == 3) {
        if (p.jref_eqwm < 2 || p.jref_eqwm > d.nyg - 1) return "jref_eqwm out of range";
        if (p.eqwm_ncv < 3 || p.eqwm_ncv > EQWM_MAXNCV) return "eqwm_ncv out of range";
    }
    if (p.nu <= 0) return "nu must be positive";
    if (p.x_mass_cte < 0 || p.x_mass_cte > 2)
        return "x_mass_cte must be 0, 1 (constant bulk flow) or 2 (constant centreline velocity)";
    if (p.y_mass_cte != 0 && p.y_mass_cte != 1) return "y_mass_cte must be 0 or 1 (constant bulk flow in y)";
    return "";
}

// ---------------------------------------------------------------------------
// Device context
// ---------------------------------------------------------------------------
// Device buffers come from hipMalloc, or from a caller-provided slab (a
// region of a serving arena: bump allocation, no per-case hipMalloc/hipFree)
// or from a dry slab that only counts the footprint.
template <typename T>
static int dalloc(T **p, size_t n, std::vector<void *> &owned, size_t &bytes, Slab *slab) {
    const size_t b = std::max<size_t>(n, 1) * sizeof(T);
    if (slab) {
        const size_t ba = (b + 255) & ~(size_t)255;
        if (!slab->dry && slab->used + ba > slab->cap) {
            set_error("device slab exhausted (footprint larger than reserved)");
            return -1;
        }
        *p = slab->dry ? nullptr : (T *)(slab->base + slab->used);
        slab->used += ba;
        bytes += ba;
        return 0;
    }
    void *q = nullptr;
    HCK(hipMalloc(&q, b));
    owned.push_back(q);
    bytes += b;
    *p = (T *)q;
    return 0;
}
template <typename T>
static int upload(T *d, const std::vector<T> &h) {
    HCK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

template <typename S>
Ctx<S>::~Ctx() {
    if (dev >= 0) hipSetDevice(dev);
    for (auto &g : graphs) if (g) hipGraphExecDestroy(g);
    for (auto &e : evpool) hipEventDestroy(e);
    for (void *p : owned) hipFree(p);
    if (stream && !stream_ext) hipStreamDestroy(stream);
}

template <typename S>
int Ctx<S>::init() {
    const Dims &d = h.d;
    const bool dry = slab && slab->dry;
    // ---- phase 1: every device buffer (a dry slab stops after this) ----
    const size_t nU = (size_t)d.nx * d.nyg * d.nzg, nV = (size_t)d.nxg * d.ny * d.nzg,
                 nW = (size_t)d.nxg * d.nyg * d.nz, nC = (size_t)d.nxg * d.nyg * d.nzg;
    n3[0] = nU; n3[1] = nV; n3[2] = nW;
    const size_t nspec = (size_t)d.mh * d.nzp * d.nyp;
    const size_t nphi = (size_t)d.nxp * d.nyp * d.nzp;
    const int npart = std::max(std::max(3 * d.nzg, 2 * d.nzg), massflow_parts(d)) + 16;
    const size_t nwx = (size_t)d.nx * 2 * d.nzg, nwz = (size_t)d.nxg * 2 * d.nz;
    const int M = d.nxp / 2;
    auto A = [&](auto **p, size_t n) { return dalloc(p, n, owned, bytes, slab); };
    for (int b = 0; b < 2; b++)
        for (int c = 0; c < 3; c++)
            if (A(&st[b][c], n3[c])) return -1;
    for (int c = 0; c < 3; c++)
        if (A(&G[c], n3[c])) return -1;
    if (ysolve_uses_scratch(d) && A(&cscr, ysolve_scratch_elems(d))) return -1;
    if (A(&mfpart, rhs_rk_tiles(d, true))) return -1;
    if (A(&dtfp, (size_t)ifftx_blocks(d, sizeof(S)) + 2 * (size_t)project_vw_blocks(d, sizeof(S)) +
                     dt_fused_mid_doubles()))
        return -1;
    if (A(&NT, nC) || A(&spec, nspec) || A(&phi1, nphi) || A(&phi23, nphi) ||
        A(&part, npart) || A(&dtscr, dt_scratch_doubles(d)) || A(&sc, 1) ||
        (prec_mixed() && A(&ubar, d.nyg + 2)) || A(&stats_out, stats_out_doubles(d)) || A(&stats_acc, 7 * d.ny + 8))
        return -1;
    if (A(&m_rdxf, h.rdxf.size()) || A(&m_rdxc, h.rdxc.size()) || A(&m_rdyf, h.rdyf.size()) ||
        A(&m_rdyc, h.rdyc.size()) || A(&m_rdzf, h.rdzf.size()) || A(&m_rdzc, h.rdzc.size()) ||
        A(&m_w0, h.w0.size()) || A(&m_w1, h.w1.size()) || A(&m_dyg, h.dyg.size()) ||
        A(&m_dxc, h.dxc.size()) || A(&m_dyc, h.dyc.size()) || A(&m_dzc, h.dzc.size()) ||
        A(&m_x, h.x1.size()) || A(&m_y, h.y1.size()) || A(&m_z, h.z1.size()) || A(&m_xg, h.xg1.size()) ||
        A(&m_yg, h.yg1.size()) || A(&m_zg, h.zg1.size()) || A(&m_kxx, h.kxx.size()) ||
        A(&m_kzz, h.kzz.size()) || A(&m_dd, h.dd.size()) || A(&m_dl, h.dl.size()) || A(&m_du, h.du.size()) ||
        A(&m_delta, h.delta.size()) || A(&m_jjz, h.jjz.size()) || A(&m_wz, h.wz.size()) ||
        A(&m_jjx, h.jjx.size()) || A(&m_wx, h.wx.size()))
        return -1;
    if (A(&twM, 2 * (size_t)M + 2) || A(&twN, (size_t)d.nxp + 1) || A(&twZ, 2 * (size_t)d.nzp + 2)) return -1;
    if (h.prm.les_model == 2 && (A(&dsm.Q, 10 * nC) || A(&dsm.Fu, nU + nV + nW) ||
                                 A(&dsm.part, dsm_part_elems(d)) || A(&dsm.avg, (size_t)d.nyg)))
        return -1;
    if (A(&alpha_x, nwx) || A(&alpha_z, nwz) || A(&azt, nwx) || A(&tau_init, nwx)) return -1;
    if (dry) return 0;

    // ---- phase 2: stream, contents ----
    dev = h.prm.device;
    HCK(hipSetDevice(dev));
    if (!stream) HCK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    else stream_ext = true;
    HCK(hipMemsetAsync(NT, 0, nC * sizeof(S), stream));
    if (dsm.avg) HCK(hipMemsetAsync(dsm.avg, 0, d.nyg * sizeof(double), stream));
    for (int b = 0; b < 2; b++)
        for (int c = 0; c < 3; c++) HCK(hipMemsetAsync(st[b][c], 0, n3[c] * sizeof(S), stream));
    HCK(hipMemsetAsync(phi1, 0, nphi * sizeof(S), stream));
    HCK(hipMemsetAsync(stats_acc, 0, (7 * d.ny + 8) * sizeof(double), stream));
    if (ubar) HCK(hipMemsetAsync(ubar, 0, (d.nyg + 2) * sizeof(double), stream));
    if (upload(m_rdxf, h.rdxf) || upload(m_rdxc, h.rdxc) || upload(m_rdyf, h.rdyf) || upload(m_rdyc, h.rdyc) ||
        upload(m_rdzf, h.rdzf) || upload(m_rdzc, h.rdzc) || upload(m_w0, h.w0) || upload(m_w1, h.w1) ||
        upload(m_dyg, h.dyg) || upload(m_dxc, h.dxc) || upload(m_dyc, h.dyc) || upload(m_dzc, h.dzc) ||
        upload(m_x, h.x1) || upload(m_y, h.y1) || upload(m_z, h.z1) || upload(m_xg, h.xg1) ||
        upload(m_yg, h.yg1) || upload(m_zg, h.zg1) || upload(m_kxx, h.kxx) || upload(m_kzz, h.kzz) ||
        upload(m_dd, h.dd) || upload(m_dl, h.dl) || upload(m_du, h.du) || upload(m_delta, h.delta) ||
        upload(m_jjz, h.jjz) || upload(m_wz, h.wz) || upload(m_jjx, h.jjx) || upload(m_wx, h.wx))
        return -1;
    met.rdxf = m_rdxf; met.rdxc = m_rdxc; met.rdyf = m_rdyf; met.rdyc = m_rdyc;
    met.rdzf = m_rdzf; met.rdzc = m_rdzc; met.w0 = m_w0; met.w1 = m_w1; met.dyg = m_dyg;
    met.dxc = m_dxc; met.dyc = m_dyc; met.dzc = m_dzc;
    wt.jjz = m_jjz; wt.wz = m_wz; wt.jjx = m_jjx; wt.wx = m_wx;
    // FFT plans and twiddles: x uses a length-M complex FFT plus the N-point split
    std::string err;
    pM = make_plan(M, err, FFT_X_RADIX8);
    pZ = make_plan(d.nzp, err, true);
    if (!err.empty()) { set_error(err); return -1; }
    auto twid = [](int n) {
        std::vector<Cx<S>> t(n + 1);
        for (int q = 0; q <= n; q++) {
            long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)n;
            t[q] = Cx<S>{(S)std::cos(a), (S)(-std::sin(a))};
        }
        return t;
    };
    // twiddles of the LDS transforms: the plain table exp(-2 pi i q/n),
    // q = 0..n, then per-pass tables: for the pass with current length Ns and
    // radix R, exp(-2 pi i r k / (Ns R)) at Ns-1 + k (R-1) + r-1 (k < Ns,
    // 1 <= r < R); n-1 entries in all
    auto pass_tw = [&](const FFTPlan &p) {
        std::vector<Cx<S>> t = twid(p.n);  // plain table, then the pass table
        t.resize(2 * (size_t)p.n + 2, Cx<S>{(S)1, (S)0});
        const int o = p.n + 1;
        int Ns = 1;
        for (int s = 0; s < p.nr; s++) {
            const int R = p.radix[s];
            for (int k = 0; k < Ns; k++)
                for (int r = 1; r < R; r++) {
                    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)(r * k) /
                                    (long double)(Ns * R);
                    t[o + Ns - 1 + k * (R - 1) + (r - 1)] = Cx<S>{(S)std::cos(a), (S)(-std::sin(a))};
                }
            Ns *= R;
        }
        return t;
    };
    if (upload(twM, pass_tw(pM)) || upload(twN, twid(d.nxp)) || upload(twZ, pass_tw(pZ))) return -1;


    // wall model buffers
    HCK(hipMemsetAsync(alpha_x, 0, nwx * sizeof(double), stream));
    HCK(hipMemsetAsync(alpha_z, 0, nwz * sizeof(double), stream));
    HCK(hipMemsetAsync(azt, 0, nwx * sizeof(double), stream));
    if (upload(tau_init, std::vector<double>(nwx, h.prm.tau_mean_eqwm))) return -1;
    // scalars
    Scalars s0{};
    s0.dt = h.prm.cfl < 0 ? -h.prm.cfl : 0.0;
    s0.dPdx = h.prm.dPdx;
    s0.dPdz = h.prm.dPdz;
    s0.norm_u = h.norm_u;
    s0.norm_v = h.norm_v;
    s0.ucl = h.prm.u_centerline;
    s0.nu = h.prm.nu;
    s0.dt_vis = h.dt_vis;
    s0.tau_mean = h.prm.tau_mean_eqwm;
    HCK(hipMemcpy(sc, &s0, sizeof(Scalars), hipMemcpyHostToDevice));
    HCK(hipStreamSynchronize(stream));
    return 0;
}

// ---- timing helpers ----
template <typename S>
void Ctx<S>::tic(const char *name) {
    if (!timing) return;
    hipEvent_t a = ev_get(), b = ev_get();
    hipEventRecord(a, stream);
    pending.push_back({name, a, b});
}
template <typename S>
void Ctx<S>::toc() {
    if (!timing || pending.empty()) return;
    hipEventRecord(pending.back().b, stream);
    done_ev.push_back(pending.back());
    pending.pop_back();
}
template <typename S>
hipEvent_t Ctx<S>::ev_get() {
    if (ev_free.empty()) {
        hipEvent_t e;
        hipEventCreate(&e);
        evpool.push_back(e);
        return e;
    }
    hipEvent_t e = ev_free.back();
    ev_free.pop_back();
    return e;
}
template <typename S>
void Ctx<S>::collect() {
    if (done_ev.empty()) return;
    hipStreamSynchronize(stream);
    for (auto &r : done_ev) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, r.a, r.b);
        auto &acc = tstat[r.name];
        acc.first += ms;
        acc.second += 1;
        ev_free.push_back(r.a);
        ev_free.push_back(r.b);
    }
    done_ev.clear();
}

// ---- building blocks ----
template <typename S>
void Ctx<S>::ghosts(S *const *F, int mask) {
    const Dims &d = h.d;
    const bool wm = h.prm.iwall_model == 3;
    const double dgb = h.yg[1] - h.yg[0], dgt = h.yg[d.nyg - 1] - h.yg[d.nyg - 2];
    S *fs[3];
    int fld[3], md[3], n = 0;
    const double *al[3];
    if (mask & 1) { fs[n] = F[0]; fld[n] = 0; md[n] = wm ? 1 : 0; al[n] = alpha_x; n++; }
    if (mask & 2) { fs[n] = F[1]; fld[n] = 1; md[n] = 0; al[n] = nullptr; n++; }
    if (mask & 4) { fs[n] = F[2]; fld[n] = 2; md[n] = wm ? 1 : 0; al[n] = alpha_z; n++; }
    tic("ghosts");
    launch_ghosts3<S>(d, fs, fld, md, al, n, dgb, dgt, stream);
    toc();
}

template <typename S>
void Ctx<S>::poisson(S *const *F, S *phi, const Scalars *corr, const DtParts *dtp, int gcopy) {
    const Dims &d = h.d;
    tic("div_fftx");
    launch_div_fftx<S, S>(d, met, F[0], F[1], F[2], spec, pM, twM, twN, stream);
    toc();
    if (cscr && zsolve_fused(d)) {  // the scratch is sized for either form when allocated
        tic("zsolve");
        launch_zsolve<S>(d, spec, cscr, pZ, twZ, m_kxx, m_kzz, m_dd, m_dl, m_du, stream);
        toc();
    } else {
        tic("fftz_fwd");
        launch_fftz<S>(d, spec, pZ, twZ, -1, stream);
        toc();
        tic("thomas");
        launch_thomas<S>(d, spec, cscr, m_kxx, m_kzz, m_dd, m_dl, m_du, stream);
        toc();
        tic("fftz_inv");
        launch_fftz<S>(d, spec, pZ, twZ, +1, stream);
        toc();
    }
    tic("ifftx_proju");
    launch_ifftx_proju<S, S, S>(d, met, spec, phi, F[0], pM, twM, twN, (double)(d.nxp * d.nzp), stream, corr, dtp,
                                gcopy);
    toc();
    tic("project_vw");
    launch_project_vw<S, S>(d, met, phi, F[1], F[2], stream, dtp ? dtp->pv : nullptr, dtp ? dtp->pw : nullptr, gcopy);
    toc();
}

template <typename S>
void Ctx<S>::eddy_viscosity(S *const *F) {
    const Dims &d = h.d;
    if (h.prm.les_model == 0) return;  // nu_t stays 0 (subgrid.f90:41-45)
    if (h.prm.les_model == 2) {
        tic("dynamic_smagorinsky");
        launch_dsm<S>(d, met, F[0], F[1], F[2], ubar, dsm, NT, h.prm.dirichlet_nu_t, stream);
        toc();
        return;
    }
    tic("smagorinsky");
    launch_smagorinsky<S>(d, met, m_y, m_yg, m_delta, F[0], F[1], F[2], NT, sc, ubar, h.prm.cs, h.prm.nu,
                          h.prm.dirichlet_nu_t, stream);
    toc();
    tic("ghosts_nut");
    launch_ghosts<S>(d, NT, 3, h.prm.dirichlet_nu_t == 1 ? 0 : 2, nullptr, 0, 0, stream);
    toc();
}

template <typename S>
void Ctx<S>::wall_model(S *const *F) {
    const Dims &d = h.d;
    if (h.prm.iwall_model != 3) return;
    EqwmParams ep{h.prm.eqwm_ncv, h.prm.eqwm_max_iter, h.prm.eqwm_stretch, h.prm.eqwm_tol, 1.0};
    const double hw = h.yg[h.prm.jref_eqwm - 1];
    const double *pred = nullptr;
    const int *use = nullptr;
    tic("reset_counters");
    launch_reset_counters(sc, stream);
    toc();
    if (surrogate) surrogate(this, F, &pred, &use);
    tic("wall_model");
    launch_wall_model<S>(d, F[0], F[2], h.prm.les_model != 0 ? NT : nullptr, wt, tau_init, alpha_x, azt,
                         alpha_z, sc, ep, h.prm.nu, hw, h.prm.jref_eqwm, h.prm.tau_mean_eqwm, pred, use,
                         ubar, stream);
    toc();
}

template <typename S>
void Ctx<S>::dt_launch() {
    const Dims &d = h.d;
    tic("dt");
    if (h.prm.cfl < 0) launch_dt_final_only(sc, h.prm.cfl, stream);
    else
        launch_dt<S>(d, m_dxc, m_dyc, m_dzc, st[cur][0], st[cur][1], st[cur][2], ubar, dtscr, sc,
                     h.prm.cfl, h.dt_vis, stream);
    toc();
}

template <typename S>
void Ctx<S>::step_launch() {
    const Dims &d = h.d;
    const double rk[3][3] = {{8.0 / 15.0, 0, 0}, {1.0 / 4.0, 5.0 / 12.0, 0}, {1.0 / 4.0, 0.0, 3.0 / 4.0}};
    const bool nut = h.prm.les_model != 0 || nut_user;
    // the x mass-flow correction (time_integration.f90:152-163) is taken from
    // the third substep's U* as k_rhs_rk writes it (the projection leaves the
    // bulk and centreline means unchanged), and added by that substep's U
    // projection; the mixed path shifts its profile instead
    const int mfm = (h.prm.x_mass_cte == 1 || h.prm.x_mass_cte == 2) ? h.prm.x_mass_cte : 0;
    // compute_dt of the next step from the final velocities as the last
    // projection writes them (the mixed path's U with its shifted profile)
    const bool fz = fuse_dt();
    // ghost copies by the projection kernels (FP64, Dirichlet walls; in the
    // mixed path the extra registers cost the FP32 V/W projection a wave per
    // SIMD, more than the ghost pass); env CF_GHOST_COPY=0 selects the ghost
    // pass (read at graph capture)
    const char *gce = getenv("CF_GHOST_COPY");
    const bool gcp = sizeof(S) == 8 && h.prm.iwall_model != 3 && (!gce || atoi(gce) != 0);
    const int ndu = ifftx_blocks(d, sizeof(S)), ndvw = project_vw_blocks(d, sizeof(S));
    const DtParts dp{dtfp, dtfp + ndu, dtfp + ndu + ndvw, ubar, sc, (ubar && mfm) ? 1 : 0};
    for (int s = 1; s <= 3; s++) {
        S *const *Fc = st[cur];
        S *const *Fn = st[1 - cur];
        eddy_viscosity(Fc);
        if (s == 1) wall_model(Fc);
        double *mfp = (s == 3 && mfm) ? mfpart : nullptr;
        tic("rhs_rk");
        launch_rhs_rk<S, S>(d, met, Fc[0], Fc[1], Fc[2], nut ? NT : nullptr, Fn[0], Fn[1], Fn[2], G[0], G[1],
                            G[2], sc, ubar, h.prm.nu, h.prm.omega_z, rk[s - 1][s - 1], rk[1][0],
                            s == 1 ? 1 : 2, stream, mfp, mfm);
        toc();
        if (mfp) {
            tic("massflow");
            launch_massflow_fused_final(d, mfpart, rhs_rk_tiles(d, nut), mfm, ubar, sc, stream);
            toc();
        }
        // the divergence reads the x and z ghost faces at their periodic
        // images, so the boundary conditions are applied once, after the
        // projection (time_integration.f90 applies them before and after);
        // with Dirichlet walls the projection kernels write the ghosts
        // themselves, with the wall model's Neumann walls a ghost pass does
        poisson(Fn, s == 1 ? phi1 : phi23, (mfp && !ubar) ? sc : nullptr, (s == 3 && fz) ? &dp : nullptr,
                gcp ? 1 : 0);
        if (!gcp) ghosts(Fn, 7);
        cur = 1 - cur;
    }
    S *const *F = st[cur];
    if (mfm && ubar) {
        // deviation storage: the uniform correction shifts the profile
        tic("massflow");
        launch_ubar_shift(sc, ubar, d.nyg, h.prm.iwall_model == 3 ? 1 : 0, stream);
        toc();
    }
    if (h.prm.y_mass_cte == 1) {
        tic("massflow_v");
        launch_massflow_v<S>(d, m_y, F[1], part, sc, 1, stream);
        launch_add_corr_v<S>(d, F[1], sc, stream);
        toc();
        ghosts(F, 2);
    }
    tic("advance_time");
    launch_advance_time(sc, stream);
    toc();
    if (fz) {
        tic("dt_fused");
        launch_dt_fused_final<S>(d, dp, ndu, ndvw, m_dyc, st[cur][1], sc, h.prm.cfl, h.dt_vis, stream);
        toc();
    }
}

// main.f90:70-73: compute_dt, then the step.  With the fused dt the step
// ends by computing the next step's dt into dt_next, and a step starts by
// taking it; a standalone compute_dt runs first whenever the state changed
// since the last step (dt_fresh false).
template <typename S>
void Ctx<S>::step_head() {
    if (fuse_dt()) launch_dt_take_next(sc, stream);
    else dt_launch();
}

template <typename S>
int Ctx<S>::run_steps(int nsteps, bool use_graph) {
    HCK(hipSetDevice(dev));
    // the launch-error state is per thread: an error left by an earlier,
    // already reported call on this thread (another context's) is not this
    // context's, so it is cleared before this call's launches
    (void)hipGetLastError();
    if (batch_member && !batch_running) {
        set_error("a batch member is stepped through its batch (cf_batch_run_steps)");
        return -1;
    }
    for (int n = 0; n < nsteps; n++) {
        if (fuse_dt() && !dt_fresh) dt_launch();
        if (use_graph && !timing) {
            hipGraphExec_t &ge = graphs[cur];
            const int start = cur;
            if (!ge) {
                hipGraph_t g;
                BatchScope bsc(bn, bcs);
                HCK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                step_head();
                step_launch();
                HCK(hipStreamEndCapture(stream, &g));
                HCK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                HCK(hipGraphDestroy(g));
                cur = start;  // capture does not execute; replay advances the state
            }
            HCK(hipGraphLaunch(ge, stream));
            cur = 1 - start;  // three substeps flip the ping-pong parity
        } else {
            BatchScope bsc(bn, bcs);
            step_head();
            step_launch();
        }
        dt_fresh = fuse_dt();
    }
    HCK(hipGetLastError());
    return 0;
}

// capture the step graphs of both ping-pong parities without running them
template <typename S>
int Ctx<S>::prepare_graphs() {
    HCK(hipSetDevice(dev));
    if (batch_member) { set_error("a batch member is stepped through its batch"); return -1; }
    const int keep = cur;
    for (int p = 0; p < 2; p++) {
        if (graphs[p]) continue;
        cur = p;
        hipGraph_t g;
        HCK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        step_head();
        step_launch();
        HCK(hipStreamEndCapture(stream, &g));
        HCK(hipGraphInstantiate(&graphs[p], g, nullptr, nullptr, 0));
        HCK(hipGraphDestroy(g));
    }
    cur = keep;
    return 0;
}

// compute_time_step_RK3 with the dt the last compute_dt set, as the
// reference's step uses the module's dt
// Step a batch (this context first, then the other members, all of one
// size class with their slabs at the byte stride cs): a member whose state
// changed since its last step takes its own compute_dt first, on the shared
// stream, then each kernel of the step runs once for every member.  The
// members keep their ping-pong parity and fresh-dt flag in step with this one.
template <typename S>
int Ctx<S>::batch_run(const std::vector<CtxBase *> &members, int64_t cs, int nsteps, bool graph) {
    HCK(hipSetDevice(dev));
    const int n = (int)members.size();
    if (n < 1 || members[0] != this) { set_error("batch_run: the first member leads the batch"); return -1; }
    if (bn != n || bcs != cs) {
        invalidate_graphs();
        bn = n;
        bcs = cs;
    }
    std::vector<Ctx<S> *> mem;
    for (CtxBase *b : members) mem.push_back(static_cast<Ctx<S> *>(b));
    for (int q = 0; q < nsteps; q++) {
        for (Ctx<S> *c : mem) {
            if (c->cur != cur || c->stream != stream) { set_error("batch_run: members out of step"); return -1; }
            if (c->fuse_dt() && !c->dt_fresh) {
                c->dt_launch();
                c->dt_fresh = true;
            }
        }
        batch_running = true;
        const int rc = run_steps(1, graph);
        batch_running = false;
        if (rc) return -1;
        for (Ctx<S> *c : mem) {
            c->cur = cur;
            c->dt_fresh = dt_fresh;
        }
    }
    return 0;
}

template <typename S>
int Ctx<S>::rk3_step() {
    HCK(hipSetDevice(dev));
    if (batch_member) { set_error("a batch member is stepped through its batch (cf_batch_run_steps)"); return -1; }
    step_launch();
    HCK(hipGetLastError());
    dt_fresh = fuse_dt();
    return 0;
}

template <typename S>
int Ctx<S>::compute_dt(double *dt) {
    HCK(hipSetDevice(dev));
    dt_launch();
    dt_fresh = true;
    HCK(hipGetLastError());
    Scalars s;
    HCK(hipMemcpyAsync(&s, sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream));
    HCK(hipStreamSynchronize(stream));
    if (dt) *dt = s.dt;
    return 0;
}

template <typename S>
int Ctx<S>::projection_step(int rk_step) {
    HCK(hipSetDevice(dev));
    dt_fresh = false;
    poisson(st[cur], rk_step == 1 ? phi1 : phi23);
    HCK(hipGetLastError());
    HCK(hipStreamSynchronize(stream));
    return 0;
}

template <typename S>
int Ctx<S>::apply_bc() {
    HCK(hipSetDevice(dev));
    dt_fresh = false;
    ghosts(st[cur], 7);
    HCK(hipGetLastError());
    HCK(hipStreamSynchronize(stream));
    return 0;
}

// host double <-> device S
template <typename S>
int Ctx<S>::put(S *dst, const double *src, size_t n) {
    if (std::is_same<S, double>::value) {
        HCK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<S> t(n);
        for (size_t q = 0; q < n; q++) t[q] = (S)src[q];
        HCK(hipMemcpy(dst, t.data(), n * sizeof(S), hipMemcpyHostToDevice));
    }
    return 0;
}
template <typename S>
int Ctx<S>::get(double *dst, const S *src, size_t n) {
    HCK(hipStreamSynchronize(stream));
    if (std::is_same<S, double>::value) {
        HCK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<S> t(n);
        HCK(hipMemcpy(t.data(), src, n * sizeof(S), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; q++) dst[q] = (double)t[q];
    }
    return 0;
}

// Mixed path: U is stored as an FP32 deviation from an FP64 profile ubar(j),
// the plane mean of the incoming U on the interior rows, with ghost rows set
// by the wall condition (so the ghost formulas hold for the deviation too).
template <typename S>
int Ctx<S>::set_profile(const double *U) {
    const Dims &d = h.d;
    std::vector<double> ub(d.nyg + 2, 0.0);
    for (int j = 2; j <= d.nyg - 1; j++) {
        double acc = 0.0;
        for (int k = 2; k <= d.nzg - 2; k++)
            for (int i = 2; i <= d.nx - 1; i++) acc += U[i3(d.nx, d.nyg, i, j, k)];
        ub[j] = acc / ((double)(d.nx - 2) * (d.nzg - 3));
    }
    const bool neu = h.prm.iwall_model == 3;
    ub[1] = neu ? ub[2] : -ub[2];
    ub[d.nyg] = neu ? ub[d.nyg - 1] : -ub[d.nyg - 1];
    HCK(hipMemcpy(ubar, ub.data(), ub.size() * sizeof(double), hipMemcpyHostToDevice));
    // bulk-flow contribution of the profile, in the order of mass_flow.f90:37-44
    double q = 0.0;
    for (int j = 3; j <= d.nyg - 1; j++) q += (ub[j] + ub[j - 1]) * 0.5 * h.dyg[j];
    Scalars s;
    HCK(hipMemcpy(&s, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    s.qbar_u = q * (double)(d.nx - 2) * (double)(d.nzg - 2);
    HCK(hipMemcpy(sc, &s, sizeof(Scalars), hipMemcpyHostToDevice));
    ubar_h = ub;
    return 0;
}

template <typename S>
int Ctx<S>::sync_profile() {
    if (!ubar) return 0;
    HCK(hipStreamSynchronize(stream));
    ubar_h.resize(h.d.nyg + 2);
    HCK(hipMemcpy(ubar_h.data(), ubar, ubar_h.size() * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

template <typename S>
int Ctx<S>::set_state(const double *U, const double *V, const double *W) {
    HCK(hipSetDevice(dev));
    dt_fresh = false;
    HCK(hipStreamSynchronize(stream));
    if (ubar) {
        if (set_profile(U)) return -1;
        const Dims &d = h.d;
        std::vector<double> Ud(U, U + n3[0]);
        for (int k = 1; k <= d.nzg; k++)
            for (int j = 1; j <= d.nyg; j++)
                for (int i = 1; i <= d.nx; i++) Ud[i3(d.nx, d.nyg, i, j, k)] -= ubar_h[j];
        if (put(st[cur][0], Ud.data(), n3[0])) return -1;
    } else if (put(st[cur][0], U, n3[0])) {
        return -1;
    }
    if (put(st[cur][1], V, n3[1]) || put(st[cur][2], W, n3[2])) return -1;
    invalidate_graphs();
    return 0;
}
template <typename S>
int Ctx<S>::get_state(double *U, double *V, double *W) {
    HCK(hipSetDevice(dev));
    if ((U && get(U, st[cur][0], n3[0])) || (V && get(V, st[cur][1], n3[1])) ||
        (W && get(W, st[cur][2], n3[2])))
        return -1;
    if (U && ubar) {
        if (sync_profile()) return -1;
        const Dims &d = h.d;
        for (int k = 1; k <= d.nzg; k++)
            for (int j = 1; j <= d.nyg; j++)
                for (int i = 1; i <= d.nx; i++) U[i3(d.nx, d.nyg, i, j, k)] += ubar_h[j];
    }
    return 0;
}

// projection.f90:129-137: P = phi/(dt*rk_coef(1,1)) with ghost copies
template <typename S>
int Ctx<S>::get_pressure(double *P) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    std::vector<double> ph((size_t)d.nxp * d.nyp * d.nzp);
    if (get(ph.data(), phi1, ph.size())) return -1;
    Scalars s;
    HCK(hipMemcpy(&s, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    const double den = s.dt * (8.0 / 15.0);
    const int nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    auto idx = [&](int i, int j, int k) { return i3(nxg, nyg, i, j, k); };
    auto phv = [&](int i, int j, int k) {  // interior incl. duplicate planes
        int ii = i == nxg - 1 ? 2 : i, kk = k == nzg - 1 ? 2 : k;
        return ph[((size_t)(kk - 2) * d.nyp + (j - 2)) * d.nxp + (ii - 2)];
    };
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) P[idx(i, j, k)] = phv(i, j, k) / den;
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) { P[idx(1, j, k)] = P[idx(2, j, k)]; P[idx(nxg, j, k)] = P[idx(nxg - 1, j, k)]; }
    for (int k = 1; k <= nzg; k++)
        for (int i = 1; i <= nxg; i++) { P[idx(i, 1, k)] = P[idx(i, 2, k)]; P[idx(i, nyg, k)] = P[idx(i, nyg - 1, k)]; }
    const size_t pl = (size_t)nxg * nyg;
    std::memcpy(P, P + pl, pl * sizeof(double));
    std::memcpy(P + pl * (nzg - 1), P + pl * (nzg - 2), pl * sizeof(double));
    return 0;
}

// compute_pressure (pressure.f90:21-341): the actual pressure of the current
// state.  The right-hand sides of the momentum equations and their
// divergence, the wall conditions dP/dy = d/dy(nu dV/dy) from one-sided
// differences of V and the plane-averaged eddy viscosity, then the same
// transforms and y solve as the projection, with the wall conditions as
// sources of the first and last rows.  Ghosts: rows j = 1, nyg from the wall
// conditions, duplicate planes by the periodic copies, and the ghost planes
// the one-rank reference leaves unset (i = 1, nxg; k = 1, nzg) as copies of
// their neighbours.  Output-only: temporaries are allocated per call.
template <typename S>
int Ctx<S>::compute_pressure(double *P) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    const int nx = d.nx, ny = d.ny, nz = d.nz, nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    const size_t nU = (size_t)nx * nyg * nzg, nV = (size_t)nxg * ny * nzg, nW = (size_t)nxg * nyg * nz;
    const size_t nb = (size_t)(nxg - 2) * (nzg - 2), npl = (size_t)d.nxp * d.nzp;
    const size_t ncell = npl * d.nyp;
    const size_t ntot = nU + nV + nW + 2 * nb + 2 * npl + ncell + 2;
    double *buf = nullptr;
    HCK(hipMalloc(&buf, ntot * sizeof(double)));
    double *Ru = buf, *Rv = Ru + nU, *Rw = Rv + nV, *bcs = Rw + nW, *ysrc = bcs + 2 * nb, *pcell = ysrc + 2 * npl,
           *means = pcell + ncell;
    hipError_t e = hipMemsetAsync(buf, 0, ntot * sizeof(double), stream);
    S *const *F = st[cur];
    // avg_nu_t at rows 1, 2, nyg-1, nyg (subgrid.f90:605-617)
    double a1 = 0, a2 = 0, an1 = 0, an = 0;
    const bool nut = h.prm.les_model != 0 || nut_user;
    if (e == hipSuccess && h.prm.les_model == 2) {
        // the dynamic model's avg_nu_t is its plane coefficient <LM>/<MM>
        // of the last evaluation (subgrid.f90:505-508, 529-535)
        std::vector<double> av(nyg);
        e = hipMemcpyAsync(av.data(), dsm.avg, nyg * sizeof(double), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        a1 = av[0];
        a2 = av[1];
        an1 = av[nyg - 2];
        an = av[nyg - 1];
    } else if (e == hipSuccess && nut) {
        launch_plane_means<S>(d, NT, 2, nyg - 1, means, stream);
        double m2[2];
        e = hipMemcpyAsync(m2, means, sizeof m2, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        a2 = std::max(m2[0], 0.0);
        an1 = std::max(m2[1], 0.0);
        a1 = h.prm.dirichlet_nu_t == 1 ? -a2 : a2;
        an = h.prm.dirichlet_nu_t == 1 ? -an1 : an1;
    }
    // one-sided metrics at the walls (pressure.f90:263-266, 302-305), dzeta = 1
    const std::vector<double> &y = h.y1, &yg = h.yg1;  // 1-based
    PressureBC c;
    c.g_b = (-3.0 * y[1] + 4.0 * y[2] - y[3]) / 2.0;
    c.g2_b = 2.0 * y[1] - 5.0 * y[2] + 4.0 * y[3] - 1.0 * y[4];
    c.g_t = (3.0 * y[ny] - 4.0 * y[ny - 1] + y[ny - 2]) / 2.0;
    c.g2_t = 2.0 * y[ny] - 5.0 * y[ny - 1] + 4.0 * y[ny - 2] - 1.0 * y[ny - 3];
    c.nui_b = h.prm.nu + 0.5 * (a1 + a2);
    c.dnu_b = (a2 - a1) / (yg[2] - yg[1]);
    c.nui_t = h.prm.nu + 0.5 * (an + an1);
    c.dnu_t = (an - an1) / (yg[nyg] - yg[nyg - 1]);
    const double coef_bc_1 = 1.0 / (y[2] - y[1]) / (yg[2] - yg[1]);                  // initialization.f90:360-364
    const double coef_bc_2 = 1.0 / (y[nyg - 1] - y[nyg - 2]) / (yg[nyg] - yg[nyg - 1]);  // :369-373
    c.src_b = coef_bc_1 * (yg[2] - yg[1]);
    c.src_t = -(coef_bc_2 * (yg[nyg] - yg[nyg - 1]));
    if (e == hipSuccess) {
        tic("compute_pressure");
        launch_rhs_only<S, S>(d, met, F[0], F[1], F[2], nut ? NT : nullptr, Ru, Rv, Rw, sc, ubar, h.prm.nu,
                              h.prm.omega_z, stream, true);
        launch_pressure_bc<S>(d, F[1], c, Rv, bcs, ysrc, stream);
        launch_div_fftx<double, S>(d, met, Ru, Rv, Rw, spec, pM, twM, twN, stream, ysrc);
        if (cscr && zsolve_fused(d)) {  // the scratch is sized for either form when allocated
            launch_zsolve<S>(d, spec, cscr, pZ, twZ, m_kxx, m_kzz, m_dd, m_dl, m_du, stream);
        } else {
            launch_fftz<S>(d, spec, pZ, twZ, -1, stream);
            launch_thomas<S>(d, spec, cscr, m_kxx, m_kzz, m_dd, m_dl, m_du, stream);
            launch_fftz<S>(d, spec, pZ, twZ, +1, stream);
        }
        launch_ifftx_proju<S, S, double>(d, met, spec, pcell, (S *)nullptr, pM, twM, twN, (double)(d.nxp * d.nzp),
                                         stream);
        toc();
        e = hipGetLastError();
    }
    std::vector<double> pc(ncell), bh(2 * nb);
    if (e == hipSuccess) e = hipMemcpyAsync(pc.data(), pcell, ncell * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(bh.data(), bcs, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    hipFree(buf);
    if (e != hipSuccess) { set_error(std::string("compute_pressure: ") + hipGetErrorString(e)); return -1; }
    auto idx = [&](int i, int j, int k) { return i3(nxg, nyg, i, j, k); };
    auto b = [&](int w, int i, int k) { return bh[(size_t)w * nb + (size_t)(k - 2) * (nxg - 2) + (i - 2)]; };
    for (int k = 2; k <= nzg - 2; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 2; i++)
                P[idx(i, j, k)] = pc[((size_t)(k - 2) * d.nyp + (j - 2)) * d.nxp + (i - 2)];
    for (int k = 2; k <= nzg - 2; k++)
        for (int i = 2; i <= nxg - 2; i++) {
            P[idx(i, 1, k)] = P[idx(i, 2, k)] - (yg[2] - yg[1]) * b(0, i, k);
            P[idx(i, nyg, k)] = P[idx(i, nyg - 1, k)] + (yg[nyg] - yg[nyg - 1]) * b(1, i, k);
        }
    for (int k = 2; k <= nzg - 2; k++)
        for (int j = 1; j <= nyg; j++) P[idx(nxg - 1, j, k)] = P[idx(2, j, k)];
    for (int j = 1; j <= nyg; j++)
        for (int i = 2; i <= nxg - 1; i++) P[idx(i, j, nzg - 1)] = P[idx(i, j, 2)];
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) { P[idx(1, j, k)] = P[idx(2, j, k)]; P[idx(nxg, j, k)] = P[idx(nxg - 1, j, k)]; }
    const size_t pl = (size_t)nxg * nyg;
    std::memcpy(P, P + pl, pl * sizeof(double));
    std::memcpy(P + pl * (nzg - 1), P + pl * (nzg - 2), pl * sizeof(double));
    return 0;
}

template <typename S>
int Ctx<S>::set_nu_t(const double *nt) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    const size_t n = (size_t)d.nxg * d.nyg * d.nzg;
    HCK(hipStreamSynchronize(stream));
    bool nz = false;
    for (size_t q = 0; q < n && !nz; q++) nz = nt[q] != 0.0;
    if (nz != nut_user) invalidate_graphs();
    nut_user = nz;
    return put(NT, nt, n);
}
template <typename S>
int Ctx<S>::get_nu_t(double *nt) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    return get(nt, NT, (size_t)d.nxg * d.nyg * d.nzg);
}
// avg_nu_t (global.f90:180): the dynamic model's plane coefficient
// <LM>/<MM> (subgrid.f90:505-535), or the clipped plane mean of nu_t of the
// completed Smagorinsky (subgrid.f90:605-617), wall rows by the nu_t rule
template <typename S>
int Ctx<S>::get_avg_nu_t(double *avg) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    const int nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    std::fill(avg, avg + nyg, 0.0);
    if (h.prm.les_model == 2 && !nut_user) {
        HCK(hipStreamSynchronize(stream));
        HCK(hipMemcpy(avg, dsm.avg, nyg * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    }
    if (h.prm.les_model == 0 && !nut_user) return 0;
    std::vector<double> nt((size_t)nxg * nyg * nzg);
    if (get_nu_t(nt.data())) return -1;
    for (int j = 2; j <= nyg - 1; j++) {
        double sum = 0.0;
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) sum += nt[i3(nxg, nyg, i, j, k)];
        avg[j - 1] = std::max(sum / (double)((nxg - 2) * (nzg - 2)), 0.0);
    }
    const double sg = h.prm.dirichlet_nu_t == 1 ? -1.0 : 1.0;
    avg[0] = sg * avg[1];
    avg[nyg - 1] = sg * avg[nyg - 2];
    return 0;
}

template <typename S>
int Ctx<S>::get_wall(double *ax, double *az, double *ti) {
    HCK(hipSetDevice(dev));
    HCK(hipStreamSynchronize(stream));
    const Dims &d = h.d;
    const size_t nwx = (size_t)d.nx * 2 * d.nzg, nwz = (size_t)d.nxg * 2 * d.nz;
    if (ax) HCK(hipMemcpy(ax, alpha_x, nwx * sizeof(double), hipMemcpyDeviceToHost));
    if (az) HCK(hipMemcpy(az, alpha_z, nwz * sizeof(double), hipMemcpyDeviceToHost));
    if (ti) HCK(hipMemcpy(ti, tau_init, nwx * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}
template <typename S>
int Ctx<S>::set_tau_init(const double *ti) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    HCK(hipStreamSynchronize(stream));
    HCK(hipMemcpy(tau_init, ti, (size_t)d.nx * 2 * d.nzg * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

template <typename S>
int Ctx<S>::set_scalars(const double *v) {
    HCK(hipSetDevice(dev));
    dt_fresh = false;
    HCK(hipStreamSynchronize(stream));
    Scalars s;
    HCK(hipMemcpy(&s, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    s.t = v[0]; s.dt = v[1]; s.dPdx = v[2]; s.dPdz = v[3]; s.qflow_x0 = v[4]; s.dPdy = v[5]; s.qflow_y0 = v[6];
    HCK(hipMemcpy(sc, &s, sizeof(Scalars), hipMemcpyHostToDevice));
    return 0;
}
template <typename S>
int Ctx<S>::get_scalars(double *v) {
    HCK(hipSetDevice(dev));
    HCK(hipStreamSynchronize(stream));
    Scalars s;
    HCK(hipMemcpy(&s, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    v[0] = s.t; v[1] = s.dt; v[2] = s.dPdx; v[3] = s.dPdz; v[4] = s.qflow_x0; v[5] = s.dPdy; v[6] = s.qflow_y0;
    return 0;
}
template <typename S>
int Ctx<S>::get_raw_scalars(Scalars *out) {
    HCK(hipSetDevice(dev));
    HCK(hipStreamSynchronize(stream));
    HCK(hipMemcpy(out, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    return 0;
}

template <typename S>
int Ctx<S>::mean_mass_flow_u(double *q) {
    HCK(hipSetDevice(dev));
    launch_massflow<S>(h.d, met, st[cur][0], part, sc, 0, stream);
    HCK(hipGetLastError());
    Scalars s;
    if (get_raw_scalars(&s)) return -1;
    *q = s.pad[0];
    return 0;
}

template <typename S>
int Ctx<S>::mean_mass_flow_v(double *q) {
    HCK(hipSetDevice(dev));
    launch_massflow_v<S>(h.d, m_y, st[cur][1], part, sc, 0, stream);
    HCK(hipGetLastError());
    Scalars s;
    if (get_raw_scalars(&s)) return -1;
    *q = s.pad[1];
    return 0;
}

template <typename S>
int Ctx<S>::mean_centerline_u(double *q) {
    HCK(hipSetDevice(dev));
    launch_centerline_u<S>(h.d, st[cur][0], ubar, part, sc, 0, stream);
    HCK(hipGetLastError());
    Scalars s;
    if (get_raw_scalars(&s)) return -1;
    *q = s.pad[0];
    return 0;
}

template <typename S>
int Ctx<S>::init_mass_flow() {
    double q;
    if (mean_mass_flow_u(&q)) return -1;
    Scalars s;
    if (get_raw_scalars(&s)) return -1;
    s.qflow_x0 = q;
    HCK(hipMemcpy(sc, &s, sizeof(Scalars), hipMemcpyHostToDevice));
    return 0;
}

template <typename S>
int Ctx<S>::check_divergence(double *ms, double *ma) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    launch_div<S>(d, m_x, m_y, m_z, st[cur][0], st[cur][1], st[cur][2], part, stream);
    HCK(hipGetLastError());
    std::vector<double> hp(2 * (d.nzg - 2));
    HCK(hipStreamSynchronize(stream));
    HCK(hipMemcpy(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost));
    double a = 0.0, b = 0.0;
    for (int k = 0; k < d.nzg - 2; k++) { a = std::max(a, hp[2 * k]); b = std::max(b, hp[2 * k + 1]); }
    if (ms) *ms = a;
    if (ma) *ma = b;
    return 0;
}

// compute_statistics (statistics.f90:22-171): the plane means, u_tau, Re_tau
// and the bulk flows; a NaN anywhere in U, V or W is an error, as the
// reference's Stop 'Error: NaNs!' (:165-167)
template <typename S>
int Ctx<S>::statistics(double *out) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    S *const *F = st[cur];
    launch_stats<S>(d, met, F[0], F[1], F[2], ubar, stats_out, stream);
    launch_stats_tail<S>(d, met, m_y, m_yg, F[0], F[1], F[2], NT, ubar, stats_out, part, sc, h.prm.nu, stream);
    HCK(hipGetLastError());
    HCK(hipStreamSynchronize(stream));
    HCK(hipMemcpy(out, stats_out, (7 * d.ny + 4) * sizeof(double), hipMemcpyDeviceToHost));
    double nan_flag = 0.0;
    HCK(hipMemcpy(&nan_flag, stats_out + 7 * d.ny + 4, sizeof(double), hipMemcpyDeviceToHost));
    // statistics.f90:107-115 normalisations
    const double cu = (double)((d.nx - 2) * (d.nzg - 2)), cv = (double)((d.nxg - 2) * (d.nzg - 2)),
                 cw = (double)((d.nxg - 2) * (d.nz - 2));
    const double den[7] = {cu, cv, cw, cu, cv, cw, cv};
    for (int c = 0; c < 7; c++)
        for (int j = 0; j < d.ny; j++) out[c * d.ny + j] /= den[c];
    if (nan_flag != 0.0) {
        set_error("compute_statistics: Error: NaNs!");
        return -1;
    }
    return 0;
}

// time-window statistics: the same plane sums as statistics(), added into
// an FP64 device accumulator (no host round trip; may be captured)
template <typename S>
int Ctx<S>::stats_accumulate() {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    tic("stats_accumulate");
    launch_stats<S>(d, met, st[cur][0], st[cur][1], st[cur][2], ubar, stats_out, stream);
    launch_stats_accum(stats_out, stats_acc, 7 * d.ny, stream);
    toc();
    HCK(hipGetLastError());
    return 0;
}

template <typename S>
int Ctx<S>::stats_window(double *out, double *count) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    std::vector<double> a(7 * d.ny + 1);
    HCK(hipStreamSynchronize(stream));
    HCK(hipMemcpy(a.data(), stats_acc, a.size() * sizeof(double), hipMemcpyDeviceToHost));
    const double n = a[7 * d.ny];
    if (count) *count = n;
    const double cu = (double)((d.nx - 2) * (d.nzg - 2)), cv = (double)((d.nxg - 2) * (d.nzg - 2)),
                 cw = (double)((d.nxg - 2) * (d.nz - 2));
    const double den[7] = {cu, cv, cw, cu, cv, cw, cv};
    for (int c = 0; c < 7; c++)
        for (int j = 0; j < d.ny; j++) out[c * d.ny + j] = n > 0 ? a[c * d.ny + j] / (den[c] * n) : 0.0;
    return 0;
}

template <typename S>
int Ctx<S>::stats_reset() {
    HCK(hipSetDevice(dev));
    HCK(hipMemsetAsync(stats_acc, 0, (7 * h.d.ny + 8) * sizeof(double), stream));
    return 0;
}

template <typename S>
int Ctx<S>::compute_rhs(const double *U, const double *V, const double *W, double *ru, double *rv,
                        double *rw) {
    HCK(hipSetDevice(dev));
    const Dims &d = h.d;
    // stage the inputs in the idle ping-pong buffers
    S *const *T = st[1 - cur];
    if (ubar) {
        if (sync_profile()) return -1;
        std::vector<double> Ud(U, U + n3[0]);
        for (int k = 1; k <= d.nzg; k++)
            for (int j = 1; j <= d.nyg; j++)
                for (int i = 1; i <= d.nx; i++) Ud[i3(d.nx, d.nyg, i, j, k)] -= ubar_h[j];
        if (put(T[0], Ud.data(), n3[0])) return -1;
    } else if (put(T[0], U, n3[0])) {
        return -1;
    }
    if (put(T[1], V, n3[1]) || put(T[2], W, n3[2])) return -1;
    const size_t nru = (size_t)(d.nx - 2) * (d.nyg - 2) * (d.nzg - 2);
    const size_t nrv = (size_t)(d.nxg - 2) * (d.ny - 2) * (d.nzg - 2);
    const size_t nrw = (size_t)(d.nxg - 2) * (d.nyg - 2) * (d.nz - 2);
    double *dr = nullptr;
    HCK(hipMalloc(&dr, (nru + nrv + nrw) * sizeof(double)));
    const bool nut = h.prm.les_model != 0 || nut_user;
    launch_rhs_only<S, S>(d, met, T[0], T[1], T[2], nut ? NT : nullptr, dr, dr + nru, dr + nru + nrv, sc,
                          ubar, h.prm.nu, h.prm.omega_z, stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess && ru) e = hipMemcpy(ru, dr, nru * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && rv) e = hipMemcpy(rv, dr + nru, nrv * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && rw) e = hipMemcpy(rw, dr + nru + nrv, nrw * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(dr);
    if (e != hipSuccess) { set_error(std::string("compute_rhs: ") + hipGetErrorString(e)); return -1; }
    return 0;
}

template <typename S>
int Ctx<S>::compute_eddy_viscosity() {
    HCK(hipSetDevice(dev));
    if (h.prm.les_model == 0) {
        HCK(hipMemsetAsync(NT, 0, (size_t)h.d.nxg * h.d.nyg * h.d.nzg * sizeof(S), stream));
        if (nut_user) invalidate_graphs();
        nut_user = false;
    }
    eddy_viscosity(st[cur]);
    HCK(hipGetLastError());
    HCK(hipStreamSynchronize(stream));
    return 0;
}

template <typename S>
int Ctx<S>::compute_wall_model(int *n_nan, int *n_iter) {
    HCK(hipSetDevice(dev));
    wall_model(st[cur]);
    HCK(hipGetLastError());
    Scalars s;
    if (get_raw_scalars(&s)) return -1;
    if (n_nan) *n_nan = s.n_nan_eqwm;
    if (n_iter) *n_iter = s.n_iter_eqwm;
    return 0;
}

// learned wall model: predicted stresses where the inputs lie inside the
// training hull; the ODE solver handles the rest (wallmodel kernel)
template <typename S>
int Ctx<S>::attach_mlp(Mlp *m) {
    HCK(hipSetDevice(dev));
    if (batch_member && m) { set_error("the learned wall model is not taken by a batch member"); return -1; }
    invalidate_graphs();
    if (!m) { surrogate = nullptr; return 0; }
    const Dims &d = h.d;
    const int n = d.nx * d.nzg * 2;
    if (!wall_u) {
        if (dalloc(&wall_u, n, owned, bytes, (Slab *)nullptr) || dalloc(&tau_pred, n, owned, bytes, (Slab *)nullptr) ||
            dalloc(&use_pred, n, owned, bytes, (Slab *)nullptr))
            return -1;
    }
    const double hw = h.yg[h.prm.jref_eqwm - 1];
    surrogate = [this, m, n, hw](Ctx<S> *, S *const *F, const double **pred, const int **use) {
        tic("mlp_wall");
        launch_wall_inputs<S>(h.d, F[0], F[2], wt, h.prm.jref_eqwm, ubar, wall_u, stream);
        m->wall_predict(n, wall_u, hw, h.prm.nu, &sc->dPdx, tau_pred, use_pred, sc, stream);
        toc();
        *pred = tau_pred;
        *use = use_pred;
    };
    return 0;
}

template <typename S>
int Ctx<S>::wall_fallbacks() {
    Scalars s;
    if (get_raw_scalars(&s)) return -1;
    return s.n_fallback_mlp;
}

template <typename S>
std::string Ctx<S>::timing_report() {
    collect();
    std::string out;
    char line[256];
    for (auto &kv : tstat) {
        std::snprintf(line, sizeof line, "%s %.6f %ld\n", kv.first.c_str(),
                      kv.second.second ? kv.second.first / kv.second.second : 0.0, (long)kv.second.second);
        out += line;
    }
    return out;
}

template struct Ctx<double>;
template struct Ctx<float>;

std::unique_ptr<CtxBase> make_ctx(const Host &h, Slab *slab, hipStream_t stream) {
    std::unique_ptr<CtxBase> c;
    if (h.prm.precision == 0) c.reset(new Ctx<double>(h));
    else c.reset(new Ctx<float>(h));
    c->slab = slab;
    c->stream = stream;
    if (c->init()) return nullptr;
    return c;
}

std::string batch_class_key(const Host &h) {
    // the struct is read as bytes: it has no padding (4 ints, 5 doubles,
    // 6 ints, 1 double, 2 ints, 3 doubles, 2 ints, 1 double)
    static_assert(sizeof(cf_params) == 136, "cf_params layout changed: revisit batch_class_key");
    const cf_params &p = h.prm;
    if (p.les_model == 2 || p.y_mass_cte != 0 || p.wall_surrogate != 0) return "";
    cf_params q = p;
    q.nu = q.dPdx = q.dPdz = q.u_centerline = q.tau_mean_eqwm = 0.0;
    std::string k(reinterpret_cast<const char *>(&q), sizeof q);
    for (const std::vector<double> *v : {&h.x, &h.y, &h.z})
        k.append(reinterpret_cast<const char *>(v->data()), v->size() * sizeof(double));
    return k;
}

size_t ctx_footprint(const Host &h) {
    Slab dry;
    dry.dry = true;
    auto c = make_ctx(h, &dry, nullptr);
    return c ? dry.used : 0;
}

}  // namespace cf

// ===========================================================================
// extern "C" surface
// ===========================================================================
using cf::g_err;
using cf::set_error;

struct cf_ctx {
    std::unique_ptr<cf::CtxBase> impl;
};

#define CTX_CHECK(c)                                           \
    do {                                                       \
        if (!(c) || !(c)->impl) {                              \
            set_error("null context");                         \
            return -1;                                         \
        }                                                      \
    } while (0)

extern "C" {

const char *cf_last_error(void) { return g_err.c_str(); }

int cf_create(const cf_params *p, const double *x, const double *y, const double *z, cf_ctx **out) {
    if (!p || !x || !y || !z || !out) { set_error("cf_create: null argument"); return -1; }
    *out = nullptr;
    cf::Host h(*p, x, y, z);
    std::string v = h.validate();
    if (!v.empty()) { set_error("cf_create: " + v); return -1; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("cf_create: no HIP device available");
        return -1;
    }
    if (p->device < 0 || p->device >= ndev) { set_error("cf_create: bad device ordinal"); return -1; }
    std::unique_ptr<cf::CtxBase> impl;
    if (p->precision == 0) impl.reset(new cf::Ctx<double>(h));
    else impl.reset(new cf::Ctx<float>(h));
    if (impl->init()) return -1;
    cf_ctx *c = new cf_ctx;
    c->impl = std::move(impl);
    *out = c;
    return 0;
}

long long cf_footprint(const cf_params *p, const double *x, const double *y, const double *z) {
    if (!p || !x || !y || !z) { set_error("cf_footprint: null argument"); return -1; }
    cf::Host h(*p, x, y, z);
    std::string v = h.validate();
    if (!v.empty()) { set_error("cf_footprint: " + v); return -1; }
    return (long long)cf::ctx_footprint(h);
}

int cf_destroy(cf_ctx *c) {
    if (!c) return 0;
    if (c->impl && c->impl->batch_member) {
        set_error("cf_destroy: a batch member is destroyed with its batch (cf_batch_destroy)");
        return -1;
    }
    delete c;
    return 0;
}

// ---- batches of one size class ----
struct cf_batch {
    int dev = 0;
    size_t fp = 0;
    char *base = nullptr;
    hipStream_t stream = nullptr;
    std::vector<cf::Slab> slabs;
    std::vector<cf_ctx *> members;
};

static void batch_free(cf_batch *b) {
    if (!b) return;
    hipSetDevice(b->dev);
    if (b->stream) hipStreamSynchronize(b->stream);
    for (cf_ctx *c : b->members) delete c;
    b->members.clear();
    if (b->base) hipFree(b->base);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}


int cf_batch_create(int n, const cf_params *p, const double *x, const double *y, const double *z, cf_batch **out) {
    if (n < 1 || !p || !x || !y || !z || !out) { set_error("cf_batch_create: bad arguments"); return -1; }
    *out = nullptr;
    const cf_params &p0 = p[0];
    std::vector<cf::Host> hosts;
    hosts.reserve(n);
    for (int m = 0; m < n; m++) {
        hosts.emplace_back(p[m], x, y, z);
        const std::string v = hosts.back().validate();
        if (!v.empty()) { set_error("cf_batch_create: " + v); return -1; }
        const std::string k = cf::batch_class_key(hosts.back());
        if (k.empty()) {
            set_error("cf_batch_create: the batched step takes no dynamic Smagorinsky model, y mass flow or learned "
                      "wall model");
            return -1;
        }
        if (m > 0 && k != cf::batch_class_key(hosts[0])) {
            set_error("cf_batch_create: members differ in an option other than nu, dPdx, dPdz, u_centerline and "
                      "tau_mean_eqwm");
            return -1;
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error("cf_batch_create: no HIP device"); return -1; }
    if (p0.device < 0 || p0.device >= ndev) { set_error("cf_batch_create: bad device ordinal"); return -1; }
    const size_t fp = cf::ctx_footprint(hosts[0]);
    if (fp == 0) return -1;
    for (int m = 1; m < n; m++)
        if (cf::ctx_footprint(hosts[m]) != fp) { set_error("cf_batch_create: member footprints differ"); return -1; }
    cf_batch *b = new cf_batch;
    b->dev = p0.device;
    b->fp = fp;
    b->slabs.resize(n);
    if (hipSetDevice(b->dev) != hipSuccess || hipMalloc(&b->base, fp * (size_t)n) != hipSuccess) {
        b->base = nullptr;
        batch_free(b);
        set_error("cf_batch_create: device allocation failed");
        return -1;
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        b->stream = nullptr;
        batch_free(b);
        set_error("cf_batch_create: stream creation failed");
        return -1;
    }
    for (int m = 0; m < n; m++) {
        b->slabs[m].base = b->base + (size_t)m * fp;
        b->slabs[m].cap = fp;
        auto impl = cf::make_ctx(hosts[m], &b->slabs[m], b->stream);
        if (!impl) { batch_free(b); return -1; }
        if (b->slabs[m].used != fp) {
            batch_free(b);
            set_error("cf_batch_create: member layout differs from the footprint");
            return -1;
        }
        impl->batch_member = true;
        cf_ctx *c = new cf_ctx;
        c->impl = std::move(impl);
        b->members.push_back(c);
    }
    *out = b;
    return 0;
}

int cf_batch_member(cf_batch *b, int m, cf_ctx **ctx) {
    if (!b || !ctx || m < 0 || m >= (int)b->members.size()) { set_error("cf_batch_member: bad arguments"); return -1; }
    *ctx = b->members[m];
    return 0;
}

int cf_batch_run_steps(cf_batch *b, int nsteps, int use_graph) {
    if (!b || nsteps < 0) { set_error("cf_batch_run_steps: bad arguments"); return -1; }
    std::vector<cf::CtxBase *> mem;
    for (cf_ctx *c : b->members) mem.push_back(c->impl.get());
    return mem[0]->batch_run(mem, (int64_t)b->fp, nsteps, use_graph != 0);
}

int cf_batch_destroy(cf_batch *b) {
    batch_free(b);
    return 0;
}

int cf_set_state(cf_ctx *c, const double *U, const double *V, const double *W) {
    CTX_CHECK(c);
    return c->impl->set_state(U, V, W);
}
int cf_get_state(cf_ctx *c, double *U, double *V, double *W) {
    CTX_CHECK(c);
    return c->impl->get_state(U, V, W);
}
int cf_get_pressure(cf_ctx *c, double *P) { CTX_CHECK(c); return c->impl->get_pressure(P); }
int cf_compute_pressure(cf_ctx *c, double *P) {
    CTX_CHECK(c);
    if (!P) { set_error("cf_compute_pressure: null output"); return -1; }
    return c->impl->compute_pressure(P);
}
int cf_set_nu_t(cf_ctx *c, const double *nt) { CTX_CHECK(c); return c->impl->set_nu_t(nt); }
int cf_get_nu_t(cf_ctx *c, double *nt) { CTX_CHECK(c); return c->impl->get_nu_t(nt); }
int cf_get_avg_nu_t(cf_ctx *c, double *avg) {
    CTX_CHECK(c);
    if (!avg) { set_error("cf_get_avg_nu_t: null output"); return -1; }
    return c->impl->get_avg_nu_t(avg);
}
int cf_get_wall(cf_ctx *c, double *ax, double *az, double *ti) { CTX_CHECK(c); return c->impl->get_wall(ax, az, ti); }
int cf_set_tau_init(cf_ctx *c, const double *ti) { CTX_CHECK(c); return c->impl->set_tau_init(ti); }
int cf_set_scalars(cf_ctx *c, const double *s) { CTX_CHECK(c); return c->impl->set_scalars(s); }
int cf_get_scalars(cf_ctx *c, double *s) { CTX_CHECK(c); return c->impl->get_scalars(s); }
int cf_init_mass_flow(cf_ctx *c) { CTX_CHECK(c); return c->impl->init_mass_flow(); }
int cf_apply_boundary_conditions(cf_ctx *c) { CTX_CHECK(c); return c->impl->apply_bc(); }
int cf_compute_dt(cf_ctx *c, double *dt) { CTX_CHECK(c); return c->impl->compute_dt(dt); }
int cf_rk3_step(cf_ctx *c) {
    CTX_CHECK(c);
    if (c->impl->rk3_step()) return -1;
    return c->impl->sync();
}
int cf_run_steps(cf_ctx *c, int nsteps, int use_graph) {
    CTX_CHECK(c);
    if (nsteps < 0) { set_error("cf_run_steps: negative step count"); return -1; }
    return c->impl->run_steps(nsteps, use_graph != 0);
}
int cf_projection_step(cf_ctx *c, int rk_step) { CTX_CHECK(c); return c->impl->projection_step(rk_step); }
int cf_compute_rhs(cf_ctx *c, const double *U, const double *V, const double *W, double *ru, double *rv,
                   double *rw) {
    CTX_CHECK(c);
    if (!U || !V || !W) { set_error("cf_compute_rhs: null input"); return -1; }
    return c->impl->compute_rhs(U, V, W, ru, rv, rw);
}
int cf_compute_eddy_viscosity(cf_ctx *c) { CTX_CHECK(c); return c->impl->compute_eddy_viscosity(); }
int cf_compute_wall_model(cf_ctx *c, int *n_nan, int *n_iter) {
    CTX_CHECK(c);
    return c->impl->compute_wall_model(n_nan, n_iter);
}
int cf_mean_mass_flow_u(cf_ctx *c, double *q) { CTX_CHECK(c); return c->impl->mean_mass_flow_u(q); }
int cf_mean_mass_flow_v(cf_ctx *c, double *q) { CTX_CHECK(c); return c->impl->mean_mass_flow_v(q); }
int cf_mean_centerline_u(cf_ctx *c, double *q) { CTX_CHECK(c); return c->impl->mean_centerline_u(q); }
int cf_check_divergence(cf_ctx *c, double *ms, double *ma) { CTX_CHECK(c); return c->impl->check_divergence(ms, ma); }
int cf_statistics(cf_ctx *c, double *out) { CTX_CHECK(c); return c->impl->statistics(out); }

int cf_eqwm_batch(int n, const double *umag, const double *nu, const double *h, const double *tau_in,
                  int ncv, double stretch, double tol, int max_iter, double *tau, int *iter) {
    if (n < 0 || (n > 0 && (!umag || !nu || !h || !tau_in || !tau || !iter))) {
        set_error("cf_eqwm_batch: bad arguments");
        return -1;
    }
    if (ncv < 3 || ncv > cf::EQWM_MAXNCV) { set_error("cf_eqwm_batch: ncv out of range [3,128]"); return -1; }
    if (n == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error("cf_eqwm_batch: no HIP device"); return -1; }
    double *d = nullptr;
    int *di = nullptr;
    const size_t b = (size_t)n * sizeof(double);
    HCK(hipMalloc(&d, 5 * b));
    if (hipMalloc(&di, (size_t)n * sizeof(int)) != hipSuccess) { hipFree(d); set_error("cf_eqwm_batch: alloc"); return -1; }
    hipError_t e = hipMemcpy(d, umag, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, nu, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 2 * (size_t)n, h, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 3 * (size_t)n, tau_in, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        cf::EqwmParams p{ncv, max_iter, stretch, tol, 1.0};
        cf::launch_eqwm_batch(n, d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, d + 4 * (size_t)n, di, p, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(tau, d + 4 * (size_t)n, b, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(iter, di, (size_t)n * sizeof(int), hipMemcpyDeviceToHost);
    hipFree(d);
    hipFree(di);
    if (e != hipSuccess) { set_error(std::string("cf_eqwm_batch: ") + hipGetErrorString(e)); return -1; }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// cf_eqwm_batch_noalloc: the staging of cf_eqwm_batch (the four input columns,
// tau and iter of n points: 5 doubles + 1 int per point) in a file-scope
// device array instead of per-call device allocations.  The array is part of
// the code object: every GPU the library runs on holds one copy of
// CF_EQWM_NOALLOC_MAX points (23 MiB) until the process ends.  Its device
// address is looked up once per device (hipGetSymbolAddress, with that device
// current) and only that address is handed to copies and kernels: the array's
// name is a host-side stand-in in host code.  Calls on one device share the
// array, so calls that overlap on one device race; nothing in it has to start
// at zero (every input is copied in and every output written before it is
// read back).
// ---------------------------------------------------------------------------
namespace {
constexpr size_t EQWM_NOALLOC_BYTES = (size_t)CF_EQWM_NOALLOC_MAX * (5 * sizeof(double) + sizeof(int));
__device__ __align__(256) char scratch[EQWM_NOALLOC_BYTES];

std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_ptr;  // device ordinal -> address of that device's scratch

// device address of `scratch` on the current device (looked up on first use)
char *scratch_on_current_device(int *dev_out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) { set_error("cf_eqwm_batch_noalloc: hipGetDevice failed"); return nullptr; }
    *dev_out = dev;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    char *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("cf_eqwm_batch_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        return nullptr;
    }
    g_scratch_ptr[dev] = p;
    return p;
}
}  // namespace

extern "C" {

int cf_eqwm_batch_noalloc(int n, const double *umag, const double *nu, const double *h, const double *tau_in,
                          int ncv, double stretch, double tol, int max_iter, double *tau, int *iter) {
    if (n < 0 || (n > 0 && (!umag || !nu || !h || !tau_in || !tau || !iter))) {
        set_error("cf_eqwm_batch_noalloc: bad arguments");
        return -1;
    }
    if (n > CF_EQWM_NOALLOC_MAX) {
        set_error("cf_eqwm_batch_noalloc: n = " + std::to_string(n) + " exceeds the maximum of " +
                  std::to_string(CF_EQWM_NOALLOC_MAX) + " points (use cf_eqwm_batch)");
        return -1;
    }
    if (ncv < 3 || ncv > cf::EQWM_MAXNCV) { set_error("cf_eqwm_batch_noalloc: ncv out of range [3,128]"); return -1; }
    if (n == 0) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("cf_eqwm_batch_noalloc: no HIP device");
        return -1;
    }
    int dev = -1;
    char *base = scratch_on_current_device(&dev);
    if (!base) return -1;
    double *d = (double *)base;  // umag | nu | h | tau_in | tau, n each
    int *di = (int *)(d + 5 * (size_t)n);
    const size_t b = (size_t)n * sizeof(double);
    hipError_t e = hipMemcpy(d, umag, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, nu, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 2 * (size_t)n, h, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 3 * (size_t)n, tau_in, b, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        cf::EqwmParams p{ncv, max_iter, stretch, tol, 1.0};
        cf::launch_eqwm_batch(n, d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, d + 4 * (size_t)n, di, p, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e == hipSuccess) e = hipMemcpy(tau, d + 4 * (size_t)n, b, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(iter, di, (size_t)n * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error(std::string("cf_eqwm_batch_noalloc: ") + hipGetErrorString(e));
        return -1;
    }
    return 0;
}

int cf_set_wall_surrogate(cf_ctx *c, cf_mlp *m) {
    CTX_CHECK(c);
    return c->impl->attach_mlp(m ? cf::mlp_of(m) : nullptr);
}
int cf_wall_fallbacks(cf_ctx *c) { CTX_CHECK(c); return c->impl->wall_fallbacks(); }

int cf_stats_accumulate(cf_ctx *c) { CTX_CHECK(c); return c->impl->stats_accumulate(); }
int cf_stats_window(cf_ctx *c, double *out, double *count) {
    CTX_CHECK(c);
    if (!out) { set_error("cf_stats_window: null output"); return -1; }
    return c->impl->stats_window(out, count);
}
int cf_stats_reset(cf_ctx *c) { CTX_CHECK(c); return c->impl->stats_reset(); }

int cf_timing_enable(cf_ctx *c, int on) { CTX_CHECK(c); c->impl->set_timing(on != 0); return 0; }
int cf_timing_reset(cf_ctx *c) { CTX_CHECK(c); c->impl->timing_reset(); return 0; }
int cf_timing_report(cf_ctx *c, char *buf, int len) {
    CTX_CHECK(c);
    std::string r = c->impl->timing_report();
    if (!buf || len <= 0) return (int)r.size() + 1;
    std::snprintf(buf, (size_t)len, "%s", r.c_str());
    return 0;
}
long long cf_device_bytes(cf_ctx *c) { return c && c->impl ? (long long)c->impl->device_bytes() : -1; }
int cf_sync(cf_ctx *c) { CTX_CHECK(c); return c->impl->sync(); }

}  // extern "C"
