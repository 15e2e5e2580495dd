// Stencil-class kernels of the RK3 step: fused RHS + Runge-Kutta update,
// ghost-cell boundary conditions, projection of V and W, bulk-flow and CFL
// reductions, divergence check, constant-coefficient Smagorinsky and plane
// statistics.  All are HBM-streaming kernels: one thread per unique cell,
// x fastest, tiles remapped per XCD.
#include <algorithm>
#include <cstdlib>

#include "cf_kernels.hpp"

namespace cf {

// ---------------------------------------------------------------------------
// Right-hand sides at one point (equations.f90:24-498), shared by the fused
// step kernel (LDS accessor) and the module-level RHS kernel (global
// accessor).  A.u(di,dj,dk) returns U(i+di, j+dj, k+dk) etc.  Per term the
// reference's association order is kept:
//   (((conv_1 + conv_2) + conv_3) + visc) + dP  [+ rotation]
// with divisions by grid spacings replaced by multiplications with
// precomputed reciprocals.
// ---------------------------------------------------------------------------
// The 1-D metrics a point needs, at stencil precision.  The i and j ones are
// loaded once per thread (they do not change along the k march), the k ones
// once per plane.  Every metric array is padded one entry past its range.
template <typename R>
struct PM {
    R w0j, w1j, w0m, w1m;      // weight_y at j and j-1
    R rxc, rxcm, rxf, rxfp;    // 1/dxc(i), 1/dxc(i-1), 1/dxf(i), 1/dxf(i+1)
    R ryf, ryfp, ryc, rycm;    // 1/dyf(j), 1/dyf(j+1), 1/dyc(j), 1/dyc(j-1)
    R rzf, rzfp, rzc, rzcm;    // 1/dzf(k), 1/dzf(k+1), 1/dzc(k), 1/dzc(k-1)
    __device__ void set_ij(const Metrics &m, int i, int j) {
        w0j = (R)m.w0[j]; w1j = (R)m.w1[j]; w0m = (R)m.w0[j - 1]; w1m = (R)m.w1[j - 1];
        rxc = (R)m.rdxc[i]; rxcm = (R)m.rdxc[i - 1]; rxf = (R)m.rdxf[i]; rxfp = (R)m.rdxf[i + 1];
        ryf = (R)m.rdyf[j]; ryfp = (R)m.rdyf[j + 1]; ryc = (R)m.rdyc[j]; rycm = (R)m.rdyc[j - 1];
    }
    __device__ void set_k(const Metrics &m, int k) {
        rzf = (R)m.rdzf[k]; rzfp = (R)m.rdzf[k + 1]; rzc = (R)m.rdzc[k]; rzcm = (R)m.rdzc[k - 1];
    }
};

template <typename R, bool NUT, bool ROT, class A>
__device__ __forceinline__ void rhs_point(const A &a, const PM<R> &pm, R nu, R om, R dPdx, R dPdz, bool do_u,
                                          bool do_v, bool do_w, R &ru, R &rv, R &rw) {
    const R half = (R)0.5, quart = (R)0.25, two = (R)2;
    const R w0j = pm.w0j, w1j = pm.w1j;
    const R w0m = pm.w0m, w1m = pm.w1m;
#define u(I, J, K) a.u(I, J, K)
#define v(I, J, K) a.v(I, J, K)
#define w(I, J, K) a.w(I, J, K)
#define nt(I, J, K) a.nt(I, J, K)
    if (do_u) {
        R c1 = half * (u(0, 0, 0) + u(1, 0, 0));
        R c0 = half * (u(-1, 0, 0) + u(0, 0, 0));
        R t1 = -(c1 * c1 - c0 * c0) * pm.rxc;
        R uv1 = (w0j * u(0, 0, 0) + w1j * u(0, 1, 0)) * (half * (v(0, 0, 0) + v(1, 0, 0)));
        R uv0 = (w0m * u(0, -1, 0) + w1m * u(0, 0, 0)) * (half * (v(0, -1, 0) + v(1, -1, 0)));
        R t2 = -(uv1 - uv0) * pm.ryf;
        R uw1 = (half * (u(0, 0, 0) + u(0, 0, 1))) * (half * (w(0, 0, 0) + w(1, 0, 0)));
        R uw0 = (half * (u(0, 0, -1) + u(0, 0, 0))) * (half * (w(0, 0, -1) + w(1, 0, -1)));
        R t3 = -(uw1 - uw0) * pm.rzf;
        R rdz1 = pm.rzcm, rdz2 = pm.rzc, rdz3 = pm.rzf;
        R rdy3 = pm.ryf;
        R rdx1 = pm.rxf, rdx2 = pm.rxfp, rdx3 = pm.rxc;
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + nt(0, 0, 0);
            nux2 = nu + nt(1, 0, 0);
            nuy1 = nu + half * (w0m * nt(0, -1, 0) + w1m * nt(0, 0, 0) + w0m * nt(1, -1, 0) + w1m * nt(1, 0, 0));
            nuy2 = nu + half * (w0j * nt(0, 0, 0) + w1j * nt(0, 1, 0) + w0j * nt(1, 0, 0) + w1j * nt(1, 1, 0));
            nuz1 = nu + quart * (nt(0, 0, 0) + nt(0, 0, -1) + nt(1, 0, 0) + nt(1, 0, -1));
            nuz2 = nu + quart * (nt(0, 0, 0) + nt(0, 0, 1) + nt(1, 0, 0) + nt(1, 0, 1));
        }
        R dudy1 = (u(0, 1, 0) - u(0, 0, 0)) * pm.ryc;
        R dudy0 = (u(0, 0, 0) - u(0, -1, 0)) * pm.rycm;
        R visc = two * rdx3 * (nux2 * (u(1, 0, 0) - u(0, 0, 0)) * rdx2 - nux1 * (u(0, 0, 0) - u(-1, 0, 0)) * rdx1) +
                 rdy3 * (nuy2 * (dudy1 + (v(1, 0, 0) - v(0, 0, 0)) * rdx2) -
                         nuy1 * (dudy0 + (v(1, -1, 0) - v(0, -1, 0)) * rdx1)) +
                 rdz3 * (nuz2 * ((u(0, 0, 1) - u(0, 0, 0)) * rdz2 + (w(1, 0, 0) - w(0, 0, 0)) * rdx2) -
                         nuz1 * ((u(0, 0, 0) - u(0, 0, -1)) * rdz1 + (w(1, 0, -1) - w(0, 0, -1)) * rdx1));
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        r = r + dPdx;
        if (ROT) {
            R aa = half * (v(0, -1, 0) + v(0, 0, 0));
            R bb = half * (v(1, -1, 0) + v(1, 0, 0));
            r = r + two * om * (half * (aa + bb));
        }
        ru = r;
    }
    if (do_v) {
        R c1 = half * (v(0, 0, 0) + v(0, 1, 0));
        R c0 = half * (v(0, -1, 0) + v(0, 0, 0));
        R t1 = -(c1 * c1 - c0 * c0) * pm.ryc;
        R uv1 = (w0j * u(0, 0, 0) + w1j * u(0, 1, 0)) * (half * (v(0, 0, 0) + v(1, 0, 0)));
        R uv0 = (w0j * u(-1, 0, 0) + w1j * u(-1, 1, 0)) * (half * (v(-1, 0, 0) + v(0, 0, 0)));
        R t2 = -(uv1 - uv0) * pm.rxf;
        R vw1 = (half * (v(0, 0, 0) + v(0, 0, 1))) * (w0j * w(0, 0, 0) + w1j * w(0, 1, 0));
        R vw0 = (half * (v(0, 0, -1) + v(0, 0, 0))) * (w0j * w(0, 0, -1) + w1j * w(0, 1, -1));
        R t3 = -(vw1 - vw0) * pm.rzf;
        R rdz1 = pm.rzcm, rdz2 = pm.rzc, rdz3 = pm.rzf;
        R rdy1 = pm.ryf, rdy2 = pm.ryfp, rdy3 = pm.ryc;
        R rdx1 = pm.rxcm, rdx2 = pm.rxc, rdx3 = pm.rxf;
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + half * (w0j * nt(-1, 0, 0) + w1j * nt(-1, 1, 0) + w0j * nt(0, 0, 0) + w1j * nt(0, 1, 0));
            nux2 = nu + half * (w0j * nt(0, 0, 0) + w1j * nt(0, 1, 0) + w0j * nt(1, 0, 0) + w1j * nt(1, 1, 0));
            nuy1 = nu + nt(0, 0, 0);
            nuy2 = nu + nt(0, 1, 0);
            nuz1 = nu + half * (w0j * nt(0, 0, -1) + w1j * nt(0, 1, -1) + w0j * nt(0, 0, 0) + w1j * nt(0, 1, 0));
            nuz2 = nu + half * (w0j * nt(0, 0, 0) + w1j * nt(0, 1, 0) + w0j * nt(0, 0, 1) + w1j * nt(0, 1, 1));
        }
        R visc = rdx3 * (nux2 * ((v(1, 0, 0) - v(0, 0, 0)) * rdx2 + (u(0, 1, 0) - u(0, 0, 0)) * rdy3) -
                         nux1 * ((v(0, 0, 0) - v(-1, 0, 0)) * rdx1 + (u(-1, 1, 0) - u(-1, 0, 0)) * rdy3)) +
                 two * rdy3 * (nuy2 * (v(0, 1, 0) - v(0, 0, 0)) * rdy2 - nuy1 * (v(0, 0, 0) - v(0, -1, 0)) * rdy1) +
                 rdz3 * (nuz2 * ((v(0, 0, 1) - v(0, 0, 0)) * rdz2 + (w(0, 1, 0) - w(0, 0, 0)) * rdy3) -
                         nuz1 * ((v(0, 0, 0) - v(0, 0, -1)) * rdz1 + (w(0, 1, -1) - w(0, 0, -1)) * rdy3));
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        if (ROT) {
            R aa = half * (u(-1, 0, 0) + u(0, 0, 0));
            R bb = half * (u(-1, 1, 0) + u(0, 1, 0));
            r = r - two * om * (w0j * aa + w1j * bb);
        }
        rv = r;
    }
    if (do_w) {
        R c1 = half * (w(0, 0, 0) + w(0, 0, 1));
        R c0 = half * (w(0, 0, -1) + w(0, 0, 0));
        R t1 = -(c1 * c1 - c0 * c0) * pm.rzc;
        R uw1 = (half * (u(0, 0, 0) + u(0, 0, 1))) * (half * (w(0, 0, 0) + w(1, 0, 0)));
        R uw0 = (half * (u(-1, 0, 0) + u(-1, 0, 1))) * (half * (w(-1, 0, 0) + w(0, 0, 0)));
        R t2 = -(uw1 - uw0) * pm.rxf;
        R vw1 = (half * (v(0, 0, 0) + v(0, 0, 1))) * (w0j * w(0, 0, 0) + w1j * w(0, 1, 0));
        R vw0 = (half * (v(0, -1, 0) + v(0, -1, 1))) * (w0m * w(0, -1, 0) + w1m * w(0, 0, 0));
        R t3 = -(vw1 - vw0) * pm.ryf;
        R rdz1 = pm.rzf, rdz2 = pm.rzfp, rdz3 = pm.rzc;
        R rdy3 = pm.ryf;
        R rdx1 = pm.rxcm, rdx2 = pm.rxc, rdx3 = pm.rxf;
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + quart * (nt(0, 0, 0) + nt(0, 0, 1) + nt(-1, 0, 0) + nt(-1, 0, 1));
            nux2 = nu + quart * (nt(0, 0, 0) + nt(0, 0, 1) + nt(1, 0, 0) + nt(1, 0, 1));
            nuy1 = nu + half * (w0m * nt(0, -1, 0) + w1m * nt(0, 0, 0) + w0m * nt(0, -1, 1) + w1m * nt(0, 0, 1));
            nuy2 = nu + half * (w0j * nt(0, 0, 0) + w1j * nt(0, 1, 0) + w0j * nt(0, 0, 1) + w1j * nt(0, 1, 1));
            nuz1 = nu + nt(0, 0, 0);
            nuz2 = nu + nt(0, 0, 1);
        }
        R dwdy1 = (w(0, 1, 0) - w(0, 0, 0)) * pm.ryc;
        R dwdy0 = (w(0, 0, 0) - w(0, -1, 0)) * pm.rycm;
        R visc = rdx3 * (nux2 * ((w(1, 0, 0) - w(0, 0, 0)) * rdx2 + (u(0, 0, 1) - u(0, 0, 0)) * rdz2) -
                         nux1 * ((w(0, 0, 0) - w(-1, 0, 0)) * rdx1 + (u(-1, 0, 1) - u(-1, 0, 0)) * rdz1)) +
                 rdy3 * (nuy2 * (dwdy1 + (v(0, 0, 1) - v(0, 0, 0)) * rdz2) -
                         nuy1 * (dwdy0 + (v(0, -1, 1) - v(0, -1, 0)) * rdz1)) +
                 two * rdz3 * (nuz2 * (w(0, 0, 1) - w(0, 0, 0)) * rdz2 - nuz1 * (w(0, 0, 0) - w(0, 0, -1)) * rdz1);
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        r = r + dPdz;
        rw = r;
    }
#undef u
#undef v
#undef w
#undef nt
}

// global-memory accessor (module-level RHS entry point)
template <typename S, typename R>
struct GAcc {
    const S *U, *V, *W, *NT;
    const double *ubar;
    int i, j, k, nx, nxg, nyg, ny;
    __device__ R u(int a, int b, int c) const {
        R x = (R)U[i3(nx, nyg, i + a, j + b, k + c)];
        return ubar ? x + (R)ubar[j + b] : x;
    }
    __device__ R v(int a, int b, int c) const { return (R)V[i3(nxg, ny, i + a, j + b, k + c)]; }
    __device__ R w(int a, int b, int c) const { return (R)W[i3(nxg, nyg, i + a, j + b, k + c)]; }
    __device__ R nt(int a, int b, int c) const { return (R)NT[i3(nxg, nyg, i + a, j + b, k + c)]; }
};

// ---------------------------------------------------------------------------
// Fused RHS + RK3 combine (time_integration.f90:104-145), 2.5-D blocked.
//
// A workgroup owns a 64 x TY tile of (i,j) and marches over a chunk of KC
// z-planes.  U, V, W (and nu_t) live in a 3-plane ring in LDS covering the
// tile plus a one-cell halo, so each value is read from HBM about once per
// chunk; all stencil reads are LDS reads.  One thread per unique cell
// (i in 2..nx-1, j in 2..nyg-1, k in 2..nzg-2); duplicate periodic planes are
// filled by the ghost kernel afterwards, as the reference's
// apply_boundary_conditions overwrites them.  Low-storage form:
//   substep 1:  U1 = U0 + dt*a11*F1,   G = U0 + dt*a21*F1
//   substep s:  Us = G  + dt*ass*Fs    (s = 2, 3; a31 = a21, a32 = 0)
// so the step keeps G instead of Uo and F1 (3 fields instead of 6).
// In the mixed path U is stored as a deviation from a per-plane FP64 profile
// ubar(j) (the RK combine is linear, so deviations update in place).
// ---------------------------------------------------------------------------
constexpr int RK_TX = 64, RK_KC = 16, RK_NSLOT = 4;

template <typename R, int TY>
struct LAcc {
    static constexpr int LX = RK_TX + 2, PL = (RK_TX + 2) * (TY + 2);
    const R *lu, *lv, *lw, *ln;  // ring base per field: [RK_NSLOT][TY+2][LX]
    int li, lj, slot[3];         // local (i,j) in the halo frame; ring slots of k-1, k, k+1
    R ub[3];                     // ubar(j-1), ubar(j), ubar(j+1)
    __device__ int at(int a, int b, int c) const { return slot[c + 1] * PL + (lj + b) * LX + (li + a); }
    __device__ R u(int a, int b, int c) const { return lu[at(a, b, c)] + ub[b + 1]; }
    __device__ R v(int a, int b, int c) const { return lv[at(a, b, c)]; }
    __device__ R w(int a, int b, int c) const { return lw[at(a, b, c)]; }
    __device__ R nt(int a, int b, int c) const { return ln[at(a, b, c)]; }
};

// The ring has four slots so one barrier per plane suffices: during plane k
// the workgroup reads slots (k-1, k, k+1) and writes plane k+2 (prefetched
// into registers one iteration earlier) into the slot of plane k-2.
// FP64 without nu_t and rotation: registers held to the four waves per
// SIMD that the LDS ring allows anyway (two 512-thread workgroups per CU);
// the mass-flow sum of the last substep would otherwise cost a wave
template <typename S, typename R, bool NUT, bool ROT, int SUB, int TY>
__global__ void __launch_bounds__(RK_TX *TY)
    __attribute__((amdgpu_waves_per_eu((sizeof(R) == 8 && !NUT && !ROT) ? 4 : 1)))
k_rhs_rk(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V,
         const S *__restrict__ W, const S *__restrict__ NT, S *__restrict__ Uo,
         S *__restrict__ Vo, S *__restrict__ Wo, S *__restrict__ Gu, S *__restrict__ Gv,
         S *__restrict__ Gw, const Scalars *__restrict__ sc, const double *__restrict__ ubar,
         double nu_d, double om_d, double a_new, double a_g, int tiles_x, int tiles_y,
         int nchunks, int ntiles, double *__restrict__ mfpart, int mf_mode, int64_t cs) {
    constexpr int LX = RK_TX + 2, PL = LX * (TY + 2), NT_ = RK_TX * TY;
    U = case_ptr(U, cs); V = case_ptr(V, cs); W = case_ptr(W, cs); NT = case_ptr(NT, cs);
    Uo = case_ptr(Uo, cs); Vo = case_ptr(Vo, cs); Wo = case_ptr(Wo, cs);
    Gu = case_ptr(Gu, cs); Gv = case_ptr(Gv, cs); Gw = case_ptr(Gw, cs);
    sc = case_ptr(sc, cs); ubar = case_ptr(ubar, cs); mfpart = case_ptr(mfpart, cs);
    if (cs) nu_d = sc->nu;
    constexpr int NPF = (PL + NT_ - 1) / NT_;  // plane elements per thread
    extern __shared__ unsigned char smem_raw[];
    R *ring = reinterpret_cast<R *>(smem_raw);
    R *lu = ring, *lv = ring + RK_NSLOT * PL, *lw = ring + 2 * RK_NSLOT * PL, *ln = ring + 3 * RK_NSLOT * PL;
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tc = t / (tiles_x * tiles_y);
    const int i0 = 2 + tx * RK_TX, j0 = 2 + ty * TY;
    const int kbeg = 2 + tc * RK_KC;
    const int kend = min(kbeg + RK_KC, d.nzg - 1);  // exclusive; unique planes 2..nzg-2
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny, nz = d.nz, nzg = d.nzg;
    const int lx = threadIdx.x, ly = threadIdx.y;
    const int tid = ly * RK_TX + lx;
    const int i = i0 + lx;
    const int j = __builtin_amdgcn_readfirstlane(j0 + ly);
    const bool active = i <= nx - 1 && j <= nyg - 1;

    R pu[NPF], pv[NPF], pw[NPF], pn[NPF];
    // in-plane offsets and bounds of this thread's halo-tile elements, fixed
    // along the k march (plane strides added per plane)
    int ou[NPF], ov[NPF], ow[NPF];
    unsigned char ok[NPF];
#pragma unroll
    for (int q = 0; q < NPF; q++) {
        const int e = tid + q * NT_;
        const int ex = e % LX, ey = e / LX;
        const int gi = i0 - 1 + ex, gj = j0 - 1 + ey;
        ou[q] = (gi - 1) + nx * (gj - 1);
        ov[q] = (gi - 1) + nxg * (gj - 1);
        ow[q] = (gi - 1) + nxg * (gj - 1);
        ok[q] = (e < PL) ? (unsigned char)(((gj <= nyg && gi <= nx) ? 1 : 0) | ((gj <= nyg && gi <= nxg) ? 2 : 0) |
                                           ((gj <= ny && gi <= nxg) ? 4 : 0))
                         : (unsigned char)0;
    }
    const int64_t psu = (int64_t)nx * nyg, psv = (int64_t)nxg * ny, psw = (int64_t)nxg * nyg;
    // The loads of fetch and gload are issued with no branch around them
    // (an element outside the field reads element 0), and the zeroing of
    // such elements waits for store: the compiler then counts the loads
    // exactly and the plane prefetched one iteration ahead stays in flight
    // through the next plane's arithmetic.
    auto inu = [&](int q, int kk) { return (ok[q] & 1) && kk <= nzg; };
    auto inw = [&](int q, int kk) { return (ok[q] & 2) && kk <= nz; };
    auto inn = [&](int q, int kk) { return (ok[q] & 2) && kk <= nzg; };
    auto inv = [&](int q, int kk) { return (ok[q] & 4) && kk <= nzg; };
    auto fetch = [&](int kk) {  // global -> registers (tile + halo of plane kk)
#pragma unroll
        for (int q = 0; q < NPF; q++) {
            pu[q] = (R)U[inu(q, kk) ? (int64_t)(kk - 1) * psu + ou[q] : 0];
            pw[q] = (R)W[inw(q, kk) ? (int64_t)(kk - 1) * psw + ow[q] : 0];
            pv[q] = (R)V[inv(q, kk) ? (int64_t)(kk - 1) * psv + ov[q] : 0];
            if (NUT) pn[q] = (R)NT[inn(q, kk) ? (int64_t)(kk - 1) * psw + ow[q] : 0];
        }
    };
    auto store = [&](int sl, int kk) {  // registers of plane kk -> ring slot
#pragma unroll
        for (int q = 0; q < NPF; q++) {
            const int e = tid + q * NT_;
            if (e < PL) {
                lu[sl * PL + e] = inu(q, kk) ? pu[q] : (R)0;
                lv[sl * PL + e] = inv(q, kk) ? pv[q] : (R)0;
                lw[sl * PL + e] = inw(q, kk) ? pw[q] : (R)0;
                if (NUT) ln[sl * PL + e] = inn(q, kk) ? pn[q] : (R)0;
            }
        }
    };
    fetch(kbeg - 1);
    store(0, kbeg - 1);
    fetch(kbeg);
    store(1, kbeg);
    fetch(kbeg + 1);
    store(2, kbeg + 1);
    if (kbeg + 2 <= kend) fetch(kbeg + 2);
    __syncthreads();

    LAcc<R, TY> a;
    a.lu = lu; a.lv = lv; a.lw = lw; a.ln = ln;
    a.li = lx + 1; a.lj = ly + 1;
    if (ubar) {
        a.ub[0] = (R)ubar[j - 1]; a.ub[1] = (R)ubar[j]; a.ub[2] = (R)ubar[min(j + 1, nyg)];
    } else {
        a.ub[0] = a.ub[1] = a.ub[2] = (R)0;
    }
    const R nu = (R)nu_d, om = (R)om_d, dPdx = (R)sc->dPdx, dPdz = (R)sc->dPdz;
    const double dt = sc->dt;
    PM<R> pm;
    pm.set_ij(m, min(i, nx), min(j, ny));
    const R cnew = (R)(dt * a_new), cg = (R)(dt * a_g);
    // substeps 2, 3 read G at the thread's own point: loaded one plane ahead
    // so the load has a whole plane of arithmetic to arrive
    const int64_t ou0 = (int64_t)(i - 1) + nx * (j - 1), ov0 = (int64_t)(i - 1) + nxg * (j - 1);
    const bool vrow = j <= ny - 1;
    R gu = 0, gv = 0, gw = 0;
    auto gload = [&](int kk, R &xu, R &xv, R &xw) {
        if (SUB != 1) {  // values of inactive threads and rows are never used
            const bool g = active && kk < kend, gv_ = g && vrow;
            xu = (R)Gu[g ? (int64_t)(kk - 1) * psu + ou0 : 0];
            xv = (R)Gv[gv_ ? (int64_t)(kk - 1) * psv + ov0 : 0];
            xw = (R)Gw[g ? (int64_t)(kk - 1) * psw + ov0 : 0];
        }
    };
    gload(kbeg, gu, gv, gw);
    // SUB 3 (the last substep): U* feeds the mass-flow correction.  The
    // projection that follows subtracts a periodic x difference of phi,
    // whose sum over the unique x faces vanishes, so the bulk (mf_mode 1)
    // and centreline (mf_mode 2) means of U* are those of the projected U.
    // Each thread sums its column over the k march (plane 2 counted twice
    // for its periodic image nzg-1 in the bulk, plane nzg-2 twice for its
    // image k = 1 on the centreline) and weights it once by its row: the
    // trapezoid weight over rows 2..nyg-1 (mass_flow.f90:116-118), or 1 on
    // row ny/2+1 (mass_flow.f90:205).
    double mf_acc = 0.0;
    const int mf_kdup = mf_mode == 1 ? 2 : nzg - 2;
    for (int k = kbeg; k < kend; k++) {
        const int r0 = k - kbeg;
        if (active) {
            a.slot[0] = r0 & 3; a.slot[1] = (r0 + 1) & 3; a.slot[2] = (r0 + 2) & 3;
            R ru = 0, rv = 0, rw = 0;
            pm.set_k(m, k);
            rhs_point<R, NUT, ROT>(a, pm, nu, om, dPdx, dPdz, true, j <= ny - 1, true, ru, rv, rw);
            const int c = a.slot[1] * PL + a.lj * LX + a.li;
            const R cu = lu[c], cv = lv[c], cw = lw[c];
            const int64_t qu = (int64_t)(k - 1) * psu + (i - 1) + nx * (j - 1),
                          qv = (int64_t)(k - 1) * psv + (i - 1) + nxg * (j - 1),
                          qw = (int64_t)(k - 1) * psw + (i - 1) + nxg * (j - 1);
            if (SUB == 1) {
                Uo[qu] = (S)(cu + cnew * ru);
                Gu[qu] = (S)(cu + cg * ru);
                if (j <= ny - 1) {
                    Vo[qv] = (S)(cv + cnew * rv);
                    Gv[qv] = (S)(cv + cg * rv);
                }
                Wo[qw] = (S)(cw + cnew * rw);
                Gw[qw] = (S)(cw + cg * rw);
            } else {
                const S un = (S)(gu + cnew * ru);
                Uo[qu] = un;
                if (j <= ny - 1) Vo[qv] = (S)(gv + cnew * rv);
                Wo[qw] = (S)(gw + cnew * rw);
                if (SUB == 3) mf_acc += (k == mf_kdup) ? 2.0 * (double)un : (double)un;
            }
        }
        // G of plane k+1 straight into the registers plane k has finished
        // with (no copy of a load in flight at the loop's back edge)
        gload(k + 1, gu, gv, gw);
        // plane k+2 into the slot of plane k-2, then prefetch plane k+3
        if (k + 2 <= kend) {
            store((r0 + 3) & 3, k + 2);
            if (k + 3 <= kend) fetch(k + 3);
        }
        __syncthreads();
    }
    if (SUB == 3) {  // one partial per workgroup, wave sums in a fixed order
        __shared__ double mf_sh[16];
        double mf_w = 0.0;
        if (active) {
            if (mf_mode == 1) mf_w = 0.5 * ((j > 2 ? m.dyg[j] : 0.0) + (j < nyg - 1 ? m.dyg[j + 1] : 0.0));
            else mf_w = j == ny / 2 + 1 ? 1.0 : 0.0;
        }
        double v = mf_w * mf_acc;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        const int wv = tid >> 6;
        if ((tid & 63) == 0) mf_sh[wv] = v;
        __syncthreads();
        if (tid == 0) {
            double r = 0.0;
            for (int q = 0; q < (RK_TX * TY) / 64; q++) r += mf_sh[q];
            mfpart[blockIdx.x] = r;
        }
    }
}

// ---------------------------------------------------------------------------
// RHS only (compute_rhs_u/v/w as separate outputs on the interior ranges of
// the reference, equations.f90:29,191,356), used by the module-level entry
// points.  Same point arithmetic as the fused kernel.
// ---------------------------------------------------------------------------
template <typename S, typename R, bool NUT, bool ROT>
__global__ void __launch_bounds__(256)
k_rhs_only(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V,
           const S *__restrict__ W, const S *__restrict__ NT, double *__restrict__ Ru,
           double *__restrict__ Rv, double *__restrict__ Rw, const Scalars *__restrict__ sc,
           const double *__restrict__ ubar, double nu_d, double om_d, int full) {
    // covers i in 2..nxg-1, j in 2..nyg-1, k in 2..nzg-1 (union of the three
    // interior boxes, duplicates included, as the reference computes them)
    const int i = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 2 + blockIdx.z;
    if (i > d.nxg - 1 || j > d.nyg - 1) return;
    GAcc<S, R> a{U, V, W, NT, ubar, i, j, k, d.nx, d.nxg, d.nyg, d.ny};
    const bool du = i <= d.nx - 1 && Ru, dv = j <= d.ny - 1 && Rv, dw = k <= d.nz - 1 && Rw;
    R ru = 0, rv = 0, rw = 0;
    PM<R> pm;
    pm.set_ij(m, i, j);
    pm.set_k(m, k);
    rhs_point<R, NUT, ROT>(a, pm, (R)nu_d, (R)om_d, (R)sc->dPdx, (R)sc->dPdz, du, dv, dw, ru, rv, rw);
    // full: arrays in the field layouts (ghosts included, left untouched);
    // otherwise the reference's interior-only rhs arrays
    if (full) {
        if (du) Ru[i3(d.nx, d.nyg, i, j, k)] = (double)ru;
        if (dv) Rv[i3(d.nxg, d.ny, i, j, k)] = (double)rv;
        if (dw) Rw[i3(d.nxg, d.nyg, i, j, k)] = (double)rw;
        return;
    }
    if (du) Ru[i3(d.nx - 2, d.nyg - 2, i - 1, j - 1, k - 1)] = (double)ru;
    if (dv) Rv[i3(d.nxg - 2, d.ny - 2, i - 1, j - 1, k - 1)] = (double)rv;
    if (dw) Rw[i3(d.nxg - 2, d.nyg - 2, i - 1, j - 1, k - 1)] = (double)rw;
}

// ---------------------------------------------------------------------------
// Ghost cells (boundary_conditions.f90:20-75, one-rank intended semantics).
// Each ghost value is written once, directly from unique cells, so the three
// phases of the reference (periodic x, periodic z, walls in y) collapse into
// one launch with no ordering hazards:
//   interior rows in y: periodic image of (i,k);
//   wall rows (j = 1, n2): Dirichlet / nonzero Neumann applied to the periodic
//   image of the first interior row, with alpha taken at the ghost's own (i,k)
//   exactly as the reference applies its y condition after the copies.
// field: 0 = U (x faces, z centres), 1 = V (y faces), 2 = W (z faces),
//        3 = cell-centred scalar (nu_t) with Neumann (0) / Dirichlet (1) walls.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int img_face(int i, int n) {  // faces 1..n, unique 2..n-1
    return i == 1 ? n - 1 : (i == n ? 2 : i);
}
__device__ __forceinline__ int img_ctr(int i, int ng) {  // centres 1..ng, unique 2..ng-2
    return i == 1 ? ng - 2 : (i == ng - 1 ? 2 : (i == ng ? 3 : i));
}

// ghost element g of one field (ordering: wall rows, z ghost planes, x ghost
// columns); wall_mode: 0 Dirichlet (negate / zero), 1 nonzero Neumann with
// alpha, 2 Neumann (copy)
template <typename S>
__device__ __forceinline__ void ghost_one(const Dims &d, S *__restrict__ F, int field, int wall_mode,
                                          const double *__restrict__ alpha, double dyg_bot, double dyg_top,
                                          int g) {
    int n1, n2, n3;
    bool xf, zf, yf;
    if (field == 0) { n1 = d.nx; n2 = d.nyg; n3 = d.nzg; xf = true; yf = false; zf = false; }
    else if (field == 1) { n1 = d.nxg; n2 = d.ny; n3 = d.nzg; xf = false; yf = true; zf = false; }
    else if (field == 2) { n1 = d.nxg; n2 = d.nyg; n3 = d.nz; xf = false; yf = false; zf = true; }
    else { n1 = d.nxg; n2 = d.nyg; n3 = d.nzg; xf = false; yf = false; zf = false; }
    const int nxgh = xf ? 2 : 3;              // ghost columns in x
    const int nzgh = zf ? 2 : 3;              // ghost planes in z
    // 32-bit index arithmetic: a field has far fewer than 2^31 ghosts
    const int nY = n1 * 2 * n3;  // wall rows
    const int nZ = n1 * (n2 - 2) * nzgh;
    const int nX = nxgh * (n2 - 2) * (n3 - nzgh);
    int i, j, k;
    if (g < nY) {
        const int q = g / n1;
        i = 1 + (g - q * n1);
        k = 1 + (q >> 1);
        j = (q & 1) == 0 ? 1 : n2;
    } else if ((g -= nY) < nZ) {
        const int q = g / n1, z = q / (n2 - 2);
        i = 1 + (g - q * n1);
        j = 2 + (q - z * (n2 - 2));
        k = zf ? (z == 0 ? 1 : n3) : (z == 0 ? 1 : n3 - 2 + z);
    } else if ((g -= nZ) < nX) {
        const int q = g / nxgh, kq = q / (n2 - 2);
        const int c = g - q * nxgh;
        j = 2 + (q - kq * (n2 - 2));
        k = 2 + kq;
        i = xf ? (c == 0 ? 1 : n1) : (c == 0 ? 1 : n1 - 2 + c);
    } else {
        return;
    }
    const int ii = xf ? img_face(i, n1) : img_ctr(i, n1);
    const int kk = zf ? img_face(k, n3) : img_ctr(k, n3);
    S val;
    if (j == 1 || j == n2) {
        if (yf) {
            val = (S)0;  // V at the walls (boundary_conditions.f90:264-265)
        } else {
            const int jin = j == 1 ? 2 : n2 - 1;
            S src = F[i3(n1, n2, ii, jin, kk)];
            if (wall_mode == 0) val = -src;
            else if (wall_mode == 2) val = src;
            else {
                const int wl = j == 1 ? 1 : 2;
                const double dd = j == 1 ? dyg_bot : dyg_top;
                val = (S)((double)src - dd * alpha[i3(n1, 2, i, wl, k)]);
            }
        }
    } else {
        val = F[i3(n1, n2, ii, j, kk)];
    }
    F[i3(n1, n2, i, j, k)] = val;
}

int64_t ghost_count(const Dims &d, int field) {
    int n1, n2, n3, nxgh, nzgh;
    if (field == 0) { n1 = d.nx; n2 = d.nyg; n3 = d.nzg; nxgh = 2; nzgh = 3; }
    else if (field == 1) { n1 = d.nxg; n2 = d.ny; n3 = d.nzg; nxgh = 3; nzgh = 3; }
    else if (field == 2) { n1 = d.nxg; n2 = d.nyg; n3 = d.nz; nxgh = 3; nzgh = 2; }
    else { n1 = d.nxg; n2 = d.nyg; n3 = d.nzg; nxgh = 3; nzgh = 3; }
    return (int64_t)n1 * 2 * n3 + (int64_t)n1 * (n2 - 2) * nzgh + (int64_t)nxgh * (n2 - 2) * (n3 - nzgh);
}

// ghosts of up to three fields in one launch: the thread range is split
// into consecutive segments, one per field (each ghost is written once from
// unique cells, so the fields are independent)
struct GhostSet {
    void *F[3];
    int field[3], mode[3];
    const double *alpha[3];
    int64_t end[3];  // cumulative counts
    int n;
};
template <typename S>
__global__ void k_ghosts(Dims d, GhostSet gs, double dyg_bot, double dyg_top, int64_t cs) {
    // the argument set is only read (a copy written with a runtime index
    // would live in scratch memory); the case offset is applied at use
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t beg = 0;
    for (int f = 0; f < gs.n; f++) {
        if (g < gs.end[f]) {
            ghost_one<S>(d, case_ptr((S *)gs.F[f], cs), gs.field[f], gs.mode[f], case_ptr(gs.alpha[f], cs), dyg_bot,
                         dyg_top, (int)(g - beg));
            return;
        }
        beg = gs.end[f];
    }
}

// ---------------------------------------------------------------------------
// project_velocity for V and W (projection.f90:213-223); U is projected in
// the inverse-x-FFT kernel that produces the same row of phi.
// phi holds unique cells: phi[((k-2)*nyp + (j-2))*nxp + (i-2)].
// ---------------------------------------------------------------------------
// Each thread marches over a chunk of kc planes at fixed (i, j), carrying
// phi(k) from one plane to the next, so phi streams through once; the
// phi(j+1) neighbour is the row the next thread row loads (cache hit).
// ---------------------------------------------------------------------------
// Block reductions (wavefront 64, LDS across wavefronts), fixed order so that
// results are reproducible run to run.
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ T block_reduce(T v, Op op, T *sh) {
    // linear thread id over 1-D or 2-D blocks; result valid in thread (0,0)
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int nthr = blockDim.x * blockDim.y;
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_down(v, o, 64));
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    const int nw = (nthr + 63) >> 6;
    if (wid == 0) {
        v = lane < nw ? sh[lane] : sh[0];
        for (int o = 32; o > 0; o >>= 1) {
            T other = __shfl_down(v, o, 64);
            if (lane + o < nw) v = op(v, other);
        }
    }
    __syncthreads();
    return v;
}
struct OpAdd { __device__ double operator()(double a, double b) const { return a + b; } };
struct OpMax { __device__ double operator()(double a, double b) const { return fmax(a, b); } };
struct OpMin { __device__ double operator()(double a, double b) const { return fmin(a, b); } };

template <typename S, typename P, int KM, bool GC>
__global__ void __launch_bounds__(256)
k_project_vw(Dims d, Metrics m, const P *__restrict__ phi, S *__restrict__ V,
             S *__restrict__ W, int tiles_x, int tiles_y, int ntiles, int kc, double *__restrict__ dtv,
             double *__restrict__ dtw, int64_t cs) {
    phi = case_ptr(phi, cs); V = case_ptr(V, cs); W = case_ptr(W, cs);
    dtv = case_ptr(dtv, cs); dtw = case_ptr(dtw, cs);
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tc = t / (tiles_x * tiles_y);
    const int i = 2 + tx * 64 + (threadIdx.x & 63);
    const int j = 2 + ty * 4 + (threadIdx.x >> 6);
    const int kb = 2 + tc * kc, ke = min(kb + kc, d.nzg - 1);  // planes kb..ke-1
    const bool act = i <= d.nx - 1 && j <= d.nyg - 1;
    // compute_dt partials: min |v| of this thread's column (one row j), min
    // |w| per plane for the first KM planes; planes beyond KM divide per value
    double vmin = INFINITY, wmin[KM];
#pragma unroll
    for (int q = 0; q < KM; q++) wmin[q] = INFINITY;
    double lw_direct = 0.0;
    if (act) {
        const int nxp = d.nxp, nyp = d.nyp;
        const int64_t pps = (int64_t)nyp * nxp;  // phi plane stride
        const P *pcol = phi + (int64_t)(j - 2) * nxp + (i - 2);
        const bool dov = j <= d.ny - 1;
        const double ry = m.rdyc[j];
        const int64_t vs = (int64_t)d.nxg * d.ny, ws = (int64_t)d.nxg * d.nyg;
        S *vp = V + i3(d.nxg, d.ny, i, j, kb);
        S *wp = W + i3(d.nxg, d.nyg, i, j, kb);
        double pc = (double)pcol[(int64_t)(kb - 2) * pps];
        for (int k = kb; k < ke; k++) {
            const int kn = (k + 1 == d.nzg - 1) ? 2 : k + 1;  // periodic image of phi(:,:,nz)
            const double pn = (double)pcol[(int64_t)(kn - 2) * pps];
            if (dov) {
                const double pj = (double)pcol[(int64_t)(k - 2) * pps + nxp];
                const S vn = (S)((double)*vp - (pj - pc) * ry);
                *vp = vn;
                vmin = fmin(vmin, fabs((double)vn));
                if (GC) {  // periodic ghost copies of V (its wall faces stay zero)
                    const int gi = ghost_of_ctr(i, d.nxg), gk = ghost_of_ctr(k, d.nzg);
                    if (gi | gk) put_ghost_copies(V, d.nxg, d.ny, i, j, k, vn, gi, 0, gk);
                }
            }
            const S wn = (S)((double)*wp - (pn - pc) * m.rdzc[k]);
            *wp = wn;
            if (GC) {  // periodic and Dirichlet-wall ghost copies of W
                const int gi = ghost_of_ctr(i, d.nxg), gk = ghost_of_face(k, d.nz);
                const int gj = j == 2 ? 1 : (j == d.nyg - 1 ? d.nyg : 0);
                if (gi | gj | gk) put_ghost_copies(W, d.nxg, d.nyg, i, j, k, wn, gi, gj, gk);
            }
            if (dtw) {
                const int q = k - kb;
                if (q < KM) {
#pragma unroll
                    for (int u = 0; u < KM; u++)
                        if (u == q) wmin[u] = fmin(wmin[u], fabs((double)wn));
                } else {
                    lw_direct = fmax(lw_direct, m.dzc[k] / fabs((double)wn));
                }
            }
            pc = pn;
            vp += vs;
            wp += ws;
        }
    }
    if (dtv) {
        // spacing/|u| is largest at the smallest |u| of equal spacing
        // (correctly rounded division is monotone): minima over each wave
        // (one row j per wave) and over the four waves (the same planes),
        // then one division per row and per plane, by thread 0
        __shared__ double shv[4], shd[4], shw[KM][4];
        for (int o = 32; o > 0; o >>= 1) {
            vmin = fmin(vmin, __shfl_down(vmin, o, 64));
            lw_direct = fmax(lw_direct, __shfl_down(lw_direct, o, 64));
#pragma unroll
            for (int u = 0; u < KM; u++) wmin[u] = fmin(wmin[u], __shfl_down(wmin[u], o, 64));
        }
        const int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            shv[wv] = (j <= d.ny - 1 && vmin < INFINITY) ? m.dyc[j] / vmin : 0.0;
            shd[wv] = lw_direct;
#pragma unroll
            for (int u = 0; u < KM; u++) shw[u][wv] = wmin[u];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double lv = fmax(fmax(shv[0], shv[1]), fmax(shv[2], shv[3]));
            double lw = fmax(fmax(shd[0], shd[1]), fmax(shd[2], shd[3]));
            for (int u = 0; u < KM && kb + u < ke; u++) {
                const double mn = fmin(fmin(shw[u][0], shw[u][1]), fmin(shw[u][2], shw[u][3]));
                if (mn < INFINITY) lw = fmax(lw, m.dzc[kb + u] / mn);
            }
            dtv[blockIdx.x] = lv;
            dtw[blockIdx.x] = lw;
        }
    }
}


// workgroup tile of the dt reduction below: 64 x 8 (x, y) by 16 z planes
constexpr int MF_TX = 64, MF_TY = 8, MF_KC = 16;

// mass_flow.f90:116-123: partial sums of the U trapezoid over k = 2..nzg-1
// (the unique planes plus the duplicate plane nzg-1, which the reference
// includes).
// one thread per (i, k) column (i = 2..nx-1, k = 2..nzg-1) marching the rows
// j = 2..nyg-1, each U(i, j, k) loaded once and weighted by its share of the
// trapezoid: 0.5 (dyg(j) + dyg(j+1)), one-sided on the first and last row
// (0.5 dyg(3), 0.5 dyg(nyg-1)); one deterministic partial per 256-thread
// block.  Rows are unrolled by MF_UNROLL so that many loads are in flight.
constexpr int MF_UNROLL = 8;
template <typename S>
__global__ void __launch_bounds__(256)
k_massflow_u_partial(Dims d, Metrics m, const S *__restrict__ U, const double *__restrict__ ubar,
                     double *__restrict__ part) {
    __shared__ double sh[8];
    const int ni = d.nx - 2, nk = d.nzg - 2, nyg = d.nyg;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (g < (int64_t)ni * nk) {
        const int i = 2 + (int)(g % ni), k = 2 + (int)(g / ni);
        const S *col = U + i3(d.nx, nyg, i, 1, k);  // row j at col[j - 1] * nx
        auto wt = [&](int j) {
            return 0.5 * ((j > 2 ? m.dyg[j] : 0.0) + (j < nyg - 1 ? m.dyg[j + 1] : 0.0));
        };
        int j = 2;
        for (; j + MF_UNROLL - 1 <= nyg - 1; j += MF_UNROLL) {
            double v[MF_UNROLL];
#pragma unroll
            for (int q = 0; q < MF_UNROLL; q++) v[q] = (double)col[(int64_t)(j + q - 1) * d.nx];
#pragma unroll
            for (int q = 0; q < MF_UNROLL; q++) acc += wt(j + q) * v[q];
        }
        for (; j <= nyg - 1; j++) acc += wt(j) * (double)col[(int64_t)(j - 1) * d.nx];
    }
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// final: Int_U over k = 2..nzg-1, i.e. the unique planes plus the duplicate
// plane nzg-1 (= plane 2 after the periodic copy, mass_flow.f90:116)
__global__ void k_massflow_u_final(const double *__restrict__ part, int nparts,
                                   Scalars *sc, int apply, int64_t cs) {
    __shared__ double sh[8];
    part = case_ptr(part, cs);
    sc = case_ptr(sc, cs);
    double acc = 0.0;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += part[q];
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) {
        const double total = acc + sc->qbar_u;
        const double mean = total / sc->norm_u;
        sc->pad[0] = mean;
        if (apply) {
            const double corr = sc->qflow_x0 - mean;  // mass_flow.f90:130
            sc->corr = corr;
            sc->dPdx = corr / sc->dt;  // time_integration.f90:156
        }
    }
}

template <typename S>
__global__ void __launch_bounds__(256)
k_add_corr_u(Dims d, S *__restrict__ U, const Scalars *__restrict__ sc) {
    const int64_t n = (int64_t)(d.nx - 2) * (d.nyg - 2) * d.nzp;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int i = 2 + (int)(g % (d.nx - 2));
    const int j = 2 + (int)((g / (d.nx - 2)) % (d.nyg - 2));
    const int k = 2 + (int)(g / ((int64_t)(d.nx - 2) * (d.nyg - 2)));
    const int64_t q = i3(d.nx, d.nyg, i, j, k);
    U[q] = (S)((double)U[q] + sc->corr);
}

// compute_dPx_for_constant_Ucenterline (mass_flow.f90:192-209): the sum of
// U(1:nx-2, jc, 1:nzg-2) with jc = ny/2 + 1 (integer division); the k range
// holds the periodic image k = 1 of plane nzg-2 besides that plane, as the
// reference's does.  One thread per (i, k), one partial per block.
template <typename S>
__global__ void __launch_bounds__(256)
k_centerline_partial(Dims d, const S *__restrict__ U, int jc, double *__restrict__ part) {
    __shared__ double sh[8];
    const int ni = d.nx - 2, nk = d.nzg - 2;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (g < (int64_t)ni * nk) {
        const int i = 1 + (int)(g % ni), k = 1 + (int)(g / ni);
        acc = (double)U[i3(d.nx, d.nyg, i, jc, k)];
    }
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// compute_mean_mass_flow_V and compute_dPy_for_constant_mass_flow
// (mass_flow.f90:61-81, 147-174): one thread per (i, k) column, i = 2..nxg-1,
// k = 2..nzg-1 (the duplicate column and plane included, as in the
// reference), marching the faces j = 1..ny; each V loaded once with its
// trapezoid weight 0.5 ((y(j)-y(j-1)) + (y(j+1)-y(j))), one-sided at the
// walls.  y is 1-based.
template <typename S>
__global__ void __launch_bounds__(256)
k_massflow_v_partial(Dims d, const double *__restrict__ y, const S *__restrict__ V, double *__restrict__ part) {
    __shared__ double sh[8];
    const int ni = d.nxg - 2, nk = d.nzg - 2, ny = d.ny;
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double acc = 0.0;
    if (g < (int64_t)ni * nk) {
        const int i = 2 + (int)(g % ni), k = 2 + (int)(g / ni);
        const S *col = V + i3(d.nxg, ny, i, 1, k);
#pragma unroll 4
        for (int j = 1; j <= ny; j++) {
            const double w = 0.5 * ((j > 1 ? y[j] - y[j - 1] : 0.0) + (j < ny ? y[j + 1] - y[j] : 0.0));
            acc += w * (double)col[(int64_t)(j - 1) * d.nxg];
        }
    }
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// final step of the centreline and bulk-V reductions (partials in block
// order).  MF_CENTERLINE_MEAN: mean of the centreline plane (count values;
// the mixed path adds the stored profile ubar(jc) to each); MF_CENTERLINE
// also the correction Ucenterline_x - mean, dPdx = corr/dt
// (time_integration.f90:158-162).
// MF_V / MF_V_APPLY: bulk V over norm_v, and for MF_V_APPLY the correction
// Qflow_y_0 - bulk, dPdy = corr/dt (time_integration.f90:166-169).
enum { MF_CENTERLINE_MEAN = 0, MF_CENTERLINE = 1, MF_V = 2, MF_V_APPLY = 3 };
__global__ void k_flow_final(const double *__restrict__ part, int nparts, Scalars *sc,
                             const double *__restrict__ ubar, int jc, double count, int mode, int64_t cs) {
    __shared__ double sh[8];
    part = case_ptr(part, cs);
    sc = case_ptr(sc, cs);
    ubar = case_ptr(ubar, cs);
    double acc = 0.0;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += part[q];
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) {
        if (mode == MF_CENTERLINE || mode == MF_CENTERLINE_MEAN) {
            const double mean = (acc + (ubar ? count * ubar[jc] : 0.0)) / count;
            sc->pad[0] = mean;
            if (mode == MF_CENTERLINE) {
                const double corr = sc->ucl - mean;
                sc->corr = corr;
                sc->dPdx = corr / sc->dt;
            }
        } else {
            const double mean = acc / sc->norm_v;
            sc->pad[1] = mean;
            if (mode == MF_V_APPLY) {
                const double corr = sc->qflow_y0 - mean;
                sc->corr_v = corr;
                sc->dPdy = corr / sc->dt;
            }
        }
    }
}

// V(2:nxg-1, 2:ny-1, 2:nzg-1) += corr_v (time_integration.f90:168); the wall
// rows keep their values
template <typename S>
__global__ void __launch_bounds__(256)
k_add_corr_v(Dims d, S *__restrict__ V, const Scalars *__restrict__ sc) {
    const int ni = d.nxg - 2, nj = d.ny - 2;
    const int64_t n = (int64_t)ni * nj * (d.nzg - 2);
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int i = 2 + (int)(g % ni);
    const int j = 2 + (int)((g / ni) % nj);
    const int k = 2 + (int)(g / ((int64_t)ni * nj));
    const int64_t q = i3(d.nxg, d.ny, i, j, k);
    V[q] = (S)((double)V[q] + sc->corr_v);
}

// time_integration.f90:188-199.  max over cells of spacing/|u| equals, per
// line of constant spacing, spacing / min|u| (correctly rounded division is
// monotone), so the kernel only takes minima: of |U| per i, |V| per j, |W|
// per k, over the reference's ranges, reduced per tile into partial arrays and
// then per line (no atomics, order independent).  Duplicate periodic planes
// hold copies of unique ones; the wall row V(:,ny,:) is read.
template <typename S>
__global__ void __launch_bounds__(MF_TX *MF_TY)
k_dt_min(Dims d, const S *__restrict__ U, const S *__restrict__ V, const S *__restrict__ W,
         const double *__restrict__ ubar, double *__restrict__ pmin, int tiles_x, int tiles_y,
         int nch) {
    __shared__ double shu[MF_TY][MF_TX];
    __shared__ double shv[MF_TY][MF_TX];
    __shared__ double shw[MF_KC][MF_TY];
    const int t = blockIdx.x;
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tc = t / (tiles_x * tiles_y);
    const int i = 2 + tx * MF_TX + threadIdx.x;
    const int j = 2 + ty * MF_TY + threadIdx.y;
    const int kb = 2 + tc * MF_KC;
    const bool act = i <= d.nx - 1 && j <= d.nyg - 1;
    const double ub = (act && ubar) ? ubar[j] : 0.0;
    double mu = INFINITY, mv = INFINITY;
#pragma unroll 4
    for (int kk = 0; kk < MF_KC; kk++) {
        const int k = kb + kk;
        double mwk = INFINITY;
        if (act && k <= d.nzg - 2) {
            mu = fmin(mu, fabs((double)U[i3(d.nx, d.nyg, i, j, k)] + ub));
            mv = fmin(mv, fabs((double)V[i3(d.nxg, d.ny, i, j, k)]));  // j <= nyg-1 = ny
            mwk = fabs((double)W[i3(d.nxg, d.nyg, i, j, k)]);
        }
        for (int o = 32; o > 0; o >>= 1) mwk = fmin(mwk, __shfl_down(mwk, o, 64));
        if (threadIdx.x == 0) shw[kk][threadIdx.y] = mwk;
    }
    shu[threadIdx.y][threadIdx.x] = mu;
    shv[threadIdx.y][threadIdx.x] = mv;
    __syncthreads();
    // per-line minima of this tile -> partial arrays (no atomics)
    const int nyz = tiles_y * nch, nxz = tiles_x * nch, nxy = tiles_x * tiles_y;
    double *pu = pmin;                                   // [nyz][nxp]
    double *pv = pu + (size_t)nyz * d.nxp;               // [nxz][nyp]
    double *pw = pv + (size_t)nxz * d.nyp;               // [nxy][nzp]
    const int tid = threadIdx.x + MF_TX * threadIdx.y;
    if (tid < MF_TX) {
        double r = INFINITY;
        for (int q = 0; q < MF_TY; q++) r = fmin(r, shu[q][tid]);
        const int ii = tx * MF_TX + tid;
        if (ii < d.nxp) pu[(size_t)(ty + tiles_y * tc) * d.nxp + ii] = r;
    } else if (tid < MF_TX + MF_TY) {
        const int q = tid - MF_TX;
        double r = INFINITY;
        for (int c = 0; c < MF_TX; c++) r = fmin(r, shv[q][c]);
        const int jj = ty * MF_TY + q;
        if (jj < d.nyp) pv[(size_t)(tx + tiles_x * tc) * d.nyp + jj] = r;
    } else if (tid < MF_TX + MF_TY + MF_KC) {
        const int q = tid - MF_TX - MF_TY;
        double r = INFINITY;
        for (int c = 0; c < MF_TY; c++) r = fmin(r, shw[q][c]);
        const int kk = tc * MF_KC + q;
        if (kk < d.nzp) pw[(size_t)(tx + tiles_x * ty) * d.nzp + kk] = r;
    }
}

// one workgroup per line: minimum over the tiles of the partial arrays
__global__ void __launch_bounds__(256)
k_dt_lines(Dims d, const double *__restrict__ pmin, double *__restrict__ mins, int tiles_x,
           int tiles_y, int nch) {
    __shared__ double sh[8];
    const int nyz = tiles_y * nch, nxz = tiles_x * nch, nxy = tiles_x * tiles_y;
    const double *pu = pmin, *pv = pu + (size_t)nyz * d.nxp, *pw = pv + (size_t)nxz * d.nyp;
    const int q = blockIdx.x;
    double r = INFINITY;
    if (q < d.nxp) {
        for (int t = threadIdx.x; t < nyz; t += 256) r = fmin(r, pu[(size_t)t * d.nxp + q]);
    } else if (q < d.nxp + d.nyp) {
        const int jj = q - d.nxp;
        for (int t = threadIdx.x; t < nxz; t += 256) r = fmin(r, pv[(size_t)t * d.nyp + jj]);
    } else {
        const int kk = q - d.nxp - d.nyp;
        for (int t = threadIdx.x; t < nxy; t += 256) r = fmin(r, pw[(size_t)t * d.nzp + kk]);
    }
    r = block_reduce(r, OpMin(), sh);
    if (threadIdx.x == 0) mins[q] = r;
}

// dt = min(CFL*max(spacing/min|u|) over the three components, viscous, 5e-3)
__global__ void k_dt_final(Dims d, const double *__restrict__ dxc, const double *__restrict__ dyc,
                           const double *__restrict__ dzc, const double *__restrict__ mins,
                           Scalars *sc, double cfl, double dt_vis) {
    __shared__ double sh[8];
    const double *mu = mins, *mv = mins + d.nxp, *mw = mv + d.nyp;  // index i-2, j-2, k-2
    double lu = 0.0, lv = 0.0, lw = 0.0;
    // U: i = 2..nxg-1 (= nx); column nx is the periodic copy of column 2
    for (int i = 2 + threadIdx.x; i <= d.nxg - 1; i += blockDim.x)
        lu = fmax(lu, dxc[i] / (i == d.nx ? mu[0] : mu[i - 2]));
    for (int j = 2 + threadIdx.x; j <= d.nyg - 1; j += blockDim.x) lv = fmax(lv, dyc[j] / mv[j - 2]);
    // W: k = 2..nzg-1 (= nz); plane nz is the periodic copy of plane 2
    for (int k = 2 + threadIdx.x; k <= d.nzg - 1; k += blockDim.x)
        lw = fmax(lw, dzc[k] / (k == d.nz ? mw[0] : mw[k - 2]));
    lu = block_reduce(lu, OpMax(), sh);
    lv = block_reduce(lv, OpMax(), sh);
    lw = block_reduce(lw, OpMax(), sh);
    if (threadIdx.x == 0) {
        const double dc = fmin(fmin(cfl * lu, cfl * lv), cfl * lw);
        double dt = fmin(dc, dt_vis);
        dt = fmin(dt, 5e-3);
        if (cfl < 0) dt = -cfl;
        sc->dt = dt;
        sc->dt_next = dt;
        sc->dt_conv[0] = lu; sc->dt_conv[1] = lv; sc->dt_conv[2] = lw;
    }
}

// compute_dt of the state a step leaves, from the partial maxima its last
// projection kernels took (DtParts) and V's wall row j = ny, which those
// kernels do not write (the reference's range j = 2..nyg-1 includes it).
// Maxima of correctly rounded divisions: the same dt as k_dt_final.
constexpr int DTF_NB = 128;  // workgroups of the first level
__global__ void __launch_bounds__(256)
k_dt_fused_mid(Dims d, const double *__restrict__ pu, int nu_, const double *__restrict__ pv,
               const double *__restrict__ pw, int nvw, const double *__restrict__ dyc, const void *__restrict__ Vp,
               int vfp32, double *__restrict__ mid, int64_t cs) {
    __shared__ double sh[8];
    pu = case_ptr(pu, cs); pv = case_ptr(pv, cs); pw = case_ptr(pw, cs);
    Vp = case_ptr(Vp, cs); mid = case_ptr(mid, cs);
    const int g = blockIdx.x * blockDim.x + threadIdx.x, gs = gridDim.x * blockDim.x;
    double lu = 0.0, lv = 0.0, lw = 0.0, vmin = INFINITY;
    for (int q = g; q < nu_; q += gs) lu = fmax(lu, pu[q]);
    for (int q = g; q < nvw; q += gs) { lv = fmax(lv, pv[q]); lw = fmax(lw, pw[q]); }
    const int ni = d.nxg - 2, nk = d.nzg - 2, ny = d.ny;
    for (int q = g; q < ni * nk; q += gs) {  // V's wall row: one spacing, its smallest |v|
        const int i = 2 + q % ni, k = 2 + q / ni;
        const int64_t e = i3(d.nxg, ny, i, ny, k);
        const double v = vfp32 ? (double)((const float *)Vp)[e] : ((const double *)Vp)[e];
        vmin = fmin(vmin, fabs(v));
    }
    lu = block_reduce(lu, OpMax(), sh);
    lv = block_reduce(lv, OpMax(), sh);
    lw = block_reduce(lw, OpMax(), sh);
    vmin = block_reduce(vmin, OpMin(), sh);
    if (threadIdx.x == 0) {
        mid[4 * blockIdx.x + 0] = lu;
        mid[4 * blockIdx.x + 1] = lv;
        mid[4 * blockIdx.x + 2] = lw;
        mid[4 * blockIdx.x + 3] = vmin;
    }
}
__global__ void __launch_bounds__(256)
k_dt_fused_final(Dims d, const double *__restrict__ mid, const double *__restrict__ dyc, Scalars *sc, double cfl,
                 double dt_vis, int64_t cs) {
    __shared__ double sh[8];
    mid = case_ptr(mid, cs);
    sc = case_ptr(sc, cs);
    if (cs) dt_vis = sc->dt_vis;
    double lu = 0.0, lv = 0.0, lw = 0.0, vmin = INFINITY;
    for (int q = threadIdx.x; q < DTF_NB; q += blockDim.x) {
        lu = fmax(lu, mid[4 * q]);
        lv = fmax(lv, mid[4 * q + 1]);
        lw = fmax(lw, mid[4 * q + 2]);
        vmin = fmin(vmin, mid[4 * q + 3]);
    }
    lu = block_reduce(lu, OpMax(), sh);
    lv = block_reduce(lv, OpMax(), sh);
    lw = block_reduce(lw, OpMax(), sh);
    vmin = block_reduce(vmin, OpMin(), sh);
    if (threadIdx.x == 0) {
        lv = fmax(lv, dyc[d.ny] / vmin);
        const double dc = fmin(fmin(cfl * lu, cfl * lv), cfl * lw);
        double dt = fmin(dc, dt_vis);
        dt = fmin(dt, 5e-3);
        if (cfl < 0) dt = -cfl;
        sc->dt_next = dt;
        sc->dt_conv[0] = lu; sc->dt_conv[1] = lv; sc->dt_conv[2] = lw;
    }
}
__global__ void k_dt_take_next(Scalars *sc, int64_t cs) {
    sc = case_ptr(sc, cs);
    sc->dt = sc->dt_next;
}

size_t dt_fused_mid_doubles() { return 4 * (size_t)DTF_NB; }
template <typename S>
void launch_dt_fused_final(const Dims &d, const DtParts &p, int nu_, int nvw, const double *dyc, const S *V,
                           Scalars *sc, double cfl, double dt_vis, hipStream_t st) {
    double *mid = p.pw + nvw;  // after the partial arrays (dt_fused_mid_doubles of room)
    const int nb = launch_batch().n;
    const int64_t cs = launch_batch().cs;
    hipLaunchKernelGGL(k_dt_fused_mid, dim3(DTF_NB, nb), dim3(256), 0, st, d, p.pu, nu_, p.pv, p.pw, nvw, dyc,
                       (const void *)V, sizeof(S) == 4 ? 1 : 0, mid, cs);
    hipLaunchKernelGGL(k_dt_fused_final, dim3(1, nb), dim3(256), 0, st, d, mid, dyc, sc, cfl, dt_vis, cs);
}
void launch_dt_take_next(Scalars *sc, hipStream_t st) {
    hipLaunchKernelGGL(k_dt_take_next, dim3(1, launch_batch().n), dim3(1), 0, st, sc, launch_batch().cs);
}
template void launch_dt_fused_final<double>(const Dims &, const DtParts &, int, int, const double *,
                                            const double *, Scalars *, double, double, hipStream_t);
template void launch_dt_fused_final<float>(const Dims &, const DtParts &, int, int, const double *, const float *,
                                           Scalars *, double, double, hipStream_t);

// projection.f90:234-256: signed maximum (from 0) and max |div| per plane
template <typename S>
__global__ void __launch_bounds__(256)
k_div_partial(Dims d, const double *__restrict__ x, const double *__restrict__ y,
              const double *__restrict__ z, const S *__restrict__ U, const S *__restrict__ V,
              const S *__restrict__ W, double *__restrict__ part) {
    __shared__ double sh[8];
    const int k = 2 + blockIdx.x;  // 2..nzg-1
    double ms = 0.0, ma = 0.0;
    const int ni = d.nxg - 2, nj = d.nyg - 2;
    for (int q = threadIdx.x; q < ni * nj; q += blockDim.x) {
        const int i = 2 + q % ni, j = 2 + q / ni;
        const double dv = ((double)U[i3(d.nx, d.nyg, i, j, k)] - (double)U[i3(d.nx, d.nyg, i - 1, j, k)]) / (x[i] - x[i - 1]) +
                          ((double)V[i3(d.nxg, d.ny, i, j, k)] - (double)V[i3(d.nxg, d.ny, i, j - 1, k)]) / (y[j] - y[j - 1]) +
                          ((double)W[i3(d.nxg, d.nyg, i, j, k)] - (double)W[i3(d.nxg, d.nyg, i, j, k - 1)]) / (z[k] - z[k - 1]);
        ms = fmax(ms, dv);
        ma = fmax(ma, fabs(dv));
    }
    ms = block_reduce(ms, OpMax(), sh);
    ma = block_reduce(ma, OpMax(), sh);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = ms; part[2 * blockIdx.x + 1] = ma; }
}

// ---------------------------------------------------------------------------
// Constant-coefficient Smagorinsky (subgrid.f90:564-619) with compute_Sij
// (subgrid.f90:891-949) fused: one thread per unique cell computes the three
// diagonal strains, the 4 edge values of each off-diagonal strain and their
// centre average, |S| = sqrt(2 SijSij), and nu_t = (Cs Delta f)^2 |S|.
// Arithmetic in R (FP64, or FP32 in the mixed path), divisions by grid
// spacings as multiplications with the precomputed reciprocals.
// ---------------------------------------------------------------------------
template <typename S, typename R>
__global__ void __launch_bounds__(256)
k_smagorinsky(Dims d, Metrics m, const double *__restrict__ y, const double *__restrict__ yg,
              const double *__restrict__ delta, const S *__restrict__ U, const S *__restrict__ V,
              const S *__restrict__ W, S *__restrict__ NT, const Scalars *__restrict__ sc,
              const double *__restrict__ ubar, double cs, double nu, int damping, int tiles_x,
              int tiles_y, int ntiles, int64_t ccs) {
    U = case_ptr(U, ccs); V = case_ptr(V, ccs); W = case_ptr(W, ccs); NT = case_ptr(NT, ccs);
    sc = case_ptr(sc, ccs); ubar = case_ptr(ubar, ccs);
    if (ccs) nu = sc->nu;
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tz = t / (tiles_x * tiles_y);
    const int i = 2 + tx * 64 + (threadIdx.x & 63);
    const int j = 2 + ty * 4 + (threadIdx.x >> 6);
    const int k = 2 + tz;
    if (i > d.nx - 1 || j > d.nyg - 1) return;
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny;
    // U of the deviation storage plus its profile at row J (a select over
    // three profile values by J would be kept in scratch memory)
#define u(I, J, K) ((R)U[i3(nx, nyg, I, J, K)] + (ubar ? (R)ubar[J] : (R)0))
#define v(I, J, K) ((R)V[i3(nxg, ny, I, J, K)])
#define w(I, J, K) ((R)W[i3(nxg, nyg, I, J, K)])
    // 1/(x(i)-x(i-1)) = rdxf(i); 1/(xg(I)-xg(I-1)) = rdxc(I-1); likewise y, z
    const R rxf = (R)m.rdxf[i], ryf = (R)m.rdyf[j], rzf = (R)m.rdzf[k];
    const R rxc0 = (R)m.rdxc[i - 1], rxc1 = (R)m.rdxc[i];
    const R ryc0 = (R)m.rdyc[j - 1], ryc1 = (R)m.rdyc[j];
    const R rzc0 = (R)m.rdzc[k - 1], rzc1 = (R)m.rdzc[k];
    const R h = (R)0.5, q = (R)0.25;
    const R s11 = (u(i, j, k) - u(i - 1, j, k)) * rxf;
    const R s22 = (v(i, j, k) - v(i, j - 1, k)) * ryf;
    const R s33 = (w(i, j, k) - w(i, j, k - 1)) * rzf;
    // edge values at (I, J, K), I in {i, i+1}, J in {j, j+1}, K in {k, k+1}
    auto e4 = [&](int I, int J, int K, R ryc, R rxc) {
        return h * ((u(I - 1, J, K) - u(I - 1, J - 1, K)) * ryc + (v(I, J - 1, K) - v(I - 1, J - 1, K)) * rxc);
    };
    auto e5 = [&](int I, int J, int K, R rzc, R rxc) {
        return h * ((u(I - 1, J, K) - u(I - 1, J, K - 1)) * rzc + (w(I, J, K - 1) - w(I - 1, J, K - 1)) * rxc);
    };
    auto e6 = [&](int I, int J, int K, R rzc, R ryc) {
        return h * ((v(I, J - 1, K) - v(I, J - 1, K - 1)) * rzc + (w(I, J, K - 1) - w(I, J - 1, K - 1)) * ryc);
    };
    const R s12 = q * (e4(i, j, k, ryc0, rxc0) + e4(i + 1, j, k, ryc0, rxc1) + e4(i, j + 1, k, ryc1, rxc0) +
                       e4(i + 1, j + 1, k, ryc1, rxc1));
    const R s13 = q * (e5(i, j, k, rzc0, rxc0) + e5(i + 1, j, k, rzc0, rxc1) + e5(i, j, k + 1, rzc1, rxc0) +
                       e5(i + 1, j, k + 1, rzc1, rxc1));
    const R s23 = q * (e6(i, j, k, rzc0, ryc0) + e6(i, j + 1, k, rzc0, ryc1) + e6(i, j, k + 1, rzc1, ryc0) +
                       e6(i, j + 1, k + 1, rzc1, ryc1));
    R s = 0;
    s = s + (R)2 * s11 * s11;
    s = s + (R)2 * s22 * s22;
    s = s + (R)2 * s33 * s33;
    s = s + (R)4 * s12 * s12;
    s = s + (R)4 * s13 * s13;
    s = s + (R)4 * s23 * s23;
    double f = 1.0;
    if (damping) {
        const double utau_t = sqrt(sc->dPdx);
        if (j < nyg / 2) f = 1.0 - exp(-(yg[j] - y[1]) / 25.0 / (nu / utau_t));
        else f = 1.0 - exp(-(y[ny] - yg[j]) / 25.0 / (nu / utau_t));
    }
    const R c = (R)(cs * delta[j] * f);
    const R val = fmax((c * c) * sqrt(s), (R)0);
    NT[i3(nxg, nyg, i, j, k)] = (S)val;
#undef u
#undef v
#undef w
}

// ---------------------------------------------------------------------------
// Plane statistics (statistics.f90:43-71) in FP64: one block per y face jj.
// out[7*ny]: Umean,Vmean,Wmean,U2mean,V2mean,W2mean,UVmean (raw sums).
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256)
k_stats_planes(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V,
               const S *__restrict__ W, const double *__restrict__ ubar, double *__restrict__ part) {
    __shared__ double sh[8];
    const int jj = 1 + blockIdx.x;
    const int nb = gridDim.y, q0 = threadIdx.x + blockIdx.y * blockDim.x, qs = blockDim.x * nb;
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny, nz = d.nz, nzg = d.nzg;
    const double a = m.w0[jj], b = m.w1[jj];
    const double ua = ubar ? ubar[jj] : 0.0, ubb = ubar ? ubar[jj + 1] : 0.0;
    double su = 0, su2 = 0, sw = 0, sw2 = 0, sv = 0, sv2 = 0, suv = 0;
    for (int q = q0; q < (nx - 2) * (nzg - 2); q += qs) {
        const int i = 1 + q % (nx - 2), k = 1 + q / (nx - 2);
        const double t = a * ((double)U[i3(nx, nyg, i, jj, k)] + ua) + b * ((double)U[i3(nx, nyg, i, jj + 1, k)] + ubb);
        su += t; su2 += t * t;
    }
    for (int q = q0; q < (nxg - 2) * (nz - 2); q += qs) {
        const int i = 1 + q % (nxg - 2), k = 1 + q / (nxg - 2);
        const double t = a * (double)W[i3(nxg, nyg, i, jj, k)] + b * (double)W[i3(nxg, nyg, i, jj + 1, k)];
        sw += t; sw2 += t * t;
    }
    for (int q = q0; q < (nxg - 2) * (nzg - 2); q += qs) {
        const int i = 2 + q % (nxg - 2), k = 2 + q / (nxg - 2);
        const double vv = (double)V[i3(nxg, ny, i, jj, k)];
        sv += vv; sv2 += vv * vv;
        const int ii = i - 1;
        const double ta = a * ((double)U[i3(nx, nyg, ii, jj, k)] + ua) + b * ((double)U[i3(nx, nyg, ii, jj + 1, k)] + ubb);
        const double tb = a * ((double)U[i3(nx, nyg, ii + 1, jj, k)] + ua) + b * ((double)U[i3(nx, nyg, ii + 1, jj + 1, k)] + ubb);
        suv += vv * (0.5 * (ta + tb));
    }
    double vals[7] = {su, sv, sw, su2, sv2, sw2, suv};
    for (int c = 0; c < 7; c++) {
        double r = block_reduce(vals[c], OpAdd(), sh);
        if (threadIdx.x == 0) part[((size_t)c * ny + (jj - 1)) * nb + blockIdx.y] = r;
    }
}

// compute_statistics wall terms (statistics.f90:121-140): sums of U over
// (2:nx-1, j, 2:nzg-1) for j = 1, 2, nyg-1, nyg (blockIdx.x = 0..3; the mixed
// path adds the stored profile ubar(j) to each value) and of nu_t over
// (2:nxg-1, 1:2, 2:nzg-1) and (2:nxg-1, nyg-1:nyg, 2:nzg-1) (blockIdx.x = 4,
// 5).  blockIdx.y splits each sum; partials are stored in block order.  The
// first block also clears the NaN flag that k_nan_scan sets afterwards.
template <typename S>
__global__ void __launch_bounds__(256)
k_stats_walls(Dims d, const S *__restrict__ U, const S *__restrict__ NT, const double *__restrict__ ubar,
              double *__restrict__ part, double *__restrict__ flag) {
    __shared__ double sh[8];
    const int q = blockIdx.x, nb = gridDim.y;
    const int t0 = threadIdx.x + blockIdx.y * 256, ts = 256 * nb;
    double acc = 0.0;
    if (q < 4) {
        const int j = q == 0 ? 1 : (q == 1 ? 2 : (q == 2 ? d.nyg - 1 : d.nyg));
        const int ni = d.nx - 2, n = ni * (d.nzg - 2);
        const double ub = ubar ? ubar[j] : 0.0;
        for (int e = t0; e < n; e += ts) {
            const int i = 2 + e % ni, k = 2 + e / ni;
            acc += (double)U[i3(d.nx, d.nyg, i, j, k)] + ub;
        }
    } else {
        const int j0 = q == 4 ? 1 : d.nyg - 1;
        const int ni = d.nxg - 2, n = ni * 2 * (d.nzg - 2);
        for (int e = t0; e < n; e += ts) {
            const int i = 2 + e % ni, r = e / ni;
            acc += (double)NT[i3(d.nxg, d.nyg, i, j0 + (r & 1), 2 + (r >> 1))];
        }
    }
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) part[q * nb + blockIdx.y] = acc;
    if (q == 0 && blockIdx.y == 0 && threadIdx.x == 0) *flag = 0.0;
}

// Any(Isnan(U)), Any(Isnan(V)), Any(Isnan(W)) over the whole arrays, ghosts
// included (statistics.f90:165-167): blockIdx.y picks the field; a thread that
// meets a NaN stores 1 into the flag (every such store writes the same value)
template <typename S>
__global__ void __launch_bounds__(256)
k_nan_scan(const S *__restrict__ U, const S *__restrict__ V, const S *__restrict__ W, int64_t nu, int64_t nv,
           int64_t nw, double *__restrict__ flag) {
    const S *a = blockIdx.y == 0 ? U : (blockIdx.y == 1 ? V : W);
    const int64_t n = blockIdx.y == 0 ? nu : (blockIdx.y == 1 ? nv : nw);
    bool bad = false;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
        const S v = a[e];
        bad = bad || (v != v);
    }
    if (bad) *flag = 1.0;
}

// u_tau, Re_tau and the bulk flows (statistics.f90:121-158) from the plane
// sums (raw UV sums at the walls in out[6 ny], out[7 ny - 1]), the wall sums
// of k_stats_walls and the bulk means the mass-flow kernels left in
// pad[0], pad[1].  One thread; writes out[7 ny .. 7 ny + 3].
__global__ void k_stats_final(Dims d, const double *__restrict__ y, const double *__restrict__ yg,
                              const double *__restrict__ wpart, int nb, double *__restrict__ out,
                              const Scalars *__restrict__ sc, double nu) {
    const int ny = d.ny, nyg = d.nyg;
    double s[6];
    for (int q = 0; q < 6; q++) {
        double a = 0.0;
        for (int b = 0; b < nb; b++) a += wpart[q * nb + b];
        s[q] = a;
    }
    const double cu = (double)((d.nx - 2) * (d.nzg - 2)), cn = (double)((d.nxg - 2) * (d.nzg - 2));
    const double U1 = s[0] / cu, U2 = s[1] / cu, Un1 = s[2] / cu, Un2 = s[3] / cu;
    const double dUb = (U2 - U1) / (yg[2] - yg[1]);
    const double dUt = (Un2 - Un1) / (yg[nyg] - yg[nyg - 1]);
    const double ntb = s[4] / cn / 2.0, ntt = s[5] / cn / 2.0;
    const double UVb = out[6 * ny] / cn, UVt = out[6 * ny + ny - 1] / cn;
    const double Ly = y[ny] - y[1];
    const double utau = sqrt((UVt - UVb - (nu + ntt) * dUt + (nu + ntb) * dUb) / Ly);
    out[7 * ny + 0] = utau;
    out[7 * ny + 1] = utau * (y[ny] - y[1]) / 2.0 / nu;
    out[7 * ny + 2] = sc->pad[0];
    out[7 * ny + 3] = sc->pad[1];
}

// out[e] = sum of the STATS_NB block partials of statistic e, in block order
__global__ void k_stats_fold(const double *__restrict__ part, double *__restrict__ out, int n, int nb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double r = 0.0;
    for (int b = 0; b < nb; b++) r += part[(size_t)e * nb + b];
    out[e] = r;
}

// time-window accumulation of the plane statistics: acc[0..n) += out, acc[n] += 1
__global__ void k_stats_accum(const double *__restrict__ out, double *__restrict__ acc, int n) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) acc[q] += out[q];
    if (q == n) acc[n] += 1.0;
}

void launch_stats_accum(const double *out, double *acc, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_stats_accum, dim3((n + 1 + 255) / 256), dim3(256), 0, st, out, acc, n);
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline int tiles_x(const Dims &d) { return (d.nxp + 63) / 64; }
static inline int tiles_y(const Dims &d) { return (d.nyp + 3) / 4; }

int rhs_rk_tiles(const Dims &d, bool nut) {
    const int TYv = nut ? 4 : 8;
    return ((d.nxp + RK_TX - 1) / RK_TX) * ((d.nyp + TYv - 1) / TYv) * ((d.nzp + RK_KC - 1) / RK_KC);
}

template <typename S, typename R>
void launch_rhs_rk(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                   const S *NT, S *Uo, S *Vo, S *Wo, S *Gu, S *Gv, S *Gw,
                   const Scalars *sc, const double *ubar, double nu, double om, double a_new,
                   double a_g, int sub, hipStream_t st, double *mfpart, int mf_mode) {
    const bool nut = NT != nullptr, rot = fabs(om) > 1e-6;
    // 64x8 tiles without eddy viscosity, 64x4 with it (the ring of four
    // fields then stays below 64 KiB of LDS in FP64)
    const int TYv = nut ? 4 : 8;
    const int tx = (d.nxp + RK_TX - 1) / RK_TX, ty = (d.nyp + TYv - 1) / TYv;
    const int nch = (d.nzp + RK_KC - 1) / RK_KC, nt = tx * ty * nch;
    const size_t shm = (size_t)(nut ? 4 : 3) * RK_NSLOT * (RK_TX + 2) * (TYv + 2) * sizeof(R);
    const int nb = launch_batch().n;
    const int64_t cs = launch_batch().cs;
#define L(NUTB, ROTB, SUBV, TYB)                                                                           \
    hipLaunchKernelGGL((k_rhs_rk<S, R, NUTB, ROTB, SUBV, TYB>), dim3(nt, nb), dim3(RK_TX, TYB), shm, st, d, m, U, V, \
                       W, NT, Uo, Vo, Wo, Gu, Gv, Gw, sc, ubar, nu, om, a_new, a_g, tx, ty, nch, nt,  \
                       mfpart, mf_mode, cs)
    if (sub == 1) {
        if (nut) { if (rot) L(true, true, 1, 4); else L(true, false, 1, 4); }
        else { if (rot) L(false, true, 1, 8); else L(false, false, 1, 8); }
    } else if (!mfpart) {
        if (nut) { if (rot) L(true, true, 2, 4); else L(true, false, 2, 4); }
        else { if (rot) L(false, true, 2, 8); else L(false, false, 2, 8); }
    } else {
        if (nut) { if (rot) L(true, true, 3, 4); else L(true, false, 3, 4); }
        else { if (rot) L(false, true, 3, 8); else L(false, false, 3, 8); }
    }
#undef L
}

template <typename S, typename R>
void launch_rhs_only(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                     const S *NT, double *Ru, double *Rv, double *Rw, const Scalars *sc,
                     const double *ubar, double nu, double om, hipStream_t st, bool full) {
    dim3 grid((d.nxg - 2 + 63) / 64, (d.nyg - 2 + 3) / 4, d.nzg - 2);
    const bool nut = NT != nullptr, rot = fabs(om) > 1e-6;
#define L(NUTB, ROTB)                                                                              \
    hipLaunchKernelGGL((k_rhs_only<S, R, NUTB, ROTB>), grid, dim3(256), 0, st, d, m, U, V, W, NT, Ru, Rv, \
                       Rw, sc, ubar, nu, om, full ? 1 : 0)
    if (nut) { if (rot) L(true, true); else L(true, false); }
    else { if (rot) L(false, true); else L(false, false); }
#undef L
}

template <typename S>
void launch_ghosts(const Dims &d, S *F, int field, int wall_mode, const double *alpha,
                   double dyg_bot, double dyg_top, hipStream_t st) {
    launch_ghosts3<S>(d, &F, &field, &wall_mode, &alpha, 1, dyg_bot, dyg_top, st);
}

template <typename S>
void launch_ghosts3(const Dims &d, S *const *F, const int *field, const int *wall_mode,
                    const double *const *alpha, int n, double dyg_bot, double dyg_top, hipStream_t st) {
    GhostSet gs{};
    int64_t tot = 0;
    for (int f = 0; f < n; f++) {
        gs.F[f] = F[f];
        gs.field[f] = field[f];
        gs.mode[f] = wall_mode[f];
        gs.alpha[f] = alpha[f];
        tot += ghost_count(d, field[f]);
        gs.end[f] = tot;
    }
    gs.n = n;
    if (tot == 0) return;
    hipLaunchKernelGGL((k_ghosts<S>), dim3((unsigned)((tot + 255) / 256), launch_batch().n), dim3(256), 0, st, d, gs,
                       dyg_bot, dyg_top, launch_batch().cs);
}

int project_vw_blocks(const Dims &d, int esize) {
    static const int kc_env = getenv("CF_PV_KC") ? atoi(getenv("CF_PV_KC")) : 0;
    const int kc = kc_env > 0 ? kc_env : (esize == 4 ? 8 : 1);
    return tiles_x(d) * tiles_y(d) * ((d.nzp + kc - 1) / kc);
}

template <typename S, typename P>
void launch_project_vw(const Dims &d, const Metrics &m, const P *phi, S *V, S *W, hipStream_t st, double *dtv,
                       double *dtw, int gcopy) {
    // planes per thread: a march hides the latency of the narrow FP32
    // accesses (8 planes measured best, 0.158 against 0.160 ms with 16);
    // FP64 rows move enough bytes per access with one plane
    static const int kc_env = getenv("CF_PV_KC") ? atoi(getenv("CF_PV_KC")) : 0;
    const int kc = kc_env > 0 ? kc_env : (sizeof(S) == 4 ? 8 : 1);
    const int tx = tiles_x(d), ty = tiles_y(d), nt = tx * ty * ((d.nzp + kc - 1) / kc);
    // minima tracked for as many planes as the default march has (FP64: 1,
    // FP32: 8)
    constexpr int KM = sizeof(S) == 4 ? 8 : 1;
    const dim3 grid(nt, launch_batch().n);
    const int64_t cs = launch_batch().cs;
    if (gcopy)
        hipLaunchKernelGGL((k_project_vw<S, P, KM, true>), grid, dim3(256), 0, st, d, m, phi, V, W, tx, ty, nt, kc,
                           dtv, dtw, cs);
    else
        hipLaunchKernelGGL((k_project_vw<S, P, KM, false>), grid, dim3(256), 0, st, d, m, phi, V, W, tx, ty, nt,
                           kc, dtv, dtw, cs);
}

template <typename S>
void launch_massflow(const Dims &d, const Metrics &m, const S *U, double *part, Scalars *sc,
                     int apply, hipStream_t st) {
    const int nb = massflow_parts(d);
    hipLaunchKernelGGL((k_massflow_u_partial<S>), dim3(nb), dim3(256), 0, st, d, m, U, nullptr, part);
    hipLaunchKernelGGL(k_massflow_u_final, dim3(1), dim3(512), 0, st, part, nb, sc, apply, (int64_t)0);
}
static int massflow_u_parts(const Dims &d) {
    const int64_t n = (int64_t)(d.nx - 2) * (d.nzg - 2);
    return (int)((n + 255) / 256);
}
static int massflow_v_parts(const Dims &d) {
    const int64_t n = (int64_t)(d.nxg - 2) * (d.nzg - 2);
    return (int)((n + 255) / 256);
}
int massflow_parts(const Dims &d) { return std::max(massflow_u_parts(d), massflow_v_parts(d)); }

// the correction from the partials k_rhs_rk left (mf_mode 1: bulk flow,
// 2: centreline); writes corr and dPdx
void launch_massflow_fused_final(const Dims &d, const double *mfpart, int nparts, int mf_mode, const double *ubar,
                                 Scalars *sc, hipStream_t st) {
    const dim3 grid(1, launch_batch().n);
    const int64_t cs = launch_batch().cs;
    if (mf_mode == 1)
        hipLaunchKernelGGL(k_massflow_u_final, grid, dim3(512), 0, st, mfpart, nparts, sc, 1, cs);
    else
        hipLaunchKernelGGL(k_flow_final, grid, dim3(512), 0, st, mfpart, nparts, sc, ubar, d.ny / 2 + 1,
                           (double)((d.nx - 2) * (d.nzg - 2)), (int)MF_CENTERLINE, cs);
}

template <typename S>
void launch_centerline_u(const Dims &d, const S *U, const double *ubar, double *part, Scalars *sc, int apply,
                         hipStream_t st) {
    const int nb = massflow_u_parts(d), jc = d.ny / 2 + 1;
    hipLaunchKernelGGL((k_centerline_partial<S>), dim3(nb), dim3(256), 0, st, d, U, jc, part);
    hipLaunchKernelGGL(k_flow_final, dim3(1), dim3(512), 0, st, part, nb, sc, ubar, jc,
                       (double)((d.nx - 2) * (d.nzg - 2)), apply ? (int)MF_CENTERLINE : (int)MF_CENTERLINE_MEAN,
                       (int64_t)0);
}
template <typename S>
void launch_massflow_v(const Dims &d, const double *y, const S *V, double *part, Scalars *sc, int apply,
                       hipStream_t st) {
    const int nb = massflow_v_parts(d);
    hipLaunchKernelGGL((k_massflow_v_partial<S>), dim3(nb), dim3(256), 0, st, d, y, V, part);
    hipLaunchKernelGGL(k_flow_final, dim3(1), dim3(512), 0, st, part, nb, sc, (const double *)nullptr, 0, 0.0,
                       apply ? (int)MF_V_APPLY : (int)MF_V, (int64_t)0);
}
template <typename S>
void launch_add_corr_v(const Dims &d, S *V, const Scalars *sc, hipStream_t st) {
    const int64_t n = (int64_t)(d.nxg - 2) * (d.ny - 2) * (d.nzg - 2);
    hipLaunchKernelGGL((k_add_corr_v<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, V, sc);
}

template <typename S>
void launch_add_corr_u(const Dims &d, S *U, const Scalars *sc, hipStream_t st) {
    const int64_t n = (int64_t)(d.nx - 2) * (d.nyg - 2) * d.nzp;
    hipLaunchKernelGGL((k_add_corr_u<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, U, sc);
}

size_t dt_scratch_doubles(const Dims &d) {
    const int tx = (d.nxp + MF_TX - 1) / MF_TX, ty = (d.nyp + MF_TY - 1) / MF_TY;
    const int nch = (d.nzp + MF_KC - 1) / MF_KC;
    return (size_t)ty * nch * d.nxp + (size_t)tx * nch * d.nyp + (size_t)tx * ty * d.nzp +
           (size_t)(d.nxp + d.nyp + d.nzp);
}

template <typename S>
void launch_dt(const Dims &d, const double *dxc, const double *dyc, const double *dzc,
               const S *U, const S *V, const S *W, const double *ubar, double *scratch,
               Scalars *sc, double cfl, double dt_vis, hipStream_t st) {
    const int tx = (d.nxp + MF_TX - 1) / MF_TX, ty = (d.nyp + MF_TY - 1) / MF_TY;
    const int nch = (d.nzp + MF_KC - 1) / MF_KC;
    double *mins = scratch + (dt_scratch_doubles(d) - (size_t)(d.nxp + d.nyp + d.nzp));
    hipLaunchKernelGGL((k_dt_min<S>), dim3(tx * ty * nch), dim3(MF_TX, MF_TY), 0, st, d, U, V, W, ubar,
                       scratch, tx, ty, nch);
    const int nl = d.nxp + d.nyp + d.nzp;
    hipLaunchKernelGGL(k_dt_lines, dim3(nl), dim3(256), 0, st, d, scratch, mins, tx, ty, nch);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, st, d, dxc, dyc, dzc, mins, sc, cfl, dt_vis);
}

template <typename S>
void launch_div(const Dims &d, const double *x, const double *y, const double *z, const S *U,
                const S *V, const S *W, double *part, hipStream_t st) {
    hipLaunchKernelGGL((k_div_partial<S>), dim3(d.nzg - 2), dim3(256), 0, st, d, x, y, z, U, V, W, part);
}

template <typename S>
void launch_smagorinsky(const Dims &d, const Metrics &m, const double *y, const double *yg, const double *delta,
                        const S *U, const S *V, const S *W, S *NT, const Scalars *sc, const double *ubar,
                        double cs, double nu, int damping, hipStream_t st) {
    const int tx = tiles_x(d), ty = tiles_y(d), nt = tx * ty * d.nzp;
    hipLaunchKernelGGL((k_smagorinsky<S, S>), dim3(nt, launch_batch().n), dim3(256), 0, st, d, m, y, yg, delta, U,
                       V, W, NT, sc, ubar, cs, nu, damping, tx, ty, nt, launch_batch().cs);
}

// each plane is split over STATS_NB workgroups, so the launch fills the GPU;
// each wall sum over STATS_WNB
constexpr int STATS_NB = 8;
constexpr int STATS_WNB = 16;
size_t stats_out_doubles(const Dims &d) { return (size_t)7 * d.ny * (STATS_NB + 1) + 8 + 6 * STATS_WNB; }

template <typename S>
void launch_stats_tail(const Dims &d, const Metrics &m, const double *y, const double *yg, const S *U, const S *V,
                       const S *W, const S *NT, const double *ubar, double *out, double *part, Scalars *sc, double nu,
                       hipStream_t st) {
    double *tail = out + 7 * (size_t)d.ny;
    double *wpart = out + 7 * (size_t)d.ny * (STATS_NB + 1) + 8;
    hipLaunchKernelGGL((k_stats_walls<S>), dim3(6, STATS_WNB), dim3(256), 0, st, d, U, NT, ubar, wpart, tail + 4);
    const int64_t nu_ = (int64_t)d.nx * d.nyg * d.nzg, nv = (int64_t)d.nxg * d.ny * d.nzg,
                  nw = (int64_t)d.nxg * d.nyg * d.nz;
    const int nbs = (int)std::min<int64_t>(2048, (std::max(std::max(nu_, nv), nw) + 255) / 256);
    hipLaunchKernelGGL((k_nan_scan<S>), dim3(nbs, 3), dim3(256), 0, st, U, V, W, nu_, nv, nw, tail + 4);
    launch_massflow<S>(d, m, U, part, sc, 0, st);
    launch_massflow_v<S>(d, y, V, part, sc, 0, st);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(1), 0, st, d, y, yg, wpart, STATS_WNB, out, sc, nu);
}

template <typename S>
void launch_stats(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                  const double *ubar, double *out, hipStream_t st) {
    double *part = out + 7 * (size_t)d.ny + 8;  // the tail of the stats_out_doubles() buffer
    hipLaunchKernelGGL((k_stats_planes<S>), dim3(d.ny, STATS_NB), dim3(256), 0, st, d, m, U, V, W, ubar, part);
    hipLaunchKernelGGL(k_stats_fold, dim3((7 * d.ny + 255) / 256), dim3(256), 0, st, part, out, 7 * d.ny, STATS_NB);
}

// ---------------------------------------------------------------------------
// Actual pressure (pressure.f90:21-341), the pieces around the Poisson chain.
// ---------------------------------------------------------------------------
// mean of F(2:nxg-1, j, 2:nzg-1) for j = j0 (block 0) and j1 (block 1), in
// FP64 with a fixed reduction order (subgrid.f90:605-607)
template <typename S>
__global__ void __launch_bounds__(256) k_plane_means(Dims d, const S *__restrict__ F, int j0, int j1,
                                                    double *__restrict__ out) {
    __shared__ double sh[8];
    const int j = blockIdx.x == 0 ? j0 : j1;
    const int ni = d.nxg - 2, n = ni * (d.nzg - 2);
    double acc = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) {
        const int k = 2 + q / ni, i = 2 + q % ni;
        acc += (double)F[i3(d.nxg, d.nyg, i, j, k)];
    }
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) out[blockIdx.x] = acc / (double)n;
}

// Wall conditions dP/dy = d/dy(nu dV/dy) (pressure.f90:256-341) at every
// (i, k) of 2..nxg-1 x 2..nzg-1: bc goes to the wall rows of the rhs_v array
// (compute_rhs_v_boundary, :240-250), to bcs (2 x (nxg-2) x (nzg-2)) and,
// times the solver coefficients, to the source of the first and last rows of
// the Poisson right-hand side at the unique cells (:137-139, applied before
// the transforms: the DFT is linear).
template <typename S>
__global__ void k_pressure_bc(Dims d, const S *__restrict__ V, PressureBC c, double *__restrict__ Rv,
                              double *__restrict__ bcs, double *__restrict__ ysrc) {
    const int ni = d.nxg - 2, n = ni * (d.nzg - 2);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = 2 + q % ni, k = 2 + q / ni;
    const int nxg = d.nxg, ny = d.ny;
    auto v = [&](int j) { return (double)V[i3(nxg, ny, i, j, k)]; };
    const double dzeta = 1.0;
    double dV = (-3.0 * v(1) + 4.0 * v(2) - v(3)) / (2.0 * dzeta) * 1.0 / c.g_b;
    double ddV = (2.0 * v(1) - 5.0 * v(2) + 4.0 * v(3) - 1.0 * v(4)) / (dzeta * dzeta);
    ddV = (ddV - dV * c.g2_b) / (c.g_b * c.g_b);
    const double b1 = c.dnu_b * dV + c.nui_b * ddV;
    dV = (3.0 * v(ny) - 4.0 * v(ny - 1) + v(ny - 2)) / (2.0 * dzeta) * 1.0 / c.g_t;
    ddV = (2.0 * v(ny) - 5.0 * v(ny - 1) + 4.0 * v(ny - 2) - 1.0 * v(ny - 3)) / (dzeta * dzeta);
    ddV = (ddV - dV * c.g2_t) / (c.g_t * c.g_t);
    const double b2 = c.dnu_t * dV + c.nui_t * ddV;
    Rv[i3(nxg, ny, i, 1, k)] = b1;
    Rv[i3(nxg, ny, i, ny, k)] = b2;
    bcs[q] = b1;
    bcs[n + q] = b2;
    if (i <= d.nxp + 1 && k <= d.nzp + 1) {
        const int u = (k - 2) * d.nxp + (i - 2);
        ysrc[u] = c.src_b * b1;
        ysrc[d.nxp * d.nzp + u] = c.src_t * b2;
    }
}

template <typename S>
void launch_plane_means(const Dims &d, const S *F, int j0, int j1, double *out, hipStream_t st) {
    hipLaunchKernelGGL((k_plane_means<S>), dim3(2), dim3(256), 0, st, d, F, j0, j1, out);
}
template <typename S>
void launch_pressure_bc(const Dims &d, const S *V, const PressureBC &c, double *Rv, double *bcs, double *ysrc,
                        hipStream_t st) {
    const int n = (d.nxg - 2) * (d.nzg - 2);
    hipLaunchKernelGGL((k_pressure_bc<S>), dim3((n + 255) / 256), dim3(256), 0, st, d, V, c, Rv, bcs, ysrc);
}

#define INST(S, R)                                                                                    \
    template void launch_rhs_rk<S, R>(const Dims &, const Metrics &, const S *, const S *, const S *, \
                                      const S *, S *, S *, S *, S *, S *, S *, const Scalars *,       \
                                      const double *, double, double, double, double, int,           \
                                      hipStream_t, double *, int);                                    \
    template void launch_rhs_only<S, R>(const Dims &, const Metrics &, const S *, const S *,          \
                                        const S *, const S *, double *, double *, double *,          \
                                        const Scalars *, const double *, double, double, hipStream_t, \
                                        bool);
INST(double, double)
INST(float, float)
#undef INST
#define INST1(S)                                                                                       \
    template void launch_ghosts<S>(const Dims &, S *, int, int, const double *, double, double,       \
                                   hipStream_t);                                                       \
    template void launch_ghosts3<S>(const Dims &, S *const *, const int *, const int *,               \
                                    const double *const *, int, double, double, hipStream_t);          \
    template void launch_massflow<S>(const Dims &, const Metrics &, const S *, double *, Scalars *,  \
                                     int, hipStream_t);                                                \
    template void launch_add_corr_u<S>(const Dims &, S *, const Scalars *, hipStream_t);              \
    template void launch_centerline_u<S>(const Dims &, const S *, const double *, double *, Scalars *, \
                                         int, hipStream_t);                                            \
    template void launch_massflow_v<S>(const Dims &, const double *, const S *, double *, Scalars *,  \
                                       int, hipStream_t);                                              \
    template void launch_add_corr_v<S>(const Dims &, S *, const Scalars *, hipStream_t);              \
    template void launch_stats_tail<S>(const Dims &, const Metrics &, const double *, const double *,  \
                                       const S *, const S *, const S *, const S *, const double *,     \
                                       double *, double *, Scalars *, double, hipStream_t);            \
    template void launch_dt<S>(const Dims &, const double *, const double *, const double *,          \
                               const S *, const S *, const S *, const double *, double *,             \
                               Scalars *, double, double, hipStream_t);                                \
    template void launch_div<S>(const Dims &, const double *, const double *, const double *,         \
                                const S *, const S *, const S *, double *, hipStream_t);              \
    template void launch_smagorinsky<S>(const Dims &, const Metrics &, const double *, const double *,   \
                                        const double *, const S *, const S *, const S *, S *,         \
                                        const Scalars *, const double *, double, double, int,         \
                                        hipStream_t);                                                  \
    template void launch_stats<S>(const Dims &, const Metrics &, const S *, const S *, const S *,     \
                                  const double *, double *, hipStream_t);                              \
    template void launch_plane_means<S>(const Dims &, const S *, int, int, double *, hipStream_t);     \
    template void launch_pressure_bc<S>(const Dims &, const S *, const PressureBC &, double *, double *, \
                                        double *, hipStream_t);
INST1(double)
INST1(float)
#undef INST1
template void launch_project_vw<double, double>(const Dims &, const Metrics &, const double *, double *,
                                                 double *, hipStream_t, double *, double *, int);
template void launch_project_vw<float, float>(const Dims &, const Metrics &, const float *, float *,
                                               float *, hipStream_t, double *, double *, int);

}  // namespace cf

// ---------------------------------------------------------------------------
// single-thread bookkeeping kernels (keep scalars on the device)
// ---------------------------------------------------------------------------
namespace cf {

__global__ void k_reset_counters(Scalars *sc, int64_t cs) {
    sc = case_ptr(sc, cs);
    sc->n_nan_eqwm = 0;  // wallmodel.f90:613-614
    sc->n_iter_eqwm = 0;
    sc->n_fallback_mlp = 0;
}
__global__ void k_dt_fixed(Scalars *sc, double cfl, int64_t cs) {
    sc = case_ptr(sc, cs);
    sc->dt = -cfl;
}
// t = to + rk_t(3)*dt with rk_t(3) = 1 (time_integration.f90:146)
__global__ void k_advance_time(Scalars *sc, int64_t cs) {
    sc = case_ptr(sc, cs);
    sc->t = sc->t + 1.0 * sc->dt;
}

// mass-flow correction in the deviation storage: U = ubar(j) + U', and the
// correction adds corr to U on the interior (time_integration.f90:155), i.e.
// to ubar(2..nyg-1); the ghost entries follow the wall condition the ghost
// kernel applies to U' (Dirichlet: antisymmetric, Neumann: symmetric).
__global__ void __launch_bounds__(256) k_ubar_shift(Scalars *sc, double *ubar, int nyg, int neumann, int64_t cs) {
    sc = case_ptr(sc, cs);
    ubar = case_ptr(ubar, cs);
    const double c = sc->corr;
    __syncthreads();  // every thread has read corr before thread 0 moves on
    for (int j = 2 + threadIdx.x; j <= nyg - 1; j += blockDim.x) ubar[j] += c;
    __syncthreads();
    if (threadIdx.x == 0) {
        ubar[1] = neumann ? ubar[2] : -ubar[2];
        ubar[nyg] = neumann ? ubar[nyg - 1] : -ubar[nyg - 1];
        sc->qbar_u = sc->qbar_u + c * sc->norm_u;
    }
}
void launch_ubar_shift(Scalars *sc, double *ubar, int nyg, int neumann, hipStream_t st) {
    hipLaunchKernelGGL(k_ubar_shift, dim3(1, launch_batch().n), dim3(256), 0, st, sc, ubar, nyg, neumann,
                       launch_batch().cs);
}

void launch_reset_counters(Scalars *sc, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1, launch_batch().n), dim3(1), 0, st, sc, launch_batch().cs);
}
void launch_dt_final_only(Scalars *sc, double cfl, hipStream_t st) {
    hipLaunchKernelGGL(k_dt_fixed, dim3(1, launch_batch().n), dim3(1), 0, st, sc, cfl, launch_batch().cs);
}
void launch_advance_time(Scalars *sc, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_time, dim3(1, launch_batch().n), dim3(1), 0, st, sc, launch_batch().cs);
}

}  // namespace cf
