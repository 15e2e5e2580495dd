// Stencil-class kernels of the RK3 step: fused RHS + Runge-Kutta update,
// ghost-cell boundary conditions, projection of V and W, bulk-flow and CFL
// reductions, divergence check, constant-coefficient Smagorinsky and plane
// statistics.  All are HBM-streaming kernels: one thread per unique cell,
// x fastest, tiles remapped per XCD.
#include "cf_kernels.hpp"

namespace cf {

// ---------------------------------------------------------------------------
// Fused RHS (equations.f90:24-498) + RK3 combine (time_integration.f90:104-145)
//
// One thread per unique cell (i in 2..nx-1, j in 2..nyg-1, k in 2..nzg-2);
// duplicate periodic planes are filled afterwards by the ghost kernel, as the
// reference's apply_boundary_conditions overwrites them.  Low-storage form:
//   substep 1:  U1 = U0 + dt*a11*F1,   G = U0 + dt*a21*F1
//   substep s:  Us = G  + dt*ass*Fs    (s = 2, 3; a31 = a21, a32 = 0)
// so the step keeps G instead of Uo and F1 (3 fields instead of 6).
// ---------------------------------------------------------------------------
template <typename S, typename R, bool NUT, bool ROT, int SUB>
__global__ void __launch_bounds__(256)
k_rhs_rk(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V,
         const S *__restrict__ W, const S *__restrict__ NT, S *__restrict__ Uo,
         S *__restrict__ Vo, S *__restrict__ Wo, S *__restrict__ Gu, S *__restrict__ Gv,
         S *__restrict__ Gw, const Scalars *__restrict__ sc, double nu_d, double om_d,
         double a_new, double a_g, int tiles_x, int tiles_y, int ntiles) {
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tz = t / (tiles_x * tiles_y);
    const int i = 2 + tx * 64 + (threadIdx.x & 63);
    const int j = 2 + ty * 4 + (threadIdx.x >> 6);
    const int k = 2 + tz;
    if (i > d.nx - 1 || j > d.nyg - 1) return;

    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny;
    const R nu = (R)nu_d, half = (R)0.5, quart = (R)0.25, two = (R)2;
#define u(I, J, K) ((R)U[i3(nx, nyg, I, J, K)])
#define v(I, J, K) ((R)V[i3(nxg, ny, I, J, K)])
#define w(I, J, K) ((R)W[i3(nxg, nyg, I, J, K)])
#define nt(I, J, K) ((R)NT[i3(nxg, nyg, I, J, K)])
    const R w0j = (R)m.w0[j], w1j = (R)m.w1[j];
    const R w0m = (R)m.w0[j - 1], w1m = (R)m.w1[j - 1];
    const double dt = sc->dt;

    // ---------------- u ----------------
    {
        R c1 = half * (u(i, j, k) + u(i + 1, j, k));
        R c0 = half * (u(i - 1, j, k) + u(i, j, k));
        R t1 = -(c1 * c1 - c0 * c0) * (R)m.rdxc[i];
        R uv1 = (w0j * u(i, j, k) + w1j * u(i, j + 1, k)) * (half * (v(i, j, k) + v(i + 1, j, k)));
        R uv0 = (w0m * u(i, j - 1, k) + w1m * u(i, j, k)) * (half * (v(i, j - 1, k) + v(i + 1, j - 1, k)));
        R t2 = -(uv1 - uv0) * (R)m.rdyf[j];
        R uw1 = (half * (u(i, j, k) + u(i, j, k + 1))) * (half * (w(i, j, k) + w(i + 1, j, k)));
        R uw0 = (half * (u(i, j, k - 1) + u(i, j, k))) * (half * (w(i, j, k - 1) + w(i + 1, j, k - 1)));
        R t3 = -(uw1 - uw0) * (R)m.rdzf[k];
        R rdz1 = (R)m.rdzc[k - 1], rdz2 = (R)m.rdzc[k], rdz3 = (R)m.rdzf[k];
        R rdy3 = (R)m.rdyf[j];
        R rdx1 = (R)m.rdxf[i], rdx2 = (R)m.rdxf[i + 1], rdx3 = (R)m.rdxc[i];
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + nt(i, j, k);
            nux2 = nu + nt(i + 1, j, k);
            nuy1 = nu + half * (w0m * nt(i, j - 1, k) + w1m * nt(i, j, k) + w0m * nt(i + 1, j - 1, k) + w1m * nt(i + 1, j, k));
            nuy2 = nu + half * (w0j * nt(i, j, k) + w1j * nt(i, j + 1, k) + w0j * nt(i + 1, j, k) + w1j * nt(i + 1, j + 1, k));
            nuz1 = nu + quart * (nt(i, j, k) + nt(i, j, k - 1) + nt(i + 1, j, k) + nt(i + 1, j, k - 1));
            nuz2 = nu + quart * (nt(i, j, k) + nt(i, j, k + 1) + nt(i + 1, j, k) + nt(i + 1, j, k + 1));
        }
        R dudy1 = (u(i, j + 1, k) - u(i, j, k)) * (R)m.rdyc[j];
        R dudy0 = (u(i, j, k) - u(i, j - 1, k)) * (R)m.rdyc[j - 1];
        R visc = two * rdx3 * (nux2 * (u(i + 1, j, k) - u(i, j, k)) * rdx2 - nux1 * (u(i, j, k) - u(i - 1, j, k)) * rdx1) +
                 rdy3 * (nuy2 * (dudy1 + (v(i + 1, j, k) - v(i, j, k)) * rdx2) -
                         nuy1 * (dudy0 + (v(i + 1, j - 1, k) - v(i, j - 1, k)) * rdx1)) +
                 rdz3 * (nuz2 * ((u(i, j, k + 1) - u(i, j, k)) * rdz2 + (w(i + 1, j, k) - w(i, j, k)) * rdx2) -
                         nuz1 * ((u(i, j, k) - u(i, j, k - 1)) * rdz1 + (w(i + 1, j, k - 1) - w(i, j, k - 1)) * rdx1));
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        r = r + (R)sc->dPdx;
        if (ROT) {
            R a = half * (v(i, j - 1, k) + v(i, j, k));
            R b = half * (v(i + 1, j - 1, k) + v(i + 1, j, k));
            r = r + two * (R)om_d * (half * (a + b));
        }
        const int64_t q = i3(nx, nyg, i, j, k);
        if (SUB == 1) {
            R u0 = u(i, j, k);
            Uo[q] = (S)(u0 + (R)(dt * a_new) * r);
            Gu[q] = (S)(u0 + (R)(dt * a_g) * r);
        } else {
            Uo[q] = (S)((R)Gu[q] + (R)(dt * a_new) * r);
        }
    }
    // ---------------- v (j <= ny-1) ----------------
    if (j <= ny - 1) {
        R c1 = half * (v(i, j, k) + v(i, j + 1, k));
        R c0 = half * (v(i, j - 1, k) + v(i, j, k));
        R t1 = -(c1 * c1 - c0 * c0) * (R)m.rdyc[j];
        R uv1 = (w0j * u(i, j, k) + w1j * u(i, j + 1, k)) * (half * (v(i, j, k) + v(i + 1, j, k)));
        R uv0 = (w0j * u(i - 1, j, k) + w1j * u(i - 1, j + 1, k)) * (half * (v(i - 1, j, k) + v(i, j, k)));
        R t2 = -(uv1 - uv0) * (R)m.rdxf[i];
        R vw1 = (half * (v(i, j, k) + v(i, j, k + 1))) * (w0j * w(i, j, k) + w1j * w(i, j + 1, k));
        R vw0 = (half * (v(i, j, k - 1) + v(i, j, k))) * (w0j * w(i, j, k - 1) + w1j * w(i, j + 1, k - 1));
        R t3 = -(vw1 - vw0) * (R)m.rdzf[k];
        R rdz1 = (R)m.rdzc[k - 1], rdz2 = (R)m.rdzc[k], rdz3 = (R)m.rdzf[k];
        R rdy1 = (R)m.rdyf[j], rdy2 = (R)m.rdyf[j + 1], rdy3 = (R)m.rdyc[j];
        R rdx1 = (R)m.rdxc[i - 1], rdx2 = (R)m.rdxc[i], rdx3 = (R)m.rdxf[i];
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + half * (w0j * nt(i - 1, j, k) + w1j * nt(i - 1, j + 1, k) + w0j * nt(i, j, k) + w1j * nt(i, j + 1, k));
            nux2 = nu + half * (w0j * nt(i, j, k) + w1j * nt(i, j + 1, k) + w0j * nt(i + 1, j, k) + w1j * nt(i + 1, j + 1, k));
            nuy1 = nu + nt(i, j, k);
            nuy2 = nu + nt(i, j + 1, k);
            nuz1 = nu + half * (w0j * nt(i, j, k - 1) + w1j * nt(i, j + 1, k - 1) + w0j * nt(i, j, k) + w1j * nt(i, j + 1, k));
            nuz2 = nu + half * (w0j * nt(i, j, k) + w1j * nt(i, j + 1, k) + w0j * nt(i, j, k + 1) + w1j * nt(i, j + 1, k + 1));
        }
        R visc = rdx3 * (nux2 * ((v(i + 1, j, k) - v(i, j, k)) * rdx2 + (u(i, j + 1, k) - u(i, j, k)) * rdy3) -
                         nux1 * ((v(i, j, k) - v(i - 1, j, k)) * rdx1 + (u(i - 1, j + 1, k) - u(i - 1, j, k)) * rdy3)) +
                 two * rdy3 * (nuy2 * (v(i, j + 1, k) - v(i, j, k)) * rdy2 - nuy1 * (v(i, j, k) - v(i, j - 1, k)) * rdy1) +
                 rdz3 * (nuz2 * ((v(i, j, k + 1) - v(i, j, k)) * rdz2 + (w(i, j + 1, k) - w(i, j, k)) * rdy3) -
                         nuz1 * ((v(i, j, k) - v(i, j, k - 1)) * rdz1 + (w(i, j + 1, k - 1) - w(i, j, k - 1)) * rdy3));
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        if (ROT) {
            R a = half * (u(i - 1, j, k) + u(i, j, k));
            R b = half * (u(i - 1, j + 1, k) + u(i, j + 1, k));
            r = r - two * (R)om_d * (w0j * a + w1j * b);
        }
        const int64_t q = i3(nxg, ny, i, j, k);
        if (SUB == 1) {
            R v0 = v(i, j, k);
            Vo[q] = (S)(v0 + (R)(dt * a_new) * r);
            Gv[q] = (S)(v0 + (R)(dt * a_g) * r);
        } else {
            Vo[q] = (S)((R)Gv[q] + (R)(dt * a_new) * r);
        }
    }
    // ---------------- w ----------------
    {
        R c1 = half * (w(i, j, k) + w(i, j, k + 1));
        R c0 = half * (w(i, j, k - 1) + w(i, j, k));
        R t1 = -(c1 * c1 - c0 * c0) * (R)m.rdzc[k];
        R uw1 = (half * (u(i, j, k) + u(i, j, k + 1))) * (half * (w(i, j, k) + w(i + 1, j, k)));
        R uw0 = (half * (u(i - 1, j, k) + u(i - 1, j, k + 1))) * (half * (w(i - 1, j, k) + w(i, j, k)));
        R t2 = -(uw1 - uw0) * (R)m.rdxf[i];
        R vw1 = (half * (v(i, j, k) + v(i, j, k + 1))) * (w0j * w(i, j, k) + w1j * w(i, j + 1, k));
        R vw0 = (half * (v(i, j - 1, k) + v(i, j - 1, k + 1))) * (w0m * w(i, j - 1, k) + w1m * w(i, j, k));
        R t3 = -(vw1 - vw0) * (R)m.rdyf[j];
        R rdz1 = (R)m.rdzf[k], rdz2 = (R)m.rdzf[k + 1], rdz3 = (R)m.rdzc[k];
        R rdy3 = (R)m.rdyf[j];
        R rdx1 = (R)m.rdxc[i - 1], rdx2 = (R)m.rdxc[i], rdx3 = (R)m.rdxf[i];
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + quart * (nt(i, j, k) + nt(i, j, k + 1) + nt(i - 1, j, k) + nt(i - 1, j, k + 1));
            nux2 = nu + quart * (nt(i, j, k) + nt(i, j, k + 1) + nt(i + 1, j, k) + nt(i + 1, j, k + 1));
            nuy1 = nu + half * (w0m * nt(i, j - 1, k) + w1m * nt(i, j, k) + w0m * nt(i, j - 1, k + 1) + w1m * nt(i, j, k + 1));
            nuy2 = nu + half * (w0j * nt(i, j, k) + w1j * nt(i, j + 1, k) + w0j * nt(i, j, k + 1) + w1j * nt(i, j + 1, k + 1));
            nuz1 = nu + nt(i, j, k);
            nuz2 = nu + nt(i, j, k + 1);
        }
        R dwdy1 = (w(i, j + 1, k) - w(i, j, k)) * (R)m.rdyc[j];
        R dwdy0 = (w(i, j, k) - w(i, j - 1, k)) * (R)m.rdyc[j - 1];
        R visc = rdx3 * (nux2 * ((w(i + 1, j, k) - w(i, j, k)) * rdx2 + (u(i, j, k + 1) - u(i, j, k)) * rdz2) -
                         nux1 * ((w(i, j, k) - w(i - 1, j, k)) * rdx1 + (u(i - 1, j, k + 1) - u(i - 1, j, k)) * rdz1)) +
                 rdy3 * (nuy2 * (dwdy1 + (v(i, j, k + 1) - v(i, j, k)) * rdz2) -
                         nuy1 * (dwdy0 + (v(i, j - 1, k + 1) - v(i, j - 1, k)) * rdz1)) +
                 two * rdz3 * (nuz2 * (w(i, j, k + 1) - w(i, j, k)) * rdz2 - nuz1 * (w(i, j, k) - w(i, j, k - 1)) * rdz1);
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        r = r + (R)sc->dPdz;
        const int64_t q = i3(nxg, nyg, i, j, k);
        if (SUB == 1) {
            R w0v = w(i, j, k);
            Wo[q] = (S)(w0v + (R)(dt * a_new) * r);
            Gw[q] = (S)(w0v + (R)(dt * a_g) * r);
        } else {
            Wo[q] = (S)((R)Gw[q] + (R)(dt * a_new) * r);
        }
    }
#undef u
#undef v
#undef w
#undef nt
}

// ---------------------------------------------------------------------------
// RHS only (compute_rhs_u/v/w as separate outputs on the interior ranges of
// the reference, equations.f90:29,191,356), used by the module-level entry
// points.  Same arithmetic as the fused kernel with dt = 0 paths removed.
// ---------------------------------------------------------------------------
template <typename S, typename R, bool NUT, bool ROT>
__global__ void __launch_bounds__(256)
k_rhs_only(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V,
           const S *__restrict__ W, const S *__restrict__ NT, double *__restrict__ Ru,
           double *__restrict__ Rv, double *__restrict__ Rw, const Scalars *__restrict__ sc,
           double nu_d, double om_d) {
    // covers i in 2..nxg-1, j in 2..nyg-1, k in 2..nzg-1 (union of the three
    // interior boxes, duplicates included, as the reference computes them)
    const int i = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 2 + blockIdx.z;
    if (i > d.nxg - 1 || j > d.nyg - 1) return;
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny, nz = d.nz, nzg = d.nzg;
    const R nu = (R)nu_d, half = (R)0.5, quart = (R)0.25, two = (R)2;
#define u(I, J, K) ((R)U[i3(nx, nyg, I, J, K)])
#define v(I, J, K) ((R)V[i3(nxg, ny, I, J, K)])
#define w(I, J, K) ((R)W[i3(nxg, nyg, I, J, K)])
#define nt(I, J, K) ((R)NT[i3(nxg, nyg, I, J, K)])
    const R w0j = (R)m.w0[j], w1j = (R)m.w1[j];
    const R w0m = (R)m.w0[j - 1], w1m = (R)m.w1[j - 1];
    if (i <= nx - 1 && Ru) {
        R c1 = half * (u(i, j, k) + u(i + 1, j, k));
        R c0 = half * (u(i - 1, j, k) + u(i, j, k));
        R t1 = -(c1 * c1 - c0 * c0) * (R)m.rdxc[i];
        R uv1 = (w0j * u(i, j, k) + w1j * u(i, j + 1, k)) * (half * (v(i, j, k) + v(i + 1, j, k)));
        R uv0 = (w0m * u(i, j - 1, k) + w1m * u(i, j, k)) * (half * (v(i, j - 1, k) + v(i + 1, j - 1, k)));
        R t2 = -(uv1 - uv0) * (R)m.rdyf[j];
        R uw1 = (half * (u(i, j, k) + u(i, j, k + 1))) * (half * (w(i, j, k) + w(i + 1, j, k)));
        R uw0 = (half * (u(i, j, k - 1) + u(i, j, k))) * (half * (w(i, j, k - 1) + w(i + 1, j, k - 1)));
        R t3 = -(uw1 - uw0) * (R)m.rdzf[k];
        R rdz1 = (R)m.rdzc[k - 1], rdz2 = (R)m.rdzc[k], rdz3 = (R)m.rdzf[k];
        R rdy3 = (R)m.rdyf[j];
        R rdx1 = (R)m.rdxf[i], rdx2 = (R)m.rdxf[i + 1], rdx3 = (R)m.rdxc[i];
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + nt(i, j, k);
            nux2 = nu + nt(i + 1, j, k);
            nuy1 = nu + half * (w0m * nt(i, j - 1, k) + w1m * nt(i, j, k) + w0m * nt(i + 1, j - 1, k) + w1m * nt(i + 1, j, k));
            nuy2 = nu + half * (w0j * nt(i, j, k) + w1j * nt(i, j + 1, k) + w0j * nt(i + 1, j, k) + w1j * nt(i + 1, j + 1, k));
            nuz1 = nu + quart * (nt(i, j, k) + nt(i, j, k - 1) + nt(i + 1, j, k) + nt(i + 1, j, k - 1));
            nuz2 = nu + quart * (nt(i, j, k) + nt(i, j, k + 1) + nt(i + 1, j, k) + nt(i + 1, j, k + 1));
        }
        R dudy1 = (u(i, j + 1, k) - u(i, j, k)) * (R)m.rdyc[j];
        R dudy0 = (u(i, j, k) - u(i, j - 1, k)) * (R)m.rdyc[j - 1];
        R visc = two * rdx3 * (nux2 * (u(i + 1, j, k) - u(i, j, k)) * rdx2 - nux1 * (u(i, j, k) - u(i - 1, j, k)) * rdx1) +
                 rdy3 * (nuy2 * (dudy1 + (v(i + 1, j, k) - v(i, j, k)) * rdx2) -
                         nuy1 * (dudy0 + (v(i + 1, j - 1, k) - v(i, j - 1, k)) * rdx1)) +
                 rdz3 * (nuz2 * ((u(i, j, k + 1) - u(i, j, k)) * rdz2 + (w(i + 1, j, k) - w(i, j, k)) * rdx2) -
                         nuz1 * ((u(i, j, k) - u(i, j, k - 1)) * rdz1 + (w(i + 1, j, k - 1) - w(i, j, k - 1)) * rdx1));
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        r = r + (R)sc->dPdx;
        if (ROT) {
            R a = half * (v(i, j - 1, k) + v(i, j, k));
            R b = half * (v(i + 1, j - 1, k) + v(i + 1, j, k));
            r = r + two * (R)om_d * (half * (a + b));
        }
        Ru[i3(nx - 2, nyg - 2, i - 1, j - 1, k - 1)] = (double)r;
    }
    if (j <= ny - 1 && Rv) {
        R c1 = half * (v(i, j, k) + v(i, j + 1, k));
        R c0 = half * (v(i, j - 1, k) + v(i, j, k));
        R t1 = -(c1 * c1 - c0 * c0) * (R)m.rdyc[j];
        R uv1 = (w0j * u(i, j, k) + w1j * u(i, j + 1, k)) * (half * (v(i, j, k) + v(i + 1, j, k)));
        R uv0 = (w0j * u(i - 1, j, k) + w1j * u(i - 1, j + 1, k)) * (half * (v(i - 1, j, k) + v(i, j, k)));
        R t2 = -(uv1 - uv0) * (R)m.rdxf[i];
        R vw1 = (half * (v(i, j, k) + v(i, j, k + 1))) * (w0j * w(i, j, k) + w1j * w(i, j + 1, k));
        R vw0 = (half * (v(i, j, k - 1) + v(i, j, k))) * (w0j * w(i, j, k - 1) + w1j * w(i, j + 1, k - 1));
        R t3 = -(vw1 - vw0) * (R)m.rdzf[k];
        R rdz1 = (R)m.rdzc[k - 1], rdz2 = (R)m.rdzc[k], rdz3 = (R)m.rdzf[k];
        R rdy1 = (R)m.rdyf[j], rdy2 = (R)m.rdyf[j + 1], rdy3 = (R)m.rdyc[j];
        R rdx1 = (R)m.rdxc[i - 1], rdx2 = (R)m.rdxc[i], rdx3 = (R)m.rdxf[i];
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + half * (w0j * nt(i - 1, j, k) + w1j * nt(i - 1, j + 1, k) + w0j * nt(i, j, k) + w1j * nt(i, j + 1, k));
            nux2 = nu + half * (w0j * nt(i, j, k) + w1j * nt(i, j + 1, k) + w0j * nt(i + 1, j, k) + w1j * nt(i + 1, j + 1, k));
            nuy1 = nu + nt(i, j, k);
            nuy2 = nu + nt(i, j + 1, k);
            nuz1 = nu + half * (w0j * nt(i, j, k - 1) + w1j * nt(i, j + 1, k - 1) + w0j * nt(i, j, k) + w1j * nt(i, j + 1, k));
            nuz2 = nu + half * (w0j * nt(i, j, k) + w1j * nt(i, j + 1, k) + w0j * nt(i, j, k + 1) + w1j * nt(i, j + 1, k + 1));
        }
        R visc = rdx3 * (nux2 * ((v(i + 1, j, k) - v(i, j, k)) * rdx2 + (u(i, j + 1, k) - u(i, j, k)) * rdy3) -
                         nux1 * ((v(i, j, k) - v(i - 1, j, k)) * rdx1 + (u(i - 1, j + 1, k) - u(i - 1, j, k)) * rdy3)) +
                 two * rdy3 * (nuy2 * (v(i, j + 1, k) - v(i, j, k)) * rdy2 - nuy1 * (v(i, j, k) - v(i, j - 1, k)) * rdy1) +
                 rdz3 * (nuz2 * ((v(i, j, k + 1) - v(i, j, k)) * rdz2 + (w(i, j + 1, k) - w(i, j, k)) * rdy3) -
                         nuz1 * ((v(i, j, k) - v(i, j, k - 1)) * rdz1 + (w(i, j + 1, k - 1) - w(i, j, k - 1)) * rdy3));
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        if (ROT) {
            R a = half * (u(i - 1, j, k) + u(i, j, k));
            R b = half * (u(i - 1, j + 1, k) + u(i, j + 1, k));
            r = r - two * (R)om_d * (w0j * a + w1j * b);
        }
        Rv[i3(nxg - 2, ny - 2, i - 1, j - 1, k - 1)] = (double)r;
    }
    if (k <= nz - 1 && Rw) {
        R c1 = half * (w(i, j, k) + w(i, j, k + 1));
        R c0 = half * (w(i, j, k - 1) + w(i, j, k));
        R t1 = -(c1 * c1 - c0 * c0) * (R)m.rdzc[k];
        R uw1 = (half * (u(i, j, k) + u(i, j, k + 1))) * (half * (w(i, j, k) + w(i + 1, j, k)));
        R uw0 = (half * (u(i - 1, j, k) + u(i - 1, j, k + 1))) * (half * (w(i - 1, j, k) + w(i, j, k)));
        R t2 = -(uw1 - uw0) * (R)m.rdxf[i];
        R vw1 = (half * (v(i, j, k) + v(i, j, k + 1))) * (w0j * w(i, j, k) + w1j * w(i, j + 1, k));
        R vw0 = (half * (v(i, j - 1, k) + v(i, j - 1, k + 1))) * (w0m * w(i, j - 1, k) + w1m * w(i, j, k));
        R t3 = -(vw1 - vw0) * (R)m.rdyf[j];
        R rdz1 = (R)m.rdzf[k], rdz2 = (R)m.rdzf[k + 1], rdz3 = (R)m.rdzc[k];
        R rdy3 = (R)m.rdyf[j];
        R rdx1 = (R)m.rdxc[i - 1], rdx2 = (R)m.rdxc[i], rdx3 = (R)m.rdxf[i];
        R nux1 = nu, nux2 = nu, nuy1 = nu, nuy2 = nu, nuz1 = nu, nuz2 = nu;
        if (NUT) {
            nux1 = nu + quart * (nt(i, j, k) + nt(i, j, k + 1) + nt(i - 1, j, k) + nt(i - 1, j, k + 1));
            nux2 = nu + quart * (nt(i, j, k) + nt(i, j, k + 1) + nt(i + 1, j, k) + nt(i + 1, j, k + 1));
            nuy1 = nu + half * (w0m * nt(i, j - 1, k) + w1m * nt(i, j, k) + w0m * nt(i, j - 1, k + 1) + w1m * nt(i, j, k + 1));
            nuy2 = nu + half * (w0j * nt(i, j, k) + w1j * nt(i, j + 1, k) + w0j * nt(i, j, k + 1) + w1j * nt(i, j + 1, k + 1));
            nuz1 = nu + nt(i, j, k);
            nuz2 = nu + nt(i, j, k + 1);
        }
        R dwdy1 = (w(i, j + 1, k) - w(i, j, k)) * (R)m.rdyc[j];
        R dwdy0 = (w(i, j, k) - w(i, j - 1, k)) * (R)m.rdyc[j - 1];
        R visc = rdx3 * (nux2 * ((w(i + 1, j, k) - w(i, j, k)) * rdx2 + (u(i, j, k + 1) - u(i, j, k)) * rdz2) -
                         nux1 * ((w(i, j, k) - w(i - 1, j, k)) * rdx1 + (u(i - 1, j, k + 1) - u(i - 1, j, k)) * rdz1)) +
                 rdy3 * (nuy2 * (dwdy1 + (v(i, j, k + 1) - v(i, j, k)) * rdz2) -
                         nuy1 * (dwdy0 + (v(i, j - 1, k + 1) - v(i, j - 1, k)) * rdz1)) +
                 two * rdz3 * (nuz2 * (w(i, j, k + 1) - w(i, j, k)) * rdz2 - nuz1 * (w(i, j, k) - w(i, j, k - 1)) * rdz1);
        R r = t1 + t2;
        r = r + t3;
        r = r + visc;
        r = r + (R)sc->dPdz;
        Rw[i3(nxg - 2, nyg - 2, i - 1, j - 1, k - 1)] = (double)r;
    }
    (void)nzg;
#undef u
#undef v
#undef w
#undef nt
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

template <typename S>
__global__ void k_ghosts(Dims d, S *__restrict__ F, int field, int wall_mode,
                         const double *__restrict__ alpha, double dyg_bot, double dyg_top) {
    // wall_mode: 0 Dirichlet (negate / zero), 1 nonzero Neumann with alpha,
    //            2 Neumann (copy)
    int n1, n2, n3;
    bool xf, zf, yf;
    if (field == 0) { n1 = d.nx; n2 = d.nyg; n3 = d.nzg; xf = true; yf = false; zf = false; }
    else if (field == 1) { n1 = d.nxg; n2 = d.ny; n3 = d.nzg; xf = false; yf = true; zf = false; }
    else if (field == 2) { n1 = d.nxg; n2 = d.nyg; n3 = d.nz; xf = false; yf = false; zf = true; }
    else { n1 = d.nxg; n2 = d.nyg; n3 = d.nzg; xf = false; yf = false; zf = false; }
    const int nxgh = xf ? 2 : 3;              // ghost columns in x
    const int nzgh = zf ? 2 : 3;              // ghost planes in z
    const int64_t nY = (int64_t)n1 * 2 * n3;  // wall rows
    const int64_t nZ = (int64_t)n1 * (n2 - 2) * nzgh;
    const int64_t nX = (int64_t)nxgh * (n2 - 2) * (n3 - nzgh);
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (g < nY) {
        i = 1 + (int)(g % n1);
        int w = (int)((g / n1) % 2);
        k = 1 + (int)(g / (2 * (int64_t)n1));
        j = w == 0 ? 1 : n2;
    } else if ((g -= nY) < nZ) {
        i = 1 + (int)(g % n1);
        j = 2 + (int)((g / n1) % (n2 - 2));
        int z = (int)(g / ((int64_t)n1 * (n2 - 2)));
        k = zf ? (z == 0 ? 1 : n3) : (z == 0 ? 1 : n3 - 2 + z);
    } else if ((g -= nZ) < nX) {
        int c = (int)(g % nxgh);
        j = 2 + (int)((g / nxgh) % (n2 - 2));
        k = 2 + (int)(g / ((int64_t)nxgh * (n2 - 2)));
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

// ---------------------------------------------------------------------------
// project_velocity for V and W (projection.f90:213-223); U is projected in
// the inverse-x-FFT kernel that produces the same row of phi.
// phi holds unique cells: phi[((k-2)*nyp + (j-2))*nxp + (i-2)].
// ---------------------------------------------------------------------------
template <typename S, typename P>
__global__ void __launch_bounds__(256)
k_project_vw(Dims d, Metrics m, const P *__restrict__ phi, S *__restrict__ V,
             S *__restrict__ W, int tiles_x, int tiles_y, int ntiles) {
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tz = t / (tiles_x * tiles_y);
    const int i = 2 + tx * 64 + (threadIdx.x & 63);
    const int j = 2 + ty * 4 + (threadIdx.x >> 6);
    const int k = 2 + tz;
    if (i > d.nx - 1 || j > d.nyg - 1) return;
    const int nxp = d.nxp, nyp = d.nyp;
    const int64_t pc = ((int64_t)(k - 2) * nyp + (j - 2)) * nxp + (i - 2);
    const double pc_v = (double)phi[pc];
    if (j <= d.ny - 1) {
        const double pn = (double)phi[pc + nxp];
        const int64_t q = i3(d.nxg, d.ny, i, j, k);
        V[q] = (S)((double)V[q] - (pn - pc_v) * m.rdyc[j]);
    }
    {
        const int kn = (k + 1 == d.nzg - 1) ? 2 : k + 1;  // periodic image of phi(:,:,nz)
        const double pn = (double)phi[((int64_t)(kn - 2) * nyp + (j - 2)) * nxp + (i - 2)];
        const int64_t q = i3(d.nxg, d.nyg, i, j, k);
        W[q] = (S)((double)W[q] - (pn - pc_v) * m.rdzc[k]);
    }
}

// ---------------------------------------------------------------------------
// Block reductions (wavefront 64, LDS across wavefronts), fixed order so that
// results are reproducible run to run.
// ---------------------------------------------------------------------------
template <typename T, typename Op>
__device__ T block_reduce(T v, Op op, T *sh) {
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_down(v, o, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    if (wid == 0) {
        v = lane < nw ? sh[lane] : sh[0];
        if (lane >= nw) v = sh[0];
        for (int o = 32; o > 0; o >>= 1) {
            T other = __shfl_down(v, o, 64);
            if (lane + o < nw) v = op(v, other);
        }
    }
    __syncthreads();
    return v;  // valid in thread 0
}
struct OpAdd { __device__ double operator()(double a, double b) const { return a + b; } };
struct OpMax { __device__ double operator()(double a, double b) const { return fmax(a, b); } };

// mass_flow.f90:116-123: per-plane partial sums of the U trapezoid
template <typename S>
__global__ void __launch_bounds__(256)
k_massflow_u_partial(Dims d, Metrics m, const S *__restrict__ U, double *__restrict__ part) {
    __shared__ double sh[8];
    const int k = 2 + blockIdx.x;  // unique planes 2..nzg-2
    double acc = 0.0;
    const int ni = d.nx - 2, nj = d.nyg - 3;
    for (int q = threadIdx.x; q < ni * nj; q += blockDim.x) {
        const int i = 2 + q % ni, j = 3 + q / ni;
        acc += ((double)U[i3(d.nx, d.nyg, i, j, k)] + (double)U[i3(d.nx, d.nyg, i, j - 1, k)]) * 0.5 * m.dyg[j];
    }
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// final: Int_U over k = 2..nzg-1, i.e. the unique planes plus the duplicate
// plane nzg-1 (= plane 2 after the periodic copy, mass_flow.f90:116)
__global__ void k_massflow_u_final(const double *__restrict__ part, int nplanes,
                                   Scalars *sc, int apply) {
    __shared__ double sh[8];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nplanes; q += blockDim.x) acc += part[q];
    acc = block_reduce(acc, OpAdd(), sh);
    if (threadIdx.x == 0) {
        const double total = acc + part[0];
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

// time_integration.f90:188-199: maxima of spacing/|u| over the reference's
// ranges (ghost duplicates included); max is order independent.
template <typename S>
__global__ void __launch_bounds__(256)
k_dt_partial(Dims d, const double *__restrict__ dxc, const double *__restrict__ dyc,
             const double *__restrict__ dzc, const S *__restrict__ U, const S *__restrict__ V,
             const S *__restrict__ W, double *__restrict__ part) {
    __shared__ double sh[8];
    const int k = 2 + blockIdx.x;  // 2..nzg-1
    double mu = 0.0, mv = 0.0, mw = 0.0;
    const int ni = d.nxg - 2, nj = d.nyg - 2;
    for (int q = threadIdx.x; q < ni * nj; q += blockDim.x) {
        const int i = 2 + q % ni, j = 2 + q / ni;
        mu = fmax(mu, dxc[i] / fabs((double)U[i3(d.nx, d.nyg, i, j, k)]));
        mv = fmax(mv, dyc[j] / fabs((double)V[i3(d.nxg, d.ny, i, j, k)]));
        mw = fmax(mw, dzc[k] / fabs((double)W[i3(d.nxg, d.nyg, i, j, k)]));
    }
    mu = block_reduce(mu, OpMax(), sh);
    mv = block_reduce(mv, OpMax(), sh);
    mw = block_reduce(mw, OpMax(), sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = mu;
        part[3 * blockIdx.x + 1] = mv;
        part[3 * blockIdx.x + 2] = mw;
    }
}

__global__ void k_dt_final(const double *__restrict__ part, int nplanes, Scalars *sc,
                           double cfl, double dt_vis) {
    __shared__ double sh[8];
    double mu = 0.0, mv = 0.0, mw = 0.0;
    for (int q = threadIdx.x; q < nplanes; q += blockDim.x) {
        mu = fmax(mu, part[3 * q]);
        mv = fmax(mv, part[3 * q + 1]);
        mw = fmax(mw, part[3 * q + 2]);
    }
    mu = block_reduce(mu, OpMax(), sh);
    mv = block_reduce(mv, OpMax(), sh);
    mw = block_reduce(mw, OpMax(), sh);
    if (threadIdx.x == 0) {
        const double dc = fmin(fmin(cfl * mu, cfl * mv), cfl * mw);
        double dt = fmin(dc, dt_vis);
        dt = fmin(dt, 5e-3);
        if (cfl < 0) dt = -cfl;
        sc->dt = dt;
        sc->dt_conv[0] = mu; sc->dt_conv[1] = mv; sc->dt_conv[2] = mw;
    }
}

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
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256)
k_smagorinsky(Dims d, const double *__restrict__ x, const double *__restrict__ y,
              const double *__restrict__ z, const double *__restrict__ xg,
              const double *__restrict__ yg, const double *__restrict__ zg,
              const double *__restrict__ delta, const S *__restrict__ U, const S *__restrict__ V,
              const S *__restrict__ W, S *__restrict__ NT, const Scalars *__restrict__ sc,
              double cs, double nu, int damping, int tiles_x, int tiles_y, int ntiles) {
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int tx = t % tiles_x, ty = (t / tiles_x) % tiles_y, tz = t / (tiles_x * tiles_y);
    const int i = 2 + tx * 64 + (threadIdx.x & 63);
    const int j = 2 + ty * 4 + (threadIdx.x >> 6);
    const int k = 2 + tz;
    if (i > d.nx - 1 || j > d.nyg - 1) return;
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny;
#define u(I, J, K) ((double)U[i3(nx, nyg, I, J, K)])
#define v(I, J, K) ((double)V[i3(nxg, ny, I, J, K)])
#define w(I, J, K) ((double)W[i3(nxg, nyg, I, J, K)])
    const double s11 = (u(i, j, k) - u(i - 1, j, k)) / (x[i] - x[i - 1]);
    const double s22 = (v(i, j, k) - v(i, j - 1, k)) / (y[j] - y[j - 1]);
    const double s33 = (w(i, j, k) - w(i, j, k - 1)) / (z[k] - z[k - 1]);
    auto e4 = [&](int I, int J, int K) {
        return 0.5 * ((u(I - 1, J, K) - u(I - 1, J - 1, K)) / (yg[J] - yg[J - 1]) +
                      (v(I, J - 1, K) - v(I - 1, J - 1, K)) / (xg[I] - xg[I - 1]));
    };
    auto e5 = [&](int I, int J, int K) {
        return 0.5 * ((u(I - 1, J, K) - u(I - 1, J, K - 1)) / (zg[K] - zg[K - 1]) +
                      (w(I, J, K - 1) - w(I - 1, J, K - 1)) / (xg[I] - xg[I - 1]));
    };
    auto e6 = [&](int I, int J, int K) {
        return 0.5 * ((v(I, J - 1, K) - v(I, J - 1, K - 1)) / (zg[K] - zg[K - 1]) +
                      (w(I, J, K - 1) - w(I, J - 1, K - 1)) / (yg[J] - yg[J - 1]));
    };
    const double s12 = 0.25 * (e4(i, j, k) + e4(i + 1, j, k) + e4(i, j + 1, k) + e4(i + 1, j + 1, k));
    const double s13 = 0.25 * (e5(i, j, k) + e5(i + 1, j, k) + e5(i, j, k + 1) + e5(i + 1, j, k + 1));
    const double s23 = 0.25 * (e6(i, j, k) + e6(i, j + 1, k) + e6(i, j, k + 1) + e6(i, j + 1, k + 1));
    double s = 0.0;
    s = s + 2.0 * s11 * s11;
    s = s + 2.0 * s22 * s22;
    s = s + 2.0 * s33 * s33;
    s = s + 4.0 * s12 * s12;
    s = s + 4.0 * s13 * s13;
    s = s + 4.0 * s23 * s23;
    double f = 1.0;
    if (damping) {
        const double utau_t = sqrt(sc->dPdx);
        if (j < nyg / 2) f = 1.0 - exp(-(yg[j] - y[1]) / 25.0 / (nu / utau_t));
        else f = 1.0 - exp(-(y[ny] - yg[j]) / 25.0 / (nu / utau_t));
    }
    const double c = cs * delta[j] * f;
    const double val = fmax((c * c) * sqrt(s), 0.0);
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
               const S *__restrict__ W, double *__restrict__ out) {
    __shared__ double sh[8];
    const int jj = 1 + blockIdx.x;
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny, nz = d.nz, nzg = d.nzg;
    const double a = m.w0[jj], b = m.w1[jj];
    double su = 0, su2 = 0, sw = 0, sw2 = 0, sv = 0, sv2 = 0, suv = 0;
    for (int q = threadIdx.x; q < (nx - 2) * (nzg - 2); q += blockDim.x) {
        const int i = 1 + q % (nx - 2), k = 1 + q / (nx - 2);
        const double t = a * (double)U[i3(nx, nyg, i, jj, k)] + b * (double)U[i3(nx, nyg, i, jj + 1, k)];
        su += t; su2 += t * t;
    }
    for (int q = threadIdx.x; q < (nxg - 2) * (nz - 2); q += blockDim.x) {
        const int i = 1 + q % (nxg - 2), k = 1 + q / (nxg - 2);
        const double t = a * (double)W[i3(nxg, nyg, i, jj, k)] + b * (double)W[i3(nxg, nyg, i, jj + 1, k)];
        sw += t; sw2 += t * t;
    }
    for (int q = threadIdx.x; q < (nxg - 2) * (nzg - 2); q += blockDim.x) {
        const int i = 2 + q % (nxg - 2), k = 2 + q / (nxg - 2);
        const double vv = (double)V[i3(nxg, ny, i, jj, k)];
        sv += vv; sv2 += vv * vv;
        const int ii = i - 1;
        const double ta = a * (double)U[i3(nx, nyg, ii, jj, k)] + b * (double)U[i3(nx, nyg, ii, jj + 1, k)];
        const double tb = a * (double)U[i3(nx, nyg, ii + 1, jj, k)] + b * (double)U[i3(nx, nyg, ii + 1, jj + 1, k)];
        suv += vv * (0.5 * (ta + tb));
    }
    double vals[7] = {su, sv, sw, su2, sv2, sw2, suv};
    for (int c = 0; c < 7; c++) {
        double r = block_reduce(vals[c], OpAdd(), sh);
        if (threadIdx.x == 0) out[c * ny + (jj - 1)] = r;
    }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static inline int tiles_x(const Dims &d) { return (d.nxp + 63) / 64; }
static inline int tiles_y(const Dims &d) { return (d.nyp + 3) / 4; }

template <typename S, typename R>
void launch_rhs_rk(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                   const S *NT, S *Uo, S *Vo, S *Wo, S *Gu, S *Gv, S *Gw,
                   const Scalars *sc, double nu, double om, double a_new, double a_g,
                   int sub, hipStream_t st) {
    const int tx = tiles_x(d), ty = tiles_y(d), nt = tx * ty * d.nzp;
    const bool nut = NT != nullptr, rot = fabs(om) > 1e-6;
#define L(NUTB, ROTB, SUBV)                                                                       \
    hipLaunchKernelGGL((k_rhs_rk<S, R, NUTB, ROTB, SUBV>), dim3(nt), dim3(256), 0, st, d, m, U, V, W, \
                       NT, Uo, Vo, Wo, Gu, Gv, Gw, sc, nu, om, a_new, a_g, tx, ty, nt)
    if (sub == 1) {
        if (nut) { if (rot) L(true, true, 1); else L(true, false, 1); }
        else { if (rot) L(false, true, 1); else L(false, false, 1); }
    } else {
        if (nut) { if (rot) L(true, true, 2); else L(true, false, 2); }
        else { if (rot) L(false, true, 2); else L(false, false, 2); }
    }
#undef L
}

template <typename S, typename R>
void launch_rhs_only(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                     const S *NT, double *Ru, double *Rv, double *Rw, const Scalars *sc,
                     double nu, double om, hipStream_t st) {
    dim3 grid((d.nxg - 2 + 63) / 64, (d.nyg - 2 + 3) / 4, d.nzg - 2);
    const bool nut = NT != nullptr, rot = fabs(om) > 1e-6;
#define L(NUTB, ROTB)                                                                              \
    hipLaunchKernelGGL((k_rhs_only<S, R, NUTB, ROTB>), grid, dim3(256), 0, st, d, m, U, V, W, NT, Ru, Rv, \
                       Rw, sc, nu, om)
    if (nut) { if (rot) L(true, true); else L(true, false); }
    else { if (rot) L(false, true); else L(false, false); }
#undef L
}

template <typename S>
void launch_ghosts(const Dims &d, S *F, int field, int wall_mode, const double *alpha,
                   double dyg_bot, double dyg_top, hipStream_t st) {
    int n1, n2, n3, nxgh, nzgh;
    if (field == 0) { n1 = d.nx; n2 = d.nyg; n3 = d.nzg; nxgh = 2; nzgh = 3; }
    else if (field == 1) { n1 = d.nxg; n2 = d.ny; n3 = d.nzg; nxgh = 3; nzgh = 3; }
    else if (field == 2) { n1 = d.nxg; n2 = d.nyg; n3 = d.nz; nxgh = 3; nzgh = 2; }
    else { n1 = d.nxg; n2 = d.nyg; n3 = d.nzg; nxgh = 3; nzgh = 3; }
    const int64_t n = (int64_t)n1 * 2 * n3 + (int64_t)n1 * (n2 - 2) * nzgh +
                      (int64_t)nxgh * (n2 - 2) * (n3 - nzgh);
    const int blocks = (int)((n + 255) / 256);
    hipLaunchKernelGGL((k_ghosts<S>), dim3(blocks), dim3(256), 0, st, d, F, field, wall_mode, alpha,
                       dyg_bot, dyg_top);
}

template <typename S, typename P>
void launch_project_vw(const Dims &d, const Metrics &m, const P *phi, S *V, S *W, hipStream_t st) {
    const int tx = tiles_x(d), ty = tiles_y(d), nt = tx * ty * d.nzp;
    hipLaunchKernelGGL((k_project_vw<S, P>), dim3(nt), dim3(256), 0, st, d, m, phi, V, W, tx, ty, nt);
}

template <typename S>
void launch_massflow(const Dims &d, const Metrics &m, const S *U, double *part, Scalars *sc,
                     int apply, hipStream_t st) {
    hipLaunchKernelGGL((k_massflow_u_partial<S>), dim3(d.nzp), dim3(256), 0, st, d, m, U, part);
    hipLaunchKernelGGL(k_massflow_u_final, dim3(1), dim3(256), 0, st, part, d.nzp, sc, apply);
}

template <typename S>
void launch_add_corr_u(const Dims &d, S *U, const Scalars *sc, hipStream_t st) {
    const int64_t n = (int64_t)(d.nx - 2) * (d.nyg - 2) * d.nzp;
    hipLaunchKernelGGL((k_add_corr_u<S>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, U, sc);
}

template <typename S>
void launch_dt(const Dims &d, const double *dxc, const double *dyc, const double *dzc,
               const S *U, const S *V, const S *W, double *part, Scalars *sc, double cfl,
               double dt_vis, hipStream_t st) {
    const int np = d.nzg - 2;
    hipLaunchKernelGGL((k_dt_partial<S>), dim3(np), dim3(256), 0, st, d, dxc, dyc, dzc, U, V, W, part);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, st, part, np, sc, cfl, dt_vis);
}

template <typename S>
void launch_div(const Dims &d, const double *x, const double *y, const double *z, const S *U,
                const S *V, const S *W, double *part, hipStream_t st) {
    hipLaunchKernelGGL((k_div_partial<S>), dim3(d.nzg - 2), dim3(256), 0, st, d, x, y, z, U, V, W, part);
}

template <typename S>
void launch_smagorinsky(const Dims &d, const double *x, const double *y, const double *z,
                        const double *xg, const double *yg, const double *zg, const double *delta,
                        const S *U, const S *V, const S *W, S *NT, const Scalars *sc, double cs,
                        double nu, int damping, hipStream_t st) {
    const int tx = tiles_x(d), ty = tiles_y(d), nt = tx * ty * d.nzp;
    hipLaunchKernelGGL((k_smagorinsky<S>), dim3(nt), dim3(256), 0, st, d, x, y, z, xg, yg, zg, delta,
                       U, V, W, NT, sc, cs, nu, damping, tx, ty, nt);
}

template <typename S>
void launch_stats(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W, double *out,
                  hipStream_t st) {
    hipLaunchKernelGGL((k_stats_planes<S>), dim3(d.ny), dim3(256), 0, st, d, m, U, V, W, out);
}

#define INST(S, R)                                                                                    \
    template void launch_rhs_rk<S, R>(const Dims &, const Metrics &, const S *, const S *, const S *, \
                                      const S *, S *, S *, S *, S *, S *, S *, const Scalars *,       \
                                      double, double, double, double, int, hipStream_t);              \
    template void launch_rhs_only<S, R>(const Dims &, const Metrics &, const S *, const S *,          \
                                        const S *, const S *, double *, double *, double *,          \
                                        const Scalars *, double, double, hipStream_t);
INST(double, double)
INST(float, float)
#undef INST
#define INST1(S)                                                                                       \
    template void launch_ghosts<S>(const Dims &, S *, int, int, const double *, double, double,       \
                                   hipStream_t);                                                       \
    template void launch_massflow<S>(const Dims &, const Metrics &, const S *, double *, Scalars *,  \
                                     int, hipStream_t);                                                \
    template void launch_add_corr_u<S>(const Dims &, S *, const Scalars *, hipStream_t);              \
    template void launch_dt<S>(const Dims &, const double *, const double *, const double *,          \
                               const S *, const S *, const S *, double *, Scalars *, double, double,  \
                               hipStream_t);                                                           \
    template void launch_div<S>(const Dims &, const double *, const double *, const double *,         \
                                const S *, const S *, const S *, double *, hipStream_t);              \
    template void launch_smagorinsky<S>(const Dims &, const double *, const double *, const double *, \
                                        const double *, const double *, const double *,               \
                                        const double *, const S *, const S *, const S *, S *,         \
                                        const Scalars *, double, double, int, hipStream_t);           \
    template void launch_stats<S>(const Dims &, const Metrics &, const S *, const S *, const S *,     \
                                  double *, hipStream_t);
INST1(double)
INST1(float)
#undef INST1
template void launch_project_vw<double, double>(const Dims &, const Metrics &, const double *, double *,
                                                 double *, hipStream_t);
template void launch_project_vw<float, float>(const Dims &, const Metrics &, const float *, float *,
                                               float *, hipStream_t);

}  // namespace cf

// ---------------------------------------------------------------------------
// single-thread bookkeeping kernels (keep scalars on the device)
// ---------------------------------------------------------------------------
namespace cf {

__global__ void k_reset_counters(Scalars *sc) {
    sc->n_nan_eqwm = 0;  // wallmodel.f90:613-614
    sc->n_iter_eqwm = 0;
}
__global__ void k_dt_fixed(Scalars *sc, double cfl) { sc->dt = -cfl; }
// t = to + rk_t(3)*dt with rk_t(3) = 1 (time_integration.f90:146)
__global__ void k_advance_time(Scalars *sc) { sc->t = sc->t + 1.0 * sc->dt; }

void launch_reset_counters(Scalars *sc, hipStream_t st) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, st, sc);
}
void launch_dt_final_only(Scalars *sc, double cfl, hipStream_t st) {
    hipLaunchKernelGGL(k_dt_fixed, dim3(1), dim3(1), 0, st, sc, cfl);
}
void launch_advance_time(Scalars *sc, hipStream_t st) {
    hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, st, sc);
}

}  // namespace cf
