// Equilibrium wall model (eqwm.f90:20-223, wallmodel.f90:591-685) on the GPU.
// One lane per wall point: each lane runs the Picard iteration of the 1-D
// equilibrium ODE on the block's shared stretched grid (FP64, bound by the
// exponentials and divisions), so lanes with different iteration counts
// simply retire at different times.
#include "cf_kernels.hpp"
#include "cf_eqwm.hpp"

namespace cf {

// Normalised wall-model grid of eqwm.f90:20-50: g_fa = y_fa/ymax, g_cv =
// y_cv/ymax (the reference scales as y/ymax*h, so y = g*h bit for bit).  The
// maximum is taken over the ncv assigned centroids (the reference's Maxval
// also reads one unassigned slot, y_cv(ncv+1)).  Also the reciprocal
// spacings r_cv[q] = 1/(g_cv[q+1]-g_cv[q]) and r_w = 1/(g_cv[0]-g_fa[0]).
// Filled by one thread into LDS.
__device__ void eqwm_grid(int ncv, double stretch, double *g_fa, double *g_cv, double *r_cv) {
    if (threadIdx.x == 0) {
        const int nfa = ncv + 1;
        g_fa[0] = 0.0;
        double inc = 1.0;
        for (int ifa = 2; ifa <= nfa; ifa++) {
            g_fa[ifa - 1] = g_fa[ifa - 2] + inc;
            g_cv[ifa - 2] = 0.5 * (g_fa[ifa - 1] + g_fa[ifa - 2]);
            inc = inc * stretch;
        }
        double ymax = g_cv[0];
        for (int q = 1; q < ncv; q++) ymax = g_cv[q] > ymax ? g_cv[q] : ymax;
        for (int q = 0; q < nfa; q++) g_fa[q] = g_fa[q] / ymax;
        for (int q = 0; q < ncv; q++) g_cv[q] = g_cv[q] / ymax;
        for (int q = 0; q + 1 < ncv; q++) r_cv[q] = 1.0 / (g_cv[q + 1] - g_cv[q]);
        r_cv[ncv - 1] = 1.0 / (g_cv[0] - g_fa[0]);
    }
    __syncthreads();
}

// solve_u_eqwm (eqwm.f90:90-223): Picard iteration on the mixing-length
// equilibrium ODE, with no pressure-gradient term (the right-hand side is 0
// except the top row, u = u_les).  The forward sweep of the tridiagonal
// solve therefore keeps d' = 0 on every row below the top, and the back
// substitution reduces to u(1) = u_les * prod(-c'(q)): each lane carries the
// running product through the sweep and needs no per-row storage.  Spacing
// divisions use the reciprocals of the normalised grid scaled by 1/h.
// (The oracle keeps the reference's stored-array form; the two agree to the
// Picard tolerance.)
__device__ double eqwm_solve(double umag_les, double mu_lam, double rho_eq, double h_wm, int ncv,
                             double tol, int max_iter, double tau_in, const double *g_fa,
                             const double *r_cv, int *iter_out) {
    const double kappa = 0.41, Aplus = 17.0;
    const double tau_wall_lam = mu_lam * umag_les / h_wm;
    const double nu_eq = mu_lam / rho_eq;
    const double rh = 1.0 / h_wm;
    const double rw = r_cv[ncv - 1] * rh;  // 1/(y_cv(1) - y_fa(1))
    int iter = 0;
    bool done = false;
    double tau_wall = tau_in;
    while (!done) {
        const double tau_prev = tau_wall;
        const double utau_eq = sqrt(tau_wall / rho_eq);
        const double ex = -utau_eq / nu_eq / Aplus;
        const double ck = rho_eq * kappa * utau_eq;
        // total viscosity at face ifa >= 2 (eqwm.f90:136-147)
        auto mu_face = [&](int ifa) -> double {
            const double yf = g_fa[ifa - 1] * h_wm;
            double D = 1.0 - exp(yf * ex);
            D = D * D;
            return mu_lam + ck * yf * D;
        };
        // row 1, the wall CV (eqwm.f90:167-171)
        double mu_lo = mu_face(2);
        const double sup = mu_lo * (r_cv[0] * rh);
        double cpp = sup / (-(sup + mu_lam * rw));
        double prod = -cpp;
        // interior rows (eqwm.f90:156-164)
        for (int icv = 2; icv <= ncv - 1; icv++) {
            const double mu_hi = mu_face(icv + 1);
            const double s_up = mu_hi * (r_cv[icv - 1] * rh);
            const double s_dn = mu_lo * (r_cv[icv - 2] * rh);
            const double m = -(s_up + s_dn) - cpp * s_dn;
            cpp = s_up / m;
            prod *= -cpp;
            mu_lo = mu_hi;
        }
        // top row u(ncv) = u_les, back substitution u(q) = -c'(q) u(q+1)
        const double u1 = umag_les * prod;
        tau_wall = mu_lam * u1 * rw;
        if (fabs((tau_wall - tau_prev) / tau_wall_lam) < tol) done = true;
        if (!done) iter = iter + 1;
        if (iter > max_iter) done = true;
    }
    *iter_out = iter;
    return tau_wall;
}

size_t eqwm_lds_bytes(int ncv, int wg) {
    (void)wg;
    return sizeof(double) * ((size_t)(ncv + 1) + 2 * (size_t)ncv);
}
int eqwm_block(int ncv) {
    (void)ncv;
    return EQWM_WG;
}

struct EqwmLds {
    double *g_fa, *g_cv, *r_cv;
};
__device__ EqwmLds eqwm_lds(int ncv) {
    extern __shared__ double eq_smem[];
    EqwmLds s;
    s.g_fa = eq_smem;
    s.g_cv = s.g_fa + (ncv + 1);
    s.r_cv = s.g_cv + ncv;
    return s;
}

__global__ void __launch_bounds__(EQWM_WG)
k_eqwm_batch(int n, const double *__restrict__ u, const double *__restrict__ nu,
             const double *__restrict__ h, const double *__restrict__ tau_in,
             double *__restrict__ tau, int *__restrict__ iter, EqwmParams p) {
    EqwmLds s = eqwm_lds(p.ncv);
    eqwm_grid(p.ncv, p.stretch, s.g_fa, s.g_cv, s.r_cv);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int it;
    tau[q] = eqwm_solve(u[q], nu[q], p.rho, h[q], p.ncv, p.tol, p.max_iter, tau_in[q], s.g_fa, s.r_cv, &it);
    iter[q] = it;
}

void launch_eqwm_batch(int n, const double *u, const double *nu, const double *h,
                       const double *tau_in, double *tau, int *iter, EqwmParams p,
                       hipStream_t st) {
    if (n <= 0) return;
    const int wg = eqwm_block(p.ncv);
    hipLaunchKernelGGL(k_eqwm_batch, dim3((n + wg - 1) / wg), dim3(wg), eqwm_lds_bytes(p.ncv, wg), st, n,
                       u, nu, h, tau_in, tau, iter, p);
}

// ---------------------------------------------------------------------------
// compute_equilibrium_wall_model (wallmodel.f90:591-664) per wall point
// (i = 1..nx, k = 1..nzg, both walls).  W is brought to the U location with
// the reference's 3-point Lagrange stencils (interpolation.f90:114-252),
// whose indices and weights depend only on the 1-D grids and are tabulated
// on the host.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(EQWM_WG) k_wall_model(Dims d, const S *__restrict__ U, const S *__restrict__ W,
                             const S *__restrict__ NT, const WallTables t,
                             double *__restrict__ tau_init, double *__restrict__ alpha_x,
                             double *__restrict__ azt, Scalars *sc, EqwmParams p, double nu,
                             double h, int jref, double tau_mean, const double *__restrict__ tau_pred,
                             const int *__restrict__ use_pred, const double *__restrict__ ubar, int64_t cs) {
    U = case_ptr(U, cs); W = case_ptr(W, cs); NT = case_ptr(NT, cs);
    tau_init = case_ptr(tau_init, cs); alpha_x = case_ptr(alpha_x, cs); azt = case_ptr(azt, cs);
    sc = case_ptr(sc, cs); ubar = case_ptr(ubar, cs);
    if (cs) {
        nu = sc->nu;
        tau_mean = sc->tau_mean;
    }
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    EqwmLds s = eqwm_lds(p.ncv);
    eqwm_grid(p.ncv, p.stretch, s.g_fa, s.g_cv, s.r_cv);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nx * nzg * 2) return;
    const int i = 1 + q % nx;
    const int k = 1 + (q / nx) % nzg;
    const int wall = q / (nx * nzg);  // 0 bottom, 1 top
    const int jw = wall == 0 ? jref : nyg - jref + 1;
    // temp_1 = interpolate_z_2nd(z, W(:,jw,:), zg) at the three x centres of
    // the x stencil, then temp_2 = interpolate_x_2nd(xg, temp_1, x)
    const int jz = t.jjz[k - 1];
    const double *wz = t.wz + 3 * (k - 1);
    const int jx = t.jjx[i - 1];
    const double *wx = t.wx + 3 * (i - 1);
    double t1[3];
    for (int a = 0; a < 3; a++) {
        const int ii = jx - 1 + a;
        t1[a] = wz[0] * (double)W[i3(nxg, nyg, ii, jw, jz - 1)] + wz[1] * (double)W[i3(nxg, nyg, ii, jw, jz)] +
                wz[2] * (double)W[i3(nxg, nyg, ii, jw, jz + 1)];
    }
    const double wt = wx[0] * t1[0] + wx[1] * t1[1] + wx[2] * t1[2];
    const double uu = (double)U[i3(nx, nyg, i, jw, k)] + (ubar ? ubar[jw] : 0.0);
    const double ul = sqrt(uu * uu + wt * wt);
    const int64_t qa = i3(nx, 2, i, wall + 1, k);
    double tau;
    int it = 0;
    bool solved = true;
    if (use_pred && use_pred[qa]) {
        tau = tau_pred[qa];
        solved = false;
    } else {
        tau = eqwm_solve(ul, nu, 1.0, h, p.ncv, p.tol, p.max_iter, tau_init[qa], s.g_fa, s.r_cv, &it);
    }
    if (solved) {
        if (isnan(tau)) { atomicAdd(&sc->n_nan_eqwm, 1); tau = tau_mean; }
        if (it == p.max_iter) { tau = tau_mean; atomicAdd(&sc->n_iter_eqwm, 1); }
    }
    tau_init[qa] = tau;
    const double tx = tau * uu / ul;
    const double tz = tau * wt / ul;
    double nik = 0.0;
    if (NT) {
        if (wall == 0)
            nik = 0.5 * (0.5 * ((double)NT[i3(nxg, nyg, i, 1, k)] + (double)NT[i3(nxg, nyg, i + 1, 1, k)]) +
                         0.5 * ((double)NT[i3(nxg, nyg, i, 2, k)] + (double)NT[i3(nxg, nyg, i + 1, 2, k)]));
        else
            nik = 0.5 * (0.5 * ((double)NT[i3(nxg, nyg, i, nyg, k)] + (double)NT[i3(nxg, nyg, i + 1, nyg, k)]) +
                         0.5 * ((double)NT[i3(nxg, nyg, i, nyg - 1, k)] + (double)NT[i3(nxg, nyg, i + 1, nyg - 1, k)]));
    }
    alpha_x[qa] = tx / (nu + nik);
    azt[qa] = tz / (nu + nik);
}

// wallmodel.f90:667-674: alpha_z from the 4-point average of alpha_z_temp,
// with the reference's periodic copies folded into the index map.
__global__ void k_alpha_z(Dims d, const double *__restrict__ azt, double *__restrict__ az, int64_t cs) {
    azt = case_ptr(azt, cs);
    az = case_ptr(az, cs);
    const int nxg = d.nxg, nz = d.nz, nx = d.nx;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nxg * 2 * nz) return;
    const int i = 1 + q % nxg;
    const int wl = 1 + (q / nxg) % 2;
    const int k = 1 + q / (2 * nxg);
    const int ii = i == nxg ? 3 : (i == nxg - 1 ? 2 : i);
    const int kk = k == nz ? 2 : k;
    az[q] = 0.25 * (azt[i3(nx, 2, ii, wl, kk)] + azt[i3(nx, 2, ii + 1, wl, kk)] +
                    azt[i3(nx, 2, ii, wl, kk + 1)] + azt[i3(nx, 2, ii + 1, wl, kk + 1)]);
}

template <typename S>
void launch_wall_model(const Dims &d, const S *U, const S *W, const S *NT, const WallTables &t,
                       double *tau_init, double *alpha_x, double *azt, double *az, Scalars *sc,
                       EqwmParams p, double nu, double h, int jref, double tau_mean,
                       const double *tau_pred, const int *use_pred, const double *ubar, hipStream_t st) {
    const int n = d.nx * d.nzg * 2;
    const int wg = eqwm_block(p.ncv);
    const int nb = launch_batch().n;
    const int64_t cs = launch_batch().cs;
    hipLaunchKernelGGL((k_wall_model<S>), dim3((n + wg - 1) / wg, nb), dim3(wg), eqwm_lds_bytes(p.ncv, wg), st, d,
                       U, W, NT, t, tau_init,
                       alpha_x, azt, sc, p, nu, h, jref, tau_mean, tau_pred, use_pred, ubar, cs);
    const int n2 = d.nxg * 2 * d.nz;
    hipLaunchKernelGGL(k_alpha_z, dim3((n2 + 255) / 256, nb), dim3(256), 0, st, d, azt, az, cs);
}

// input of the wall model at every wall point, for the learned surrogate:
// the wall-parallel speed at the matching height, stored at the index
// (i, wall, k) of the (nx, 2, nzg) wall arrays that k_wall_model reads
template <typename S>
__global__ void k_wall_inputs(Dims d, const S *__restrict__ U, const S *__restrict__ W,
                              const WallTables t, int jref, const double *__restrict__ ubar,
                              double *__restrict__ out) {
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nx * nzg * 2) return;
    const int i = 1 + q % nx;
    const int k = 1 + (q / nx) % nzg;
    const int wall = q / (nx * nzg);
    const int jw = wall == 0 ? jref : nyg - jref + 1;
    const int jz = t.jjz[k - 1];
    const double *wz = t.wz + 3 * (k - 1);
    const int jx = t.jjx[i - 1];
    const double *wx = t.wx + 3 * (i - 1);
    double t1[3];
    for (int a = 0; a < 3; a++) {
        const int ii = jx - 1 + a;
        t1[a] = wz[0] * (double)W[i3(nxg, nyg, ii, jw, jz - 1)] + wz[1] * (double)W[i3(nxg, nyg, ii, jw, jz)] +
                wz[2] * (double)W[i3(nxg, nyg, ii, jw, jz + 1)];
    }
    const double wt = wx[0] * t1[0] + wx[1] * t1[1] + wx[2] * t1[2];
    const double uu = (double)U[i3(nx, nyg, i, jw, k)] + (ubar ? ubar[jw] : 0.0);
    out[i3(nx, 2, i, wall + 1, k)] = sqrt(uu * uu + wt * wt);
}

template <typename S>
void launch_wall_inputs(const Dims &d, const S *U, const S *W, const WallTables &t, int jref,
                        const double *ubar, double *out, hipStream_t st) {
    const int n = d.nx * d.nzg * 2;
    hipLaunchKernelGGL((k_wall_inputs<S>), dim3((n + 255) / 256), dim3(256), 0, st, d, U, W, t, jref, ubar,
                       out);
}

template void launch_wall_model<double>(const Dims &, const double *, const double *, const double *,
                                        const WallTables &, double *, double *, double *, double *,
                                        Scalars *, EqwmParams, double, double, int, double,
                                        const double *, const int *, const double *, hipStream_t);
template void launch_wall_model<float>(const Dims &, const float *, const float *, const float *,
                                       const WallTables &, double *, double *, double *, double *,
                                       Scalars *, EqwmParams, double, double, int, double,
                                       const double *, const int *, const double *, hipStream_t);
template void launch_wall_inputs<double>(const Dims &, const double *, const double *, const WallTables &,
                                         int, const double *, double *, hipStream_t);
template void launch_wall_inputs<float>(const Dims &, const float *, const float *, const WallTables &, int,
                                        const double *, double *, hipStream_t);

}  // namespace cf
