// Kernel launcher declarations shared by the HIP translation units.
#pragma once

#include "cf_common.hpp"

namespace cf {

// Factorisation of one FFT length into radices 8, 4, 2, 3 (applied in order).
struct FFTPlan {
    int n;
    int nr;
    int radix[16];
};
// radix 8 in the plan of the x transforms (the z transforms always use it)
constexpr bool FFT_X_RADIX8 = true;

// ---- batches (several cases of one size class stepped by each launch) ----
// The launchers of the step kernels launch launch_batch().n copies of their
// grid along y (one per case) and pass the byte stride cs between cases;
// BatchScope sets them for the launches of one batched step on this thread.
struct Batch {
    int n = 1;
    int64_t cs = 0;
};
Batch &launch_batch();
struct BatchScope {
    Batch prev;
    BatchScope(int n, int64_t cs) : prev(launch_batch()) { launch_batch() = Batch{n, cs}; }
    ~BatchScope() { launch_batch() = prev; }
};

// ---- dynamic Smagorinsky (cf_dsm.hip) ----
struct DsmWork {
    double *Q;     // 10 centre arrays (nxg*nyg*nzg each): centre velocities, |S|Sij, |S|
    double *Fu;    // filtered U, V, W on their own grids, back to back
    double *part;  // dsm_part_elems(d): per-wave plane partials of LM, MM
    double *avg;   // nyg: <LM>/<MM> per row (rows 1 and nyg are ghosts)
};
size_t dsm_part_elems(const Dims &d);
template <typename S>
void launch_dsm(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W, const double *ubar,
                const DsmWork &w, S *NT, int dirichlet, hipStream_t st);

// ---- stencils (cf_stencil.hip) ----
// mfpart (substeps 2 and 3 only; null: none): one partial per workgroup of
// the bulk (mf_mode 1) or centreline (mf_mode 2) sum of the new U, for
// launch_massflow_fused_final; rhs_rk_tiles(d, nut) workgroups
template <typename S, typename R>
void launch_rhs_rk(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                   const S *NT, S *Uo, S *Vo, S *Wo, S *Gu, S *Gv, S *Gw,
                   const Scalars *sc, const double *ubar, double nu, double om, double a_new,
                   double a_g, int sub, hipStream_t st, double *mfpart = nullptr, int mf_mode = 0);
int rhs_rk_tiles(const Dims &d, bool nut);
void launch_massflow_fused_final(const Dims &d, const double *mfpart, int nparts, int mf_mode, const double *ubar,
                                 Scalars *sc, hipStream_t st);
template <typename S, typename R>
void launch_rhs_only(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                     const S *NT, double *Ru, double *Rv, double *Rw, const Scalars *sc,
                     const double *ubar, double nu, double om, hipStream_t st, bool full = false);
// actual pressure (pressure.f90): plane means of a centred field at two y
// rows, and the wall conditions d/dy(nu dV/dy)
template <typename S>
void launch_plane_means(const Dims &d, const S *F, int j0, int j1, double *out, hipStream_t st);
struct PressureBC {
    double g_b, g2_b, g_t, g2_t;      // one-sided metric factors dy/dzeta, d2y/dzeta2 at the walls
    double nui_b, dnu_b, nui_t, dnu_t;  // nu + avg nu_t at the wall face and its y derivative
    double src_b, src_t;              // coef_bc_1 (yg2-yg1) and -coef_bc_2 (yg(nyg)-yg(nyg-1))
};
template <typename S>
void launch_pressure_bc(const Dims &d, const S *V, const PressureBC &c, double *Rv, double *bcs, double *ysrc,
                        hipStream_t st);
template <typename S>
void launch_ghosts(const Dims &d, S *F, int field, int wall_mode, const double *alpha,
                   double dyg_bot, double dyg_top, hipStream_t st);
template <typename S>
void launch_ghosts3(const Dims &d, S *const *F, const int *field, const int *wall_mode,
                    const double *const *alpha, int n, double dyg_bot, double dyg_top, hipStream_t st);
// dtv, dtw (optional): per-workgroup maxima of compute_dt's spacing/|v| and
// spacing/|w| over the projected values, project_vw_blocks(d, sizeof(S))
template <typename S, typename P>
void launch_project_vw(const Dims &d, const Metrics &m, const P *phi, S *V, S *W, hipStream_t st,
                       double *dtv = nullptr, double *dtw = nullptr, int gcopy = 0);
int project_vw_blocks(const Dims &d, int esize);
int ifftx_blocks(const Dims &d, int esize);
// dt_next from the partial maxima of a step's last projection (DtParts)
template <typename S>
void launch_dt_fused_final(const Dims &d, const DtParts &p, int nu_, int nvw, const double *dyc, const S *V,
                           Scalars *sc, double cfl, double dt_vis, hipStream_t st);
void launch_dt_take_next(Scalars *sc, hipStream_t st);
size_t dt_fused_mid_doubles();  // room launch_dt_fused_final needs after the partial arrays
template <typename S>
void launch_massflow(const Dims &d, const Metrics &m, const S *U, double *part, Scalars *sc,
                     int apply, hipStream_t st);
template <typename S>
void launch_add_corr_u(const Dims &d, S *U, const Scalars *sc, hipStream_t st);
// the centreline mean of U (pad[0]); apply (x_mass_cte = 2): also the
// correction (corr) and dPdx
template <typename S>
void launch_centerline_u(const Dims &d, const S *U, const double *ubar, double *part, Scalars *sc, int apply,
                         hipStream_t st);
// bulk V (pad[1]); apply: the y correction corr_v and dPdy
template <typename S>
void launch_massflow_v(const Dims &d, const double *y, const S *V, double *part, Scalars *sc, int apply,
                       hipStream_t st);
template <typename S>
void launch_add_corr_v(const Dims &d, S *V, const Scalars *sc, hipStream_t st);
template <typename S>
void launch_dt(const Dims &d, const double *dxc, const double *dyc, const double *dzc,
               const S *U, const S *V, const S *W, const double *ubar, double *scratch,
               Scalars *sc, double cfl, double dt_vis, hipStream_t st);
size_t dt_scratch_doubles(const Dims &d);
int massflow_parts(const Dims &d);
template <typename S>
void launch_div(const Dims &d, const double *x, const double *y, const double *z, const S *U,
                const S *V, const S *W, double *part, hipStream_t st);
template <typename S>
void launch_smagorinsky(const Dims &d, const Metrics &m, const double *y, const double *yg, const double *delta,
                        const S *U, const S *V, const S *W, S *NT, const Scalars *sc, const double *ubar,
                        double cs, double nu, int damping, hipStream_t st);
template <typename S>
void launch_stats(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                  const double *ubar, double *out, hipStream_t st);

void launch_stats_accum(const double *out, double *acc, int n, hipStream_t st);
size_t stats_out_doubles(const Dims &d);  // the `out` buffer of launch_stats (results + block partials)
// after launch_stats: u_tau, Re_tau, bulk U and V into out[7 ny .. 7 ny + 3] and
// the NaN flag of U, V, W into out[7 ny + 4] (y, yg 1-based; `part` is the
// mass-flow partial buffer)
template <typename S>
void launch_stats_tail(const Dims &d, const Metrics &m, const double *y, const double *yg, const S *U, const S *V,
                       const S *W, const S *NT, const double *ubar, double *out, double *part, Scalars *sc, double nu,
                       hipStream_t st);

// ---- Poisson chain (cf_fft.hip) ----
int fft_rows_per_wg(int M, int elems_per_row, int esize);
// ysrc (optional): 2 x nzp x nxp values added to the divergence of the first
// and the last interior y row (the actual-pressure wall sources)
template <typename S, typename T>
void launch_div_fftx(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                     Cx<T> *spec, const FFTPlan &pM, const Cx<T> *twM, const Cx<T> *twN,
                     hipStream_t st, const double *ysrc = nullptr);
int fft_cols_per_wg(int n, int esize);
bool thomas_supported(const Dims &d);
template <typename T>
void launch_fftz(const Dims &d, Cx<T> *spec, const FFTPlan &pZ, const Cx<T> *twZ, int sign,
                 hipStream_t st);
// whether the y solve of nmodes (kx, kz) systems takes the lane form, which
// needs the FP64 checkpoint scratch of thomas_scratch_elems
bool thomas_uses_scratch(int nmodes);
size_t thomas_scratch_elems(const Dims &d);  // FP64 checkpoints of the lane-form y solve

template <typename T>
void launch_thomas(const Dims &d, Cx<T> *spec, double *cscr, const double *kxx, const double *kzz,
                   const double *dd, const double *dl, const double *du, hipStream_t st);
// forward z transform, y solve and inverse z transform in two launches (one
// workgroup per kx slab) in place of launch_fftz, launch_thomas, launch_fftz
// where zsolve_fused(d); needs the FP64 checkpoint scratch
bool zsolve_fused(const Dims &d);
template <typename T>
void launch_zsolve(const Dims &d, Cx<T> *spec, double *cscr, const FFTPlan &pZ, const Cx<T> *twZ,
                   const double *kxx, const double *kzz, const double *dd, const double *dl, const double *du,
                   hipStream_t st);
// the y solve's FP64 checkpoint scratch (lane form or fused form): whether
// it is needed and its size in doubles
bool ysolve_uses_scratch(const Dims &d);
size_t ysolve_scratch_elems(const Dims &d);
// U null: phi only, no projection of U; corr (optional): the U projection
// also adds corr->corr (the mass-flow correction, time_integration.f90:155)
// dtp (optional, dtp.pu non-null): per-workgroup maxima of compute_dt's
// spacing/|u| over the final U, ifftx_blocks(d, sizeof(T)) of them.
// gcopy (here and in launch_project_vw): each projected value is also
// written to the periodic and Dirichlet-wall ghost cells it is the source
// of, so no ghost pass is needed after the projection (walls without a wall
// model)
template <typename S, typename T, typename P>
void launch_ifftx_proju(const Dims &d, const Metrics &m, const Cx<T> *spec, P *phi, S *U,
                        const FFTPlan &pM, const Cx<T> *twM, const Cx<T> *twN, double norm,
                        hipStream_t st, const Scalars *corr = nullptr, const DtParts *dtp = nullptr,
                        int gcopy = 0);

// ---- wall model (cf_eqwm.hip) ----
struct EqwmParams {
    int ncv;
    int max_iter;
    double stretch;
    double tol;
    double rho;
};
void launch_eqwm_batch(int n, const double *u, const double *nu, const double *h,
                       const double *tau_in, double *tau, int *iter, EqwmParams p,
                       hipStream_t st);

}  // namespace cf
