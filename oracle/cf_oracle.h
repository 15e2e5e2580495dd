/*
 * CPU oracle: an FP64 restatement of the reference channel-flow time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * C-ABI or the host package) links, loads or calls this code.  It is used by
 * tests/, by __graft_entry__.smoke() as the checker, and by bench.py's CPU
 * baseline leg.
 *
 * Arrays follow the reference layout exactly: Fortran order (x fastest, then
 * y, then z), 1-based extents with ghosts:
 *   U (nx ,nyg,nzg)   V (nxg,ny ,nzg)   W (nxg,nyg,nz )   P (nxg,nyg,nzg)
 *   nu_t (nxg,nyg,nzg)
 *   rhs_u (2:nx-1,2:nyg-1,2:nzg-1)  rhs_v (2:nxg-1,2:ny-1,2:nzg-1)
 *   rhs_w (2:nxg-1,2:nyg-1,2:nz-1)
 *   alpha_x (nx,2,nzg)  alpha_z (nxg,2,nz)
 * with nxg = nx+1, nyg = ny+1, nzg = nz+1 (one rank holds the whole domain).
 */
#ifndef CF_ORACLE_H
#define CF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    double cfl;            /* <0: fixed dt = -cfl (time_integration.f90:223) */
    double nu;
    double dPdx, dPdz;     /* forcing carried between steps                  */
    double omega_z;        /* spanwise rotation (equations.f90:169)          */
    double cs;             /* Smagorinsky coefficient (subgrid.f90:580)      */
    double eqwm_stretch;   /* wallmodel.f90:621 passes 1.10                  */
    double eqwm_tol;       /* eqwm.f90:109 hard-codes 1e-7                   */
    double tau_mean_eqwm;  /* fallback stress (initialization.f90:524)       */
    int x_mass_cte, y_mass_cte;
    int les_model;         /* 0 none, 1 constant-coefficient, 2 dynamic Smagorinsky */
    int iwall_model;       /* 0 no-slip, 3 equilibrium wall model            */
    int dirichlet_nu_t;
    int jref_eqwm;
    int eqwm_ncv;          /* wallmodel.f90:621 passes 100                   */
    int eqwm_max_iter;     /* eqwm.f90:110 hard-codes 100                    */
    int nthreads;          /* OpenMP threads for the restatement (1 = serial)*/
    double u_centerline;   /* x_mass_cte = 2 target Ucenterline_x
                              (initialization.f90:453-455)                  */
} orc_params;

typedef struct orc orc_t;

orc_t *orc_create(int nx, int ny, int nz, const double *x, const double *y,
                  const double *z, const orc_params *p);
void orc_destroy(orc_t *o);

/* grid-derived operators (initialization.f90:164-215,281-296,338-388) */
void orc_get_grid(const orc_t *o, double *xg, double *yg, double *zg,
                  double *kxx, double *kzz, double *w0, double *w1,
                  double *dmin3);
void orc_get_dyy(const orc_t *o, double *diag, double *lower, double *upper);

void orc_set_state(orc_t *o, const double *U, const double *V, const double *W);
void orc_get_state(const orc_t *o, double *U, double *V, double *W, double *P);
void orc_set_nu_t(orc_t *o, const double *nu_t);
void orc_get_nu_t(const orc_t *o, double *nu_t, double *avg_nu_t);
void orc_get_alpha(const orc_t *o, double *alpha_x, double *alpha_z,
                   double *tau_init);
void orc_set_tau_init(orc_t *o, const double *tau_init);

/* scalars: [t, dt, dPdx, dPdz, Qflow_x_0, dPdy, Qflow_y_0] */
void orc_set_scalars(orc_t *o, const double *s);
void orc_get_scalars(const orc_t *o, double *s);

/* one routine each, as in the reference modules */
void orc_apply_boundary_conditions(orc_t *o);
void orc_projection_step(orc_t *o, int rk_step);
void orc_compute_rhs(orc_t *o, const double *U, const double *V,
                     const double *W, double *rhs_u, double *rhs_v,
                     double *rhs_w);
double orc_compute_dt(orc_t *o);
void orc_rk3_step(orc_t *o);
double orc_mean_mass_flow_U(const orc_t *o, const double *U);
double orc_mean_mass_flow_V(const orc_t *o, const double *V);
double orc_mean_centerline_U(const orc_t *o, const double *U);
void orc_eddy_viscosity(orc_t *o);
void orc_wall_model(orc_t *o, int *n_nan, int *n_iter);
void orc_check_divergence(const orc_t *o, double *max_signed, double *max_abs);
/* pressure.f90:21-341 on the current state and avg_nu_t: P (nxg,nyg,nzg);
   optional rhs_p (2:nxg-1,2:nyg-1,2:nzg-1) and wall conditions
   bc_1, bc_2 (2:nxg-1,2:nzg-1) */
void orc_compute_pressure(orc_t *o, double *P, double *rhs_p, double *bc_1, double *bc_2);
/* out: Umean,Vmean,Wmean,U2mean,V2mean,W2mean,UVmean (ny each), then
   [utau, Retau, Qflow_x, Qflow_y] */
void orc_statistics(const orc_t *o, double *out);

/* stand-alone pieces */
void orc_eqwm(double umag, double mu_lam, double rho, double h, int ncv,
              double stretch, double tol, int max_iter, double tau_in,
              double *tau, int *iter);
void orc_fft(int n, double *re, double *im, int sign);
void orc_fft2d(int n1, int n2, double *re, double *im, int sign);
int orc_zgtsv(int n, double *dl, double *d, double *du, double *bre,
              double *bim);
void orc_interp_2nd(int ny, const double *y, int n, const double *yi,
                    int ge, int *jj, double *w);

#ifdef __cplusplus
}
#endif
#endif
