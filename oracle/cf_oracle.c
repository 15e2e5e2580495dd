/*
 * CPU oracle -- FP64 restatement of the reference channel-flow time step.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() as the
 * checker, and as bench.py's CPU baseline.  The product never links it.
 *
 * Every routine follows the reference operation by operation for the points
 * it computes (same association order of sums and products, divisions kept
 * as divisions), so that its results track the reference at round-off level.
 * Single-rank MPI semantics are the intended ones: periodic ghosts are plain
 * copies of interior planes (the one-rank self-exchanges of
 * boundary_conditions.f90:156-202 and projection.f90:150-160 would truncate
 * or drop them).
 *
 * Compile with -ffp-contract=off so that no multiply-add is fused.
 *
 * ORC_PLANT (0 in every normal build) selects one deliberate mistake at a
 * PLANT(n, wrong, right) site.  tests/test_oracle_planted.py builds the
 * oracle with -DORC_PLANT=n for each n and checks that the pin guarding that
 * site fails, which shows the pins can tell a plausible mistake apart.
 */
#include "cf_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#define PLANT(n, wrong, right) (ORC_PLANT == (n) ? (wrong) : (right))

struct orc {
    int nx, ny, nz, nxg, nyg, nzg, nxp, nzp, nr;
    orc_params p;
    double *x, *y, *z, *xg, *yg, *zg, *w0, *w1, *kxx, *kzz;
    double *dyy_d, *dyy_l, *dyy_u; /* Dyy(j,j), Dyy(j+1,j), Dyy(j,j+1), j=2.. */
    double dxmin, dymin, dzmin, dx, dz;
    double rk_coef[3][3], rk_t[3];
    double t, dt, dPdx, dPdz, dPdy, Qflow_x_0, Qflow_y_0;
    size_t nU, nV, nW, nC, nRu, nRv, nRw;
    double *U, *V, *W, *P, *Uo, *Vo, *Wo, *nu_t, *avg_nu_t;
    double *Fu[3], *Fv[3], *Fw[3];
    double *rhs_p; /* (2:nxg-1, 2:nyg-1, 2:nzg) */
    double *alpha_x, *alpha_z, *alpha_z_temp, *tau_init;
    int *ix_jj, *iz_jj;   /* 2nd-order interpolation stencils (wall model) */
    double *ix_w, *iz_w;
};

/* ---- index helpers: 1-based Fortran order ------------------------------ */
#define I3(n1, n2, i, j, k) \
    ((size_t)((i)-1) + (size_t)(n1) * ((size_t)((j)-1) + (size_t)(n2) * (size_t)((k)-1)))
#define IU(i, j, k) I3(o->nx, o->nyg, i, j, k)
#define IV(i, j, k) I3(o->nxg, o->ny, i, j, k)
#define IW(i, j, k) I3(o->nxg, o->nyg, i, j, k)
#define IC(i, j, k) I3(o->nxg, o->nyg, i, j, k)
/* rhs arrays start at index 2 in every direction */
#define IRU(i, j, k) I3(o->nx - 2, o->nyg - 2, (i)-1, (j)-1, (k)-1)
#define IRV(i, j, k) I3(o->nxg - 2, o->ny - 2, (i)-1, (j)-1, (k)-1)
#define IRW(i, j, k) I3(o->nxg - 2, o->nyg - 2, (i)-1, (j)-1, (k)-1)
/* rhs_p(2:nxg-1, 2:nyg-1, 2:nzg) */
#define IP(i, j, k) I3(o->nxg - 2, o->nyg - 2, (i)-1, (j)-1, (k)-1)
#define AX(i, w, k) I3(o->nx, 2, i, w, k)
#define AZ(i, w, k) I3(o->nxg, 2, i, w, k)

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static double sq(double a) { return a * a; }

/* initialization.f90:385-388: weights on y faces */
static void build_operators(orc_t *o) {
    int nx = o->nx, ny = o->ny, nz = o->nz;
    int nxm = nx - 1, nym = ny - 1, nzm = nz - 1;
    double *xm = dalloc(nxm), *ym = dalloc(nym), *zm = dalloc(nzm);
    /* initialization.f90:169-199 */
    for (int i = 1; i <= nxm; i++) xm[i - 1] = 0.5 * (o->x[i - 1] + o->x[i]);
    for (int j = 1; j <= nym; j++) ym[j - 1] = 0.5 * (o->y[j - 1] + o->y[j]);
    for (int k = 1; k <= nzm; k++) zm[k - 1] = 0.5 * (o->z[k - 1] + o->z[k]);
    for (int i = 2; i <= nxm + 1; i++) o->xg[i - 1] = xm[i - 2];
    o->xg[0] = xm[0] - 2.0 * (xm[0] - o->x[0]);
    o->xg[nxm + 1] = xm[nxm - 1] + 2.0 * (o->x[nx - 1] - xm[nxm - 1]);
    for (int j = 2; j <= nym + 1; j++) o->yg[j - 1] = ym[j - 2];
    o->yg[0] = ym[0] - 2.0 * (ym[0] - o->y[0]);
    o->yg[nym + 1] = ym[nym - 1] + 2.0 * (o->y[ny - 1] - ym[nym - 1]);
    for (int k = 2; k <= nzm + 1; k++) o->zg[k - 1] = zm[k - 2];
    o->zg[0] = zm[0] - 2.0 * (zm[0] - o->z[0]);
    o->zg[nzm + 1] = zm[nzm - 1] + 2.0 * (o->z[nz - 1] - zm[nzm - 1]);
    free(xm); free(ym); free(zm);

    /* initialization.f90:208-210 */
    double m;
    m = 1e300; for (int i = 2; i <= o->nxg; i++) { double d = o->xg[i - 1] - o->xg[i - 2]; if (d < m) m = d; } o->dxmin = m;
    m = 1e300; for (int j = 2; j <= o->nyg; j++) { double d = o->yg[j - 1] - o->yg[j - 2]; if (d < m) m = d; } o->dymin = m;
    m = 1e300; for (int k = 2; k <= o->nzg; k++) { double d = o->zg[k - 1] - o->zg[k - 2]; if (d < m) m = d; } o->dzmin = m;
    o->dx = o->dxmin;
    o->dz = o->dzmin;

    /* initialization.f90:281-296: modified wavenumbers */
    const double pi = 4.0 * atan(1.0);
    int nxp = o->nxp, nzp = o->nzp;
    int cx = (nxp + 1) / 2, cz = (nzp + 1) / 2; /* Ceiling(n/2) */
    for (int i = 0; i < nxp; i++) {
        double a = (i <= cx) ? (double)i : (double)(-nxp + i);
        o->kxx[i] = 2.0 * (cos(2.0 * pi * a / (double)nxp) - 1.0) / (o->dx * o->dx);
    }
    for (int k = 0; k < nzp; k++) {
        double a = (k <= cz) ? (double)k : (double)(-nzp + k);
        o->kzz[k] = 2.0 * (cos(2.0 * pi * a / (double)nzp) - 1.0) / (o->dz * o->dz);
    }

    /* initialization.f90:343-373: Dyy rows with Neumann closure */
    const double *y = o->y, *yg = o->yg;
    int nyg = o->nyg;
#define Y(j) y[(j)-1]
#define YG(j) yg[(j)-1]
    for (int j = 2; j <= nyg - 1; j++) {
        double a = 1.0 / (Y(j) - Y(j - 1)) / (YG(j + 1) - YG(j));
        double b = 1.0 / (Y(j) - Y(j - 1)) * (-1.0 / (YG(j + 1) - YG(j)) - 1.0 / (YG(j) - YG(j - 1)));
        double c = 1.0 / (Y(j) - Y(j - 1)) / (YG(j) - YG(j - 1));
        o->dyy_d[j - 2] = b;
        if (j < nyg - 1) o->dyy_u[j - 2] = a; /* Dyy(j,j+1) */
        if (j > 2) o->dyy_l[j - 3] = c;       /* Dyy(j,j-1) */
        if (j == 2) o->dyy_d[0] = b + c;
        if (j == nyg - 1) o->dyy_d[j - 2] = a + b;
    }
    /* initialization.f90:386-388 */
    for (int j = 1; j <= ny; j++) {
        o->w0[j - 1] = (YG(j + 1) - Y(j)) / (YG(j + 1) - YG(j));
        o->w1[j - 1] = 1.0 - o->w0[j - 1];
    }
#undef Y
#undef YG
    /* initialization.f90:407-417 */
    memset(o->rk_coef, 0, sizeof(o->rk_coef));
    o->rk_t[0] = 8.0 / 15.0; o->rk_t[1] = 2.0 / 3.0; o->rk_t[2] = 1.0;
    o->rk_coef[0][0] = 8.0 / 15.0;
    o->rk_coef[1][0] = 1.0 / 4.0; o->rk_coef[1][1] = 5.0 / 12.0;
    o->rk_coef[2][0] = 1.0 / 4.0; o->rk_coef[2][1] = 0.0; o->rk_coef[2][2] = 3.0 / 4.0;
}

/* interpolation.f90:114-252: 3-point Lagrange stencil search.
   ge=1: x version (search y(j2) >= yref), ge=0: z version (y(j2) > yref).
   Returns jj (1-based centre of the stencil jj-1,jj,jj+1) and 3 weights
   per target point. */
void orc_interp_2nd(int ny, const double *y, int n, const double *yi, int ge,
                    int *jjo, double *w) {
#define Yv(j) y[(j)-1]
    int jj = ny - 1;
    for (int t = 1; t <= n; t++) {
        double yref = yi[t - 1];
        if (t == 1) jj = 2;
        if (t == n) jj = ny - 1;
        for (int j2 = 3; j2 <= ny; j2++) {
            int hit = ge ? (Yv(j2) >= yref) : (Yv(j2) > yref);
            if (hit) { jj = j2 - 1; break; }
        }
        if (t == 1 && ge && jj == 0) jj = ny;
        double w0 = (yref - Yv(jj)) * (yref - Yv(jj + 1)) / ((Yv(jj - 1) - Yv(jj)) * (Yv(jj - 1) - Yv(jj + 1)));
        double w1 = (yref - Yv(jj - 1)) * (yref - Yv(jj + 1)) / ((Yv(jj) - Yv(jj - 1)) * (Yv(jj) - Yv(jj + 1)));
        double w2 = (yref - Yv(jj - 1)) * (yref - Yv(jj)) / ((Yv(jj + 1) - Yv(jj - 1)) * (Yv(jj + 1) - Yv(jj)));
        jjo[t - 1] = jj;
        w[3 * (t - 1) + 0] = w0;
        w[3 * (t - 1) + 1] = w1;
        w[3 * (t - 1) + 2] = w2;
    }
#undef Yv
}

orc_t *orc_create(int nx, int ny, int nz, const double *x, const double *y,
                  const double *z, const orc_params *p) {
    orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
    o->nx = nx; o->ny = ny; o->nz = nz;
    o->nxg = nx + 1; o->nyg = ny + 1; o->nzg = nz + 1;
    o->nxp = nx - 2; o->nzp = nz - 2; o->nr = o->nyg - 2;
    o->p = *p;
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
    o->x = dalloc(nx); o->y = dalloc(ny); o->z = dalloc(nz);
    memcpy(o->x, x, nx * sizeof(double));
    memcpy(o->y, y, ny * sizeof(double));
    memcpy(o->z, z, nz * sizeof(double));
    o->xg = dalloc(o->nxg); o->yg = dalloc(o->nyg); o->zg = dalloc(o->nzg);
    o->w0 = dalloc(ny); o->w1 = dalloc(ny);
    o->kxx = dalloc(o->nxp); o->kzz = dalloc(o->nzp);
    o->dyy_d = dalloc(o->nr); o->dyy_l = dalloc(o->nr); o->dyy_u = dalloc(o->nr);
    build_operators(o);

    o->nU = (size_t)o->nx * o->nyg * o->nzg;
    o->nV = (size_t)o->nxg * o->ny * o->nzg;
    o->nW = (size_t)o->nxg * o->nyg * o->nz;
    o->nC = (size_t)o->nxg * o->nyg * o->nzg;
    o->nRu = (size_t)(o->nx - 2) * (o->nyg - 2) * (o->nzg - 2);
    o->nRv = (size_t)(o->nxg - 2) * (o->ny - 2) * (o->nzg - 2);
    o->nRw = (size_t)(o->nxg - 2) * (o->nyg - 2) * (o->nz - 2);
    o->U = dalloc(o->nU); o->V = dalloc(o->nV); o->W = dalloc(o->nW);
    o->Uo = dalloc(o->nU); o->Vo = dalloc(o->nV); o->Wo = dalloc(o->nW);
    o->P = dalloc(o->nC); o->nu_t = dalloc(o->nC); o->avg_nu_t = dalloc(o->nyg);
    for (int s = 0; s < 3; s++) {
        o->Fu[s] = dalloc(o->nRu); o->Fv[s] = dalloc(o->nRv); o->Fw[s] = dalloc(o->nRw);
    }
    o->rhs_p = dalloc((size_t)(o->nxg - 2) * (o->nyg - 2) * (o->nzg - 1));
    o->alpha_x = dalloc((size_t)o->nx * 2 * o->nzg);
    o->alpha_z = dalloc((size_t)o->nxg * 2 * o->nz);
    o->alpha_z_temp = dalloc((size_t)o->nx * 2 * o->nzg);
    o->tau_init = dalloc((size_t)o->nx * 2 * o->nzg);
    for (size_t q = 0; q < (size_t)o->nx * 2 * o->nzg; q++) o->tau_init[q] = p->tau_mean_eqwm;
    /* wallmodel.f90:605-610 stencils: z faces -> z centres, x centres -> x faces */
    o->iz_jj = (int *)calloc(o->nzg, sizeof(int)); o->iz_w = dalloc(3 * o->nzg);
    o->ix_jj = (int *)calloc(o->nx, sizeof(int)); o->ix_w = dalloc(3 * o->nx);
    orc_interp_2nd(o->nz, o->z, o->nzg, o->zg, 0, o->iz_jj, o->iz_w);
    orc_interp_2nd(o->nxg, o->xg, o->nx, o->x, 1, o->ix_jj, o->ix_w);
    o->dPdx = p->dPdx; o->dPdz = p->dPdz;
    return o;
}

void orc_destroy(orc_t *o) {
    if (!o) return;
    double *bufs[] = {o->x, o->y, o->z, o->xg, o->yg, o->zg, o->w0, o->w1, o->kxx, o->kzz,
                      o->dyy_d, o->dyy_l, o->dyy_u, o->U, o->V, o->W, o->Uo, o->Vo, o->Wo,
                      o->P, o->nu_t, o->avg_nu_t, o->rhs_p, o->alpha_x, o->alpha_z,
                      o->alpha_z_temp, o->tau_init, o->iz_w, o->ix_w};
    for (size_t q = 0; q < sizeof(bufs) / sizeof(bufs[0]); q++) free(bufs[q]);
    for (int s = 0; s < 3; s++) { free(o->Fu[s]); free(o->Fv[s]); free(o->Fw[s]); }
    free(o->iz_jj); free(o->ix_jj);
    free(o);
}

void orc_get_grid(const orc_t *o, double *xg, double *yg, double *zg, double *kxx,
                  double *kzz, double *w0, double *w1, double *dmin3) {
    if (xg) memcpy(xg, o->xg, o->nxg * sizeof(double));
    if (yg) memcpy(yg, o->yg, o->nyg * sizeof(double));
    if (zg) memcpy(zg, o->zg, o->nzg * sizeof(double));
    if (kxx) memcpy(kxx, o->kxx, o->nxp * sizeof(double));
    if (kzz) memcpy(kzz, o->kzz, o->nzp * sizeof(double));
    if (w0) memcpy(w0, o->w0, o->ny * sizeof(double));
    if (w1) memcpy(w1, o->w1, o->ny * sizeof(double));
    if (dmin3) { dmin3[0] = o->dxmin; dmin3[1] = o->dymin; dmin3[2] = o->dzmin; }
}

void orc_get_dyy(const orc_t *o, double *diag, double *lower, double *upper) {
    memcpy(diag, o->dyy_d, o->nr * sizeof(double));
    memcpy(lower, o->dyy_l, (o->nr - 1) * sizeof(double));
    memcpy(upper, o->dyy_u, (o->nr - 1) * sizeof(double));
}

void orc_set_state(orc_t *o, const double *U, const double *V, const double *W) {
    memcpy(o->U, U, o->nU * sizeof(double));
    memcpy(o->V, V, o->nV * sizeof(double));
    memcpy(o->W, W, o->nW * sizeof(double));
}
void orc_get_state(const orc_t *o, double *U, double *V, double *W, double *P) {
    if (U) memcpy(U, o->U, o->nU * sizeof(double));
    if (V) memcpy(V, o->V, o->nV * sizeof(double));
    if (W) memcpy(W, o->W, o->nW * sizeof(double));
    if (P) memcpy(P, o->P, o->nC * sizeof(double));
}
void orc_set_nu_t(orc_t *o, const double *nu_t) { memcpy(o->nu_t, nu_t, o->nC * sizeof(double)); }
void orc_get_nu_t(const orc_t *o, double *nu_t, double *avg) {
    if (nu_t) memcpy(nu_t, o->nu_t, o->nC * sizeof(double));
    if (avg) memcpy(avg, o->avg_nu_t, o->nyg * sizeof(double));
}
void orc_get_alpha(const orc_t *o, double *ax, double *az, double *ti) {
    if (ax) memcpy(ax, o->alpha_x, (size_t)o->nx * 2 * o->nzg * sizeof(double));
    if (az) memcpy(az, o->alpha_z, (size_t)o->nxg * 2 * o->nz * sizeof(double));
    if (ti) memcpy(ti, o->tau_init, (size_t)o->nx * 2 * o->nzg * sizeof(double));
}
void orc_set_tau_init(orc_t *o, const double *ti) {
    memcpy(o->tau_init, ti, (size_t)o->nx * 2 * o->nzg * sizeof(double));
}
void orc_set_scalars(orc_t *o, const double *s) {
    o->t = s[0]; o->dt = s[1]; o->dPdx = s[2]; o->dPdz = s[3]; o->Qflow_x_0 = s[4];
    o->dPdy = s[5]; o->Qflow_y_0 = s[6];
}
void orc_get_scalars(const orc_t *o, double *s) {
    s[0] = o->t; s[1] = o->dt; s[2] = o->dPdx; s[3] = o->dPdz; s[4] = o->Qflow_x_0;
    s[5] = o->dPdy; s[6] = o->Qflow_y_0;
}

/* ======================================================================
 * boundary_conditions.f90:20-75 with one-rank (intended) semantics
 * ====================================================================== */
static void periodic_x(orc_t *o, double *F, int n1, int n2, int n3, int faces) {
    /* boundary_conditions.f90:87-106 */
    for (int k = 1; k <= n3; k++)
        for (int j = 1; j <= n2; j++) {
            double *r = F + I3(n1, n2, 1, j, k);
            if (faces) { r[0] = r[o->nx - 2]; r[o->nx - 1] = r[1]; }
            else { r[0] = r[o->nxg - 3]; r[o->nxg - 2] = r[1]; r[o->nxg - 1] = r[2]; }
        }
}
static void periodic_z(double *F, int n1, int n2, int n3, int faces) {
    /* boundary_conditions.f90:111-244 on one rank: the self exchange is a
       plain copy of the saved planes */
    size_t pl = (size_t)n1 * n2;
    if (faces) { /* id 3: F(:,:,1)=F(:,:,nz-1), F(:,:,nz)=F(:,:,2) */
        memcpy(F, F + pl * (n3 - 2), pl * sizeof(double));
        memcpy(F + pl * (n3 - 1), F + pl * 1, pl * sizeof(double));
    } else { /* id 1,2,4: F(1)=F(nzg-2), F(nzg-1)=F(2), F(nzg)=F(3) */
        memcpy(F, F + pl * (n3 - 3), pl * sizeof(double));
        memcpy(F + pl * (n3 - 2), F + pl * 1, pl * sizeof(double));
        memcpy(F + pl * (n3 - 1), F + pl * 2, pl * sizeof(double));
    }
}
static void dirichlet_y(double *F, int n1, int n2, int n3, int faces) {
    /* boundary_conditions.f90:257-272 */
    for (int k = 1; k <= n3; k++)
        for (int i = 1; i <= n1; i++) {
            if (faces) { F[I3(n1, n2, i, 1, k)] = 0.0; F[I3(n1, n2, i, n2, k)] = 0.0; }
            else {
                F[I3(n1, n2, i, 1, k)] = -F[I3(n1, n2, i, 2, k)];
                F[I3(n1, n2, i, n2, k)] = -F[I3(n1, n2, i, n2 - 1, k)];
            }
        }
}
static void neumann_y(double *F, int n1, int n2, int n3) {
    /* boundary_conditions.f90:284-299, centres */
    for (int k = 1; k <= n3; k++)
        for (int i = 1; i <= n1; i++) {
            F[I3(n1, n2, i, 1, k)] = F[I3(n1, n2, i, 2, k)];
            F[I3(n1, n2, i, n2, k)] = F[I3(n1, n2, i, n2 - 1, k)];
        }
}
static void nonzero_neumann_y(const orc_t *o, double *F, const double *alpha, int n1, int n2, int n3) {
    /* boundary_conditions.f90:311-327, centres (id 2) */
    double d1 = o->yg[1] - o->yg[0];
    double d2 = o->yg[o->nyg - 1] - o->yg[o->nyg - 2];
    for (int k = 1; k <= n3; k++)
        for (int i = 1; i <= n1; i++) {
            F[I3(n1, n2, i, 1, k)] = F[I3(n1, n2, i, 2, k)] - d1 * alpha[I3(n1, 2, i, 1, k)];
            F[I3(n1, n2, i, n2, k)] = F[I3(n1, n2, i, n2 - 1, k)] - d2 * alpha[I3(n1, 2, i, 2, k)];
        }
}

void orc_apply_boundary_conditions(orc_t *o) {
    periodic_x(o, o->U, o->nx, o->nyg, o->nzg, 1);
    periodic_x(o, o->V, o->nxg, o->ny, o->nzg, 0);
    periodic_x(o, o->W, o->nxg, o->nyg, o->nz, 0);
    periodic_z(o->U, o->nx, o->nyg, o->nzg, 0);
    periodic_z(o->V, o->nxg, o->ny, o->nzg, 0);
    periodic_z(o->W, o->nxg, o->nyg, o->nz, 1);
    if (o->p.iwall_model == 0) dirichlet_y(o->U, o->nx, o->nyg, o->nzg, 0);
    else nonzero_neumann_y(o, o->U, o->alpha_x, o->nx, o->nyg, o->nzg);
    dirichlet_y(o->V, o->nxg, o->ny, o->nzg, 1);
    if (o->p.iwall_model == 0) dirichlet_y(o->W, o->nxg, o->nyg, o->nz, 0);
    else nonzero_neumann_y(o, o->W, o->alpha_z, o->nxg, o->nyg, o->nz);
}

/* ======================================================================
 * equations.f90:24-498 -- right-hand sides, evaluated point by point with
 * the reference's term order:  (((conv_1 + conv_2) + conv_3) + visc) + dP
 * ====================================================================== */
void orc_compute_rhs(orc_t *o, const double *U, const double *V, const double *W,
                     double *rhs_u, double *rhs_v, double *rhs_w) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    const double *w0 = o->w0, *w1 = o->w1, *nut = o->nu_t;
    const double nu = o->p.nu, dPdx = o->dPdx, dPdz = o->dPdz, Om = o->p.omega_z;
    const int rot = fabs(Om) > 1e-6;
#define X(i) x[(i)-1]
#define Yf(j) y[(j)-1]
#define Z(k) z[(k)-1]
#define XG(i) xg[(i)-1]
#define YG(j) yg[(j)-1]
#define ZG(k) zg[(k)-1]
#define W0(j) w0[(j)-1]
#define W1(j) w1[(j)-1]
#define u(i, j, k) U[IU(i, j, k)]
#define v(i, j, k) V[IV(i, j, k)]
#define w(i, j, k) W[IW(i, j, k)]
#define nt(i, j, k) nut[IC(i, j, k)]
    if (rhs_u) {
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nx - 1; i++) {
                    /* 1) -du^2/dx (:43-57) */
                    double c1 = 0.5 * (u(i, j, k) + u(i + 1, j, k));
                    double c0 = 0.5 * (u(i - 1, j, k) + u(i, j, k));
                    double t1 = -(sq(c1) - sq(c0)) / (XG(i + 1) - XG(i));
                    /* 2) -duv/dy (:62-79) */
                    double uv1 = PLANT(4, W1(j) * u(i, j, k) + W0(j) * u(i, j + 1, k),
                                       W0(j) * u(i, j, k) + W1(j) * u(i, j + 1, k)) *
                                 (0.5 * (v(i, j, k) + v(i + 1, j, k)));
                    double uv0 = PLANT(4, W1(j - 1) * u(i, j - 1, k) + W0(j - 1) * u(i, j, k),
                                       W0(j - 1) * u(i, j - 1, k) + W1(j - 1) * u(i, j, k)) *
                                 (0.5 * (v(i, j - 1, k) + v(i + 1, j - 1, k)));
                    double t2 = -(uv1 - uv0) / (Yf(j) - Yf(j - 1));
                    /* 3) -duw/dz (:84-101) */
                    double uw1 = (0.5 * (u(i, j, k) + u(i, j, k + 1))) * (0.5 * (w(i, j, k) + w(i + 1, j, k)));
                    double uw0 = (0.5 * (u(i, j, k - 1) + u(i, j, k))) * (0.5 * (w(i, j, k - 1) + w(i + 1, j, k - 1)));
                    double t3 = -(uw1 - uw0) / (Z(k) - Z(k - 1));
                    /* 4) viscous (:108-161) */
                    double dz_1 = ZG(k) - ZG(k - 1), dz_2 = ZG(k + 1) - ZG(k), dz_3 = Z(k) - Z(k - 1);
                    double dy_3 = Yf(j) - Yf(j - 1);
                    double dx_1 = X(i) - X(i - 1), dx_2 = X(i + 1) - X(i), dx_3 = XG(i + 1) - XG(i);
                    double nu_x1 = nu + nt(i, j, k), nu_x2 = nu + nt(i + 1, j, k);
                    const int jw = PLANT(1, j, j - 1);
                    double nu_y1 = nu + 0.5 * (W0(jw) * nt(i, j - 1, k) + W1(jw) * nt(i, j, k) +
                                               W0(jw) * nt(i + 1, j - 1, k) + W1(jw) * nt(i + 1, j, k));
                    double nu_y2 = nu + 0.5 * (W0(j) * nt(i, j, k) + W1(j) * nt(i, j + 1, k) +
                                               W0(j) * nt(i + 1, j, k) + W1(j) * nt(i + 1, j + 1, k));
                    double nu_z1 = nu + 0.25 * (nt(i, j, k) + nt(i, j, k - 1) + nt(i + 1, j, k) + nt(i + 1, j, k - 1));
                    double nu_z2 = nu + 0.25 * (nt(i, j, k) + nt(i, j, k + 1) + nt(i + 1, j, k) + nt(i + 1, j, k + 1));
                    double dudy1 = (u(i, j + 1, k) - u(i, j, k)) / (YG(j + 1) - YG(j));
                    double dudy0 = (u(i, j, k) - u(i, j - 1, k)) / (YG(j) - YG(j - 1));
                    double visc =
                        2.0 / dx_3 * (nu_x2 * (u(i + 1, j, k) - u(i, j, k)) / dx_2 - nu_x1 * (u(i, j, k) - u(i - 1, j, k)) / dx_1) +
                        1.0 / dy_3 * (nu_y2 * (dudy1 + (v(i + 1, j, k) - v(i, j, k)) / dx_2) -
                                      nu_y1 * (dudy0 + (v(i + 1, j - 1, k) - v(i, j - 1, k)) / dx_1)) +
                        1.0 / dz_3 * (nu_z2 * ((u(i, j, k + 1) - u(i, j, k)) / dz_2 + (w(i + 1, j, k) - w(i, j, k)) / dx_2) -
                                      nu_z1 * ((u(i, j, k) - u(i, j, k - 1)) / dz_1 + (w(i + 1, j, k - 1) - w(i, j, k - 1)) / dx_1));
                    double r = t1 + t2;
                    r = r + t3;
                    r = r + visc;
                    r = r + dPdx;
                    if (rot) { /* :169-177 */
                        const int jc = PLANT(7, j + 1, j);
                        double a = 0.5 * (v(i, jc - 1, k) + v(i, jc, k));
                        double b = 0.5 * (v(i + 1, jc - 1, k) + v(i + 1, jc, k));
                        r = r + 2.0 * Om * (0.5 * (a + b));
                    }
                    rhs_u[IRU(i, j, k)] = r;
                }
    }
    if (rhs_v) {
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    /* 1) -dv^2/dy (:205-223) */
                    double c1 = 0.5 * (v(i, j, k) + v(i, j + 1, k));
                    double c0 = 0.5 * (v(i, j - 1, k) + v(i, j, k));
                    double t1 = -(sq(c1) - sq(c0)) / PLANT(8, Yf(j + 1) - Yf(j), YG(j + 1) - YG(j));
                    /* 2) -duv/dx (:228-245) */
                    double uv1 = (W0(j) * u(i, j, k) + W1(j) * u(i, j + 1, k)) * (0.5 * (v(i, j, k) + v(i + 1, j, k)));
                    double uv0 = (W0(j) * u(i - 1, j, k) + W1(j) * u(i - 1, j + 1, k)) * (0.5 * (v(i - 1, j, k) + v(i, j, k)));
                    double t2 = -(uv1 - uv0) / (X(i) - X(i - 1));
                    /* 3) -dvw/dz (:250-267) */
                    double vw1 = (0.5 * (v(i, j, k) + v(i, j, k + 1))) * (W0(j) * w(i, j, k) + W1(j) * w(i, j + 1, k));
                    double vw0 = (0.5 * (v(i, j, k - 1) + v(i, j, k))) * (W0(j) * w(i, j, k - 1) + W1(j) * w(i, j + 1, k - 1));
                    double t3 = -(vw1 - vw0) / (Z(k) - Z(k - 1));
                    /* 4) viscous (:276-325) */
                    double dz_1 = ZG(k) - ZG(k - 1), dz_2 = ZG(k + 1) - ZG(k), dz_3 = Z(k) - Z(k - 1);
                    double dy_1 = PLANT(3, Yf(j + 1) - Yf(j), Yf(j) - Yf(j - 1));
                    double dy_2 = PLANT(3, Yf(j) - Yf(j - 1), Yf(j + 1) - Yf(j));
                    double dy_3 = PLANT(2, Yf(j + 1) - Yf(j), YG(j + 1) - YG(j));
                    double dx_1 = XG(i) - XG(i - 1), dx_2 = XG(i + 1) - XG(i), dx_3 = X(i) - X(i - 1);
                    const int jx = PLANT(14, j - 1, j);
                    double nu_x1 = nu + 0.5 * (W0(jx) * nt(i - 1, j, k) + W1(jx) * nt(i - 1, j + 1, k) +
                                               W0(jx) * nt(i, j, k) + W1(jx) * nt(i, j + 1, k));
                    double nu_x2 = nu + 0.5 * (W0(j) * nt(i, j, k) + W1(j) * nt(i, j + 1, k) +
                                               W0(j) * nt(i + 1, j, k) + W1(j) * nt(i + 1, j + 1, k));
                    double nu_y1 = nu + PLANT(9, nt(i, j - 1, k), nt(i, j, k)), nu_y2 = nu + nt(i, j + 1, k);
                    double nu_z1 = nu + 0.5 * (W0(j) * nt(i, j, k - 1) + W1(j) * nt(i, j + 1, k - 1) +
                                               W0(j) * nt(i, j, k) + W1(j) * nt(i, j + 1, k));
                    double nu_z2 = nu + 0.5 * (W0(j) * nt(i, j, k) + W1(j) * nt(i, j + 1, k) +
                                               W0(j) * nt(i, j, k + 1) + W1(j) * nt(i, j + 1, k + 1));
                    double visc =
                        1.0 / dx_3 * (nu_x2 * ((v(i + 1, j, k) - v(i, j, k)) / dx_2 + (u(i, j + 1, k) - u(i, j, k)) / dy_3) -
                                      nu_x1 * ((v(i, j, k) - v(i - 1, j, k)) / dx_1 + (u(i - 1, j + 1, k) - u(i - 1, j, k)) / dy_3)) +
                        2.0 / dy_3 * (nu_y2 * (v(i, j + 1, k) - v(i, j, k)) / dy_2 - nu_y1 * (v(i, j, k) - v(i, j - 1, k)) / dy_1) +
                        1.0 / dz_3 * (nu_z2 * ((v(i, j, k + 1) - v(i, j, k)) / dz_2 + (w(i, j + 1, k) - w(i, j, k)) / dy_3) -
                                      nu_z1 * ((v(i, j, k) - v(i, j, k - 1)) / dz_1 + (w(i, j + 1, k - 1) - w(i, j, k - 1)) / dy_3));
                    double r = t1 + t2;
                    r = r + t3;
                    r = r + visc;
                    if (rot) { /* :334-342 */
                        double a = 0.5 * (u(i - 1, j, k) + u(i, j, k));
                        double b = 0.5 * (u(i - 1, j + 1, k) + u(i, j + 1, k));
                        r = PLANT(6, r + 2.0 * Om * (W0(j) * a + W1(j) * b), r - 2.0 * Om * (W0(j) * a + W1(j) * b));
                    }
                    rhs_v[IRV(i, j, k)] = r;
                }
    }
    if (rhs_w) {
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    /* 1) -dw^2/dz (:370-384) */
                    double c1 = 0.5 * (w(i, j, k) + w(i, j, k + 1));
                    double c0 = 0.5 * (w(i, j, k - 1) + w(i, j, k));
                    double t1 = -(sq(c1) - sq(c0)) / (ZG(k + 1) - ZG(k));
                    /* 2) -duw/dx (:389-406) */
                    double uw1 = (0.5 * (u(i, j, k) + u(i, j, k + 1))) * (0.5 * (w(i, j, k) + w(i + 1, j, k)));
                    double uw0 = (0.5 * (u(i - 1, j, k) + u(i - 1, j, k + 1))) * (0.5 * (w(i - 1, j, k) + w(i, j, k)));
                    double t2 = -(uw1 - uw0) / (X(i) - X(i - 1));
                    /* 3) -dvw/dy (:411-428) */
                    double vw1 = (0.5 * (v(i, j, k) + v(i, j, k + 1))) * (W0(j) * w(i, j, k) + W1(j) * w(i, j + 1, k));
                    const int jv = PLANT(5, j, j - 1);
                    double vw0 = (0.5 * (v(i, j - 1, k) + v(i, j - 1, k + 1))) * (W0(jv) * w(i, j - 1, k) + W1(jv) * w(i, j, k));
                    double t3 = -(vw1 - vw0) / (Yf(j) - Yf(j - 1));
                    /* 4) viscous (:435-491) */
                    double dz_1 = Z(k) - Z(k - 1), dz_2 = Z(k + 1) - Z(k), dz_3 = ZG(k + 1) - ZG(k);
                    double dy_3 = Yf(j) - Yf(j - 1);
                    double dx_1 = XG(i) - XG(i - 1), dx_2 = XG(i + 1) - XG(i), dx_3 = X(i) - X(i - 1);
                    double nu_x1 = nu + 0.25 * (nt(i, j, k) + nt(i, j, k + 1) + nt(i - 1, j, k) + nt(i - 1, j, k + 1));
                    double nu_x2 = nu + 0.25 * (nt(i, j, k) + nt(i, j, k + 1) + nt(i + 1, j, k) + nt(i + 1, j, k + 1));
                    double nu_y1 = nu + 0.5 * (W0(j - 1) * nt(i, j - 1, k) + W1(j - 1) * nt(i, j, k) +
                                               W0(j - 1) * nt(i, j - 1, k + 1) + W1(j - 1) * nt(i, j, k + 1));
                    double nu_y2 = nu + 0.5 * (W0(j) * nt(i, j, k) + W1(j) * nt(i, j + 1, k) +
                                               W0(j) * nt(i, j, k + 1) + W1(j) * nt(i, j + 1, k + 1));
                    double nu_z1 = nu + nt(i, j, k), nu_z2 = nu + nt(i, j, k + 1);
                    double dwdy1 = (w(i, j + 1, k) - w(i, j, k)) / (YG(j + 1) - YG(j));
                    double dwdy0 = (w(i, j, k) - w(i, j - 1, k)) / (YG(j) - YG(j - 1));
                    double visc =
                        1.0 / dx_3 * (nu_x2 * ((w(i + 1, j, k) - w(i, j, k)) / dx_2 + (u(i, j, k + 1) - u(i, j, k)) / dz_2) -
                                      nu_x1 * ((w(i, j, k) - w(i - 1, j, k)) / dx_1 + (u(i - 1, j, k + 1) - u(i - 1, j, k)) / dz_1)) +
                        1.0 / dy_3 * (nu_y2 * (dwdy1 + (v(i, j, k + 1) - v(i, j, k)) / dz_2) -
                                      nu_y1 * (dwdy0 + (v(i, j - 1, k + 1) - v(i, j - 1, k)) / dz_1)) +
                        2.0 / dz_3 * (nu_z2 * (w(i, j, k + 1) - w(i, j, k)) / dz_2 - nu_z1 * (w(i, j, k) - w(i, j, k - 1)) / dz_1);
                    double r = t1 + t2;
                    r = r + t3;
                    r = r + visc;
                    r = r + dPdz;
                    rhs_w[IRW(i, j, k)] = r;
                }
    }
}

/* ======================================================================
 * FFT (restating FFTW's unnormalised DFT: sum x_j exp(sign 2 pi i jk/n))
 * mixed-radix recursive decimation in time
 * ====================================================================== */
static void fft_rec(int n, const cplx *xin, int s, cplx *yout, int sign) {
    if (n == 1) { yout[0] = xin[0]; return; }
    int p = n;
    for (int f = 2; f * f <= n; f++) if (n % f == 0) { p = f; break; }
    int m = n / p;
    for (int r = 0; r < p; r++) fft_rec(m, xin + (size_t)r * s, s * p, yout + (size_t)r * m, sign);
    const double tp = 2.0 * 3.14159265358979323846 / (double)n;
    cplx tmp[64];
    cplx *t = p <= 64 ? tmp : (cplx *)malloc(p * sizeof(cplx));
    cplx *acc = (cplx *)malloc(p * sizeof(cplx));
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            long e = ((long)r * k) % n;
            cplx wv = cos(tp * e) + sign * I * sin(tp * e);
            t[r] = yout[r * m + k] * (r ? wv : 1.0);
        }
        for (int q = 0; q < p; q++) {
            cplx sum = 0;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * q % p) * m; /* W_p^{rq} = W_n^{rq m} */
                cplx wv = cos(tp * e) + sign * I * sin(tp * e);
                sum += t[r] * (e ? wv : 1.0);
            }
            acc[q] = sum;
        }
        for (int q = 0; q < p; q++) yout[q * m + k] = acc[q];
    }
    if (t != tmp) free(t);
    free(acc);
}

static void fft_strided(int n, cplx *a, int stride, int sign, cplx *w1, cplx *w2) {
    for (int q = 0; q < n; q++) w1[q] = a[(size_t)q * stride];
    fft_rec(n, w1, 1, w2, sign);
    for (int q = 0; q < n; q++) a[(size_t)q * stride] = w2[q];
}

void orc_fft(int n, double *re, double *im, int sign) {
    cplx *a = (cplx *)malloc(n * sizeof(cplx)), *b = (cplx *)malloc(n * sizeof(cplx));
    for (int q = 0; q < n; q++) a[q] = re[q] + I * im[q];
    fft_rec(n, a, 1, b, sign);
    for (int q = 0; q < n; q++) { re[q] = creal(b[q]); im[q] = cimag(b[q]); }
    free(a); free(b);
}

static void fft2d_plane(int n1, int n2, cplx *a, int sign, cplx *w1, cplx *w2) {
    for (int k = 0; k < n2; k++) fft_strided(n1, a + (size_t)k * n1, 1, sign, w1, w2);
    for (int i = 0; i < n1; i++) fft_strided(n2, a + i, n1, sign, w1, w2);
}

void orc_fft2d(int n1, int n2, double *re, double *im, int sign) {
    size_t n = (size_t)n1 * n2;
    int mx = n1 > n2 ? n1 : n2;
    cplx *a = (cplx *)malloc(n * sizeof(cplx));
    cplx *w1 = (cplx *)malloc(mx * sizeof(cplx)), *w2 = (cplx *)malloc(mx * sizeof(cplx));
    for (size_t q = 0; q < n; q++) a[q] = re[q] + I * im[q];
    fft2d_plane(n1, n2, a, sign, w1, w2);
    for (size_t q = 0; q < n; q++) { re[q] = creal(a[q]); im[q] = cimag(a[q]); }
    free(a); free(w1); free(w2);
}

/* LAPACK zgtsv (Gaussian elimination with partial pivoting) for a real
   tridiagonal matrix and one complex right-hand side.  dl, d, du are
   overwritten as in LAPACK.  Returns INFO. */
int orc_zgtsv(int n, double *dl, double *d, double *du, double *br, double *bi) {
    for (int k = 0; k < n - 1; k++) {
        if (dl[k] == 0.0) {
            if (d[k] == 0.0) return k + 1;
        } else if (fabs(d[k]) >= fabs(dl[k])) {
            double mult = dl[k] / d[k];
            d[k + 1] = d[k + 1] - mult * du[k];
            br[k + 1] = br[k + 1] - mult * br[k];
            bi[k + 1] = bi[k + 1] - mult * bi[k];
            if (k < n - 2) dl[k] = 0.0;
        } else {
            double mult = d[k] / dl[k];
            d[k] = dl[k];
            double temp = d[k + 1];
            d[k + 1] = du[k] - mult * temp;
            if (k < n - 2) { dl[k] = du[k + 1]; du[k + 1] = -mult * dl[k]; }
            du[k] = temp;
            double tr = br[k], ti = bi[k];
            br[k] = br[k + 1]; bi[k] = bi[k + 1];
            br[k + 1] = tr - mult * br[k + 1];
            bi[k + 1] = ti - mult * bi[k + 1];
        }
    }
    if (d[n - 1] == 0.0) return n;
    br[n - 1] = br[n - 1] / d[n - 1];
    bi[n - 1] = bi[n - 1] / d[n - 1];
    if (n > 1) {
        br[n - 2] = (br[n - 2] - du[n - 2] * br[n - 1]) / d[n - 2];
        bi[n - 2] = (bi[n - 2] - du[n - 2] * bi[n - 1]) / d[n - 2];
    }
    for (int k = n - 3; k >= 0; k--) {
        br[k] = (br[k] - du[k] * br[k + 1] - dl[k] * br[k + 2]) / d[k];
        bi[k] = (bi[k] - du[k] * bi[k + 1] - dl[k] * bi[k + 2]) / d[k];
    }
    return 0;
}

/* ======================================================================
 * projection.f90:20-225
 * ====================================================================== */
void orc_projection_step(orc_t *o, int rk_step) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    double *U = o->U, *V = o->V, *W = o->W, *rp = o->rhs_p;
    const int nxp = o->nxp, nzp = o->nzp, nr = o->nr;
    /* compute_pseudo_pressure_rhs (:40-55) */
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                rp[IP(i, j, k)] = (u(i, j, k) - u(i - 1, j, k)) / (X(i) - X(i - 1)) +
                                  (v(i, j, k) - v(i, j - 1, k)) / (Yf(j) - Yf(j - 1)) +
                                  (w(i, j, k) - w(i, j, k - 1)) / (Z(k) - Z(k - 1));
    /* solve_poisson_equation (:83-120): 2D DFT per plane, zgtsv per mode */
    size_t npl = (size_t)nxp * nzp;
    cplx *hat = (cplx *)malloc(npl * nr * sizeof(cplx)); /* [j][kz][kx] */
    int mx = nxp > nzp ? nxp : nzp;
#pragma omp parallel
    {
        cplx *w1 = (cplx *)malloc(mx * sizeof(cplx)), *w2 = (cplx *)malloc(mx * sizeof(cplx));
#pragma omp for schedule(static)
        for (int j = 2; j <= nyg - 1; j++) {
            cplx *pl = hat + (size_t)(j - 2) * npl;
            for (int kk = 1; kk <= nzp; kk++)
                for (int ii = 1; ii <= nxp; ii++)
                    pl[(size_t)(kk - 1) * nxp + (ii - 1)] = rp[IP(ii + 1, j, kk + 1)];
            fft2d_plane(nxp, nzp, pl, -1, w1, w2);
        }
        double *D = dalloc(nr), *DL = dalloc(nr), *DU = dalloc(nr), *br = dalloc(nr), *bi = dalloc(nr);
#pragma omp for schedule(static)
        for (int kz = 0; kz < nzp; kz++)
            for (int kx = 0; kx < nxp; kx++) {
                for (int q = 0; q < nr; q++) D[q] = o->dyy_d[q] + o->kxx[kx] + o->kzz[kz];
                for (int q = 0; q < nr - 1; q++) { DL[q] = o->dyy_l[q]; DU[q] = o->dyy_u[q]; }
                if (kx == 0 && kz == 0) D[0] = 3.0 / 2.0 * D[0];
                for (int q = 0; q < nr; q++) {
                    cplx h = hat[(size_t)q * npl + (size_t)kz * nxp + kx];
                    br[q] = creal(h); bi[q] = cimag(h);
                }
                orc_zgtsv(nr, DL, D, DU, br, bi);
                for (int q = 0; q < nr; q++) hat[(size_t)q * npl + (size_t)kz * nxp + kx] = br[q] + I * bi[q];
            }
        free(D); free(DL); free(DU); free(br); free(bi);
        const double norm = (double)(nxp * nzp);
#pragma omp for schedule(static)
        for (int j = 2; j <= nyg - 1; j++) {
            cplx *pl = hat + (size_t)(j - 2) * npl;
            fft2d_plane(nxp, nzp, pl, +1, w1, w2);
            for (int kk = 1; kk <= nzp; kk++)
                for (int ii = 1; ii <= nxp; ii++)
                    rp[IP(ii + 1, j, kk + 1)] = creal(pl[(size_t)(kk - 1) * nxp + (ii - 1)] / norm);
        }
        free(w1); free(w2);
    }
    free(hat);
    /* apply_periodic_xz_pressure (:144-160), intended one-rank semantics */
    for (int k = 2; k <= nzg; k++)
        for (int j = 2; j <= nyg - 1; j++) rp[IP(nxg - 1, j, k)] = rp[IP(2, j, k)];
    for (int j = 2; j <= nyg - 1; j++)
        for (int i = 2; i <= nxg - 1; i++) rp[IP(i, j, nzp + 2)] = rp[IP(i, j, 2)];
    /* save pressure (:129-137) */
    if (rk_step == 1) {
        double den = o->dt * o->rk_coef[0][0];
        double *P = o->P;
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) P[IC(i, j, k)] = rp[IP(i, j, k)] / den;
        for (int k = 1; k <= nzg; k++)
            for (int j = 1; j <= nyg; j++) { P[IC(1, j, k)] = P[IC(2, j, k)]; P[IC(nxg, j, k)] = P[IC(nxg - 1, j, k)]; }
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg; i++) { P[IC(i, 1, k)] = P[IC(i, 2, k)]; P[IC(i, nyg, k)] = P[IC(i, nyg - 1, k)]; }
        size_t pl = (size_t)nxg * nyg;
        memcpy(P, P + pl, pl * sizeof(double));
        memcpy(P + pl * (nzg - 1), P + pl * (nzg - 2), pl * sizeof(double));
    }
    /* project_velocity (:203-225) */
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nx - 1; i++)
                u(i, j, k) = u(i, j, k) - (rp[IP(i + 1, j, k)] - rp[IP(i, j, k)]) / (XG(i + 1) - XG(i));
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= ny - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                v(i, j, k) = v(i, j, k) - (rp[IP(i, j + 1, k)] - rp[IP(i, j, k)]) / (YG(j + 1) - YG(j));
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nz - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                w(i, j, k) = w(i, j, k) - (rp[IP(i, j, k + 1)] - rp[IP(i, j, k)]) / (ZG(k + 1) - ZG(k));
}

/* ======================================================================
 * pressure.f90:21-341, compute_pressure: the actual pressure from
 *   Laplacian(P) = div(rhs_u, rhs_v, rhs_w),
 *   dP/dy = d/dy(nu dV/dy) at the walls (bc_1 bottom, bc_2 top).
 * Optional outputs: rhs_p (2:nxg-1, 2:nyg-1, 2:nzg-1) and the two wall
 * conditions (2:nxg-1, 2:nzg-1) each, for tests.  Ghosts the one-rank
 * reference leaves unset (i = 1, nxg; k = 1, nzg) are copies of the
 * neighbouring plane, as the pseudo-pressure save does (projection.f90:130-135).
 * ====================================================================== */
#define IB(i, k) I3(o->nxg - 2, 1, (i)-1, 1, (k)-1)
void orc_compute_pressure(orc_t *o, double *Pout, double *rhs_p_out, double *bc1_out, double *bc2_out) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *x = o->x, *y = o->y, *z = o->z, *yg = o->yg;
    const double *V = o->V, *avg = o->avg_nu_t;
    const int nxp = o->nxp, nzp = o->nzp, nr = o->nr;
    const size_t nb = (size_t)(nxg - 2) * (nzg - 2);
    double *bc1 = dalloc(nb), *bc2 = dalloc(nb);
    /* compute_pressure_BC (:256-341) */
    {
        const double dzeta = 1.0;
        double g = (-3.0 * Yf(1) + 4.0 * Yf(2) - Yf(3)) / (2.0 * dzeta);
        double g2 = (2.0 * Yf(1) - 5.0 * Yf(2) + 4.0 * Yf(3) - 1.0 * Yf(4)) / pow(dzeta, 2.0);
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) {
                double nui = o->p.nu + 0.5 * (avg[0] + avg[1]);
                double dnu = (avg[1] - avg[0]) / (YG(2) - YG(1));
                double dV = (-3.0 * v(i, 1, k) + 4.0 * v(i, 2, k) - v(i, 3, k)) / (2.0 * dzeta) * 1.0 / g;
                double ddV = (2.0 * v(i, 1, k) - 5.0 * v(i, 2, k) + 4.0 * v(i, 3, k) - 1.0 * v(i, 4, k)) / pow(dzeta, 2.0);
                ddV = (ddV - dV * g2) / pow(g, 2.0);
                bc1[IB(i, k)] = dnu * dV + nui * ddV;
            }
        g = (3.0 * Yf(ny) - 4.0 * Yf(ny - 1) + Yf(ny - 2)) / (2.0 * dzeta);
        g2 = (2.0 * Yf(ny) - 5.0 * Yf(ny - 1) + 4.0 * Yf(ny - 2) - 1.0 * Yf(ny - 3)) / pow(dzeta, 2.0);
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) {
                double nui = o->p.nu + 0.5 * (avg[nyg - 1] + avg[nyg - 2]);
                double dnu = (avg[nyg - 1] - avg[nyg - 2]) / (YG(nyg) - YG(nyg - 1));
                double dV = (3.0 * v(i, ny, k) - 4.0 * v(i, ny - 1, k) + v(i, ny - 2, k)) / (2.0 * dzeta) * 1.0 / g;
                double ddV = (2.0 * v(i, ny, k) - 5.0 * v(i, ny - 1, k) + 4.0 * v(i, ny - 2, k) - 1.0 * v(i, ny - 3, k)) /
                             pow(dzeta, 2.0);
                ddV = (ddV - dV * g2) / pow(g, 2.0);
                bc2[IB(i, k)] = dnu * dV + nui * ddV;
            }
    }
    /* compute_pressure_rhs (:43-73): rhs_u, rhs_v, rhs_w on full arrays with
       the x-periodic rhs_u, the wall rows of rhs_v set to the conditions and
       the z-periodic rhs_w, then their divergence */
    double *ruf = dalloc(o->nU), *rvf = dalloc(o->nV), *rwf = dalloc(o->nW);
    {
        double *ru = dalloc(o->nRu), *rv = dalloc(o->nRv), *rw = dalloc(o->nRw);
        orc_compute_rhs(o, o->U, o->V, o->W, ru, rv, rw);
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nx - 1; i++) ruf[IU(i, j, k)] = ru[IRU(i, j, k)];
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) rvf[IV(i, j, k)] = rv[IRV(i, j, k)];
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) rwf[IW(i, j, k)] = rw[IRW(i, j, k)];
        free(ru); free(rv); free(rw);
    }
    periodic_x(o, ruf, nx, nyg, nzg, 1);
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nxg - 1; i++) { rvf[IV(i, 1, k)] = bc1[IB(i, k)]; rvf[IV(i, ny, k)] = bc2[IB(i, k)]; }
    periodic_z(rwf, nxg, nyg, nz, 1);
    double *rp = o->rhs_p;
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                rp[IP(i, j, k)] = (ruf[IU(i, j, k)] - ruf[IU(i - 1, j, k)]) / (X(i) - X(i - 1)) +
                                  (rvf[IV(i, j, k)] - rvf[IV(i, j - 1, k)]) / (Yf(j) - Yf(j - 1)) +
                                  (rwf[IW(i, j, k)] - rwf[IW(i, j, k - 1)]) / (Z(k) - Z(k - 1));
    free(ruf); free(rvf); free(rwf);
    if (rhs_p_out)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++)
                    rhs_p_out[I3(nxg - 2, nyg - 2, i - 1, j - 1, k - 1)] = rp[IP(i, j, k)];
    if (bc1_out) memcpy(bc1_out, bc1, nb * sizeof(double));
    if (bc2_out) memcpy(bc2_out, bc2, nb * sizeof(double));
    /* solve_poisson_equation_for_pressure (:94-151) */
    const double cb1 = 1.0 / (Yf(2) - Yf(1)) / (YG(2) - YG(1));               /* coef_bc_1 (initialization.f90:364) */
    const double cb2 = 1.0 / (Yf(nyg - 1) - Yf(nyg - 2)) / (YG(nyg) - YG(nyg - 1)); /* coef_bc_2 (:373) */
    size_t npl = (size_t)nxp * nzp;
    cplx *hat = (cplx *)malloc(npl * (nr + 2) * sizeof(cplx)); /* [j][kz][kx], then bc_1, bc_2 */
    int mx = nxp > nzp ? nxp : nzp;
#pragma omp parallel
    {
        cplx *w1 = (cplx *)malloc(mx * sizeof(cplx)), *w2 = (cplx *)malloc(mx * sizeof(cplx));
#pragma omp for schedule(static)
        for (int j = 2; j <= nyg + 1; j++) { /* the nr planes, then the two conditions */
            cplx *pl = hat + (size_t)(j - 2) * npl;
            for (int kk = 1; kk <= nzp; kk++)
                for (int ii = 1; ii <= nxp; ii++) {
                    double val = j <= nyg - 1 ? rp[IP(ii + 1, j, kk + 1)]
                                              : (j == nyg ? bc1 : bc2)[IB(ii + 1, kk + 1)];
                    pl[(size_t)(kk - 1) * nxp + (ii - 1)] = val;
                }
            fft2d_plane(nxp, nzp, pl, -1, w1, w2);
        }
        double *D = dalloc(nr), *DL = dalloc(nr), *DU = dalloc(nr), *br = dalloc(nr), *bi = dalloc(nr);
#pragma omp for schedule(static)
        for (int kz = 0; kz < nzp; kz++)
            for (int kx = 0; kx < nxp; kx++) {
                for (int q = 0; q < nr; q++) D[q] = o->dyy_d[q] + o->kxx[kx] + o->kzz[kz];
                for (int q = 0; q < nr - 1; q++) { DL[q] = o->dyy_l[q]; DU[q] = o->dyy_u[q]; }
                if (kx == 0 && kz == 0) D[0] = 3.0 / 2.0 * D[0];
                for (int q = 0; q < nr; q++) {
                    cplx h = hat[(size_t)q * npl + (size_t)kz * nxp + kx];
                    br[q] = creal(h); bi[q] = cimag(h);
                }
                const cplx b1 = hat[(size_t)nr * npl + (size_t)kz * nxp + kx];
                const cplx b2 = hat[(size_t)(nr + 1) * npl + (size_t)kz * nxp + kx];
                const cplx r0 = (br[0] + I * bi[0]) + cb1 * (YG(2) - YG(1)) * b1;
                const cplx rn = (br[nr - 1] + I * bi[nr - 1]) - cb2 * (YG(nyg) - YG(nyg - 1)) * b2;
                br[0] = creal(r0); bi[0] = cimag(r0);
                br[nr - 1] = creal(rn); bi[nr - 1] = cimag(rn);
                orc_zgtsv(nr, DL, D, DU, br, bi);
                for (int q = 0; q < nr; q++) hat[(size_t)q * npl + (size_t)kz * nxp + kx] = br[q] + I * bi[q];
            }
        free(D); free(DL); free(DU); free(br); free(bi);
        const double norm = (double)(nxp * nzp);
#pragma omp for schedule(static)
        for (int j = 2; j <= nyg - 1; j++) {
            cplx *pl = hat + (size_t)(j - 2) * npl;
            fft2d_plane(nxp, nzp, pl, +1, w1, w2);
            for (int kk = 1; kk <= nzp; kk++)
                for (int ii = 1; ii <= nxp; ii++)
                    Pout[IC(ii + 1, j, kk + 1)] = creal(pl[(size_t)(kk - 1) * nxp + (ii - 1)] / norm);
        }
        free(w1); free(w2);
    }
    free(hat);
    /* ghost rows in y from the wall conditions (:155-158), then the periodic
       copies (:177-193, intended one-rank semantics) */
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nxg - 1; i++) {
            Pout[IC(i, 1, k)] = Pout[IC(i, 2, k)] - (YG(2) - YG(1)) * bc1[IB(i, k)];
            Pout[IC(i, nyg, k)] = Pout[IC(i, nyg - 1, k)] + (YG(nyg) - YG(nyg - 1)) * bc2[IB(i, k)];
        }
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) Pout[IC(nxg - 1, j, k)] = Pout[IC(2, j, k)];
    for (int j = 1; j <= nyg; j++)
        for (int i = 2; i <= nxg - 1; i++) Pout[IC(i, j, nzg - 1)] = Pout[IC(i, j, 2)];
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) { Pout[IC(1, j, k)] = Pout[IC(2, j, k)]; Pout[IC(nxg, j, k)] = Pout[IC(nxg - 1, j, k)]; }
    {
        size_t pl = (size_t)nxg * nyg;
        memcpy(Pout, Pout + pl, pl * sizeof(double));
        memcpy(Pout + pl * (nzg - 1), Pout + pl * (nzg - 2), pl * sizeof(double));
    }
    free(bc1); free(bc2);
}
#undef IB

/* projection.f90:234-256 (signed max, starting from 0) plus max |div| */
void orc_check_divergence(const orc_t *o, double *max_signed, double *max_abs) {
    const int nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *x = o->x, *y = o->y, *z = o->z, *U = o->U, *V = o->V, *W = o->W;
    double ds = 0.0, da = 0.0;
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                double d = (u(i, j, k) - u(i - 1, j, k)) / (X(i) - X(i - 1)) +
                           (v(i, j, k) - v(i, j - 1, k)) / (Yf(j) - Yf(j - 1)) +
                           (w(i, j, k) - w(i, j, k - 1)) / (Z(k) - Z(k - 1));
                if (d > ds) ds = d;
                if (fabs(d) > da) da = fabs(d);
            }
    *max_signed = ds;
    *max_abs = da;
}

/* ======================================================================
 * mass_flow.f90:25-132
 * ====================================================================== */
double orc_mean_mass_flow_U(const orc_t *o, const double *U) {
    const double *yg = o->yg;
    double q = 0.0, nrm = 0.0;
    for (int k = 2; k <= o->nzg - 1; k++)
        for (int j = 3; j <= o->nyg - 1; j++)
            for (int i = 2; i <= o->nx - 1; i++) {
                q = q + (u(i, j, k) + u(i, j - 1, k)) * 0.5 * (YG(j) - YG(j - 1));
                nrm = nrm + (1.0 + 1.0) * 0.5 * (YG(j) - YG(j - 1));
            }
    return q / nrm;
}
/* mass_flow.f90:192-209: the plane mean of U(1:nx-2, jc, 1:nzg-2) at
 * jc = Floor(Real(ny/2))+1 (integer division first), over the count
 * (nx_global-2)(nzg_global-2) -- the k range holds the periodic image k = 1
 * of plane nzg-2 as well as that plane */
double orc_mean_centerline_U(const orc_t *o, const double *U) {
    const int jc = PLANT(16, o->ny / 2, o->ny / 2 + 1);
    double s = 0.0;
    for (int k = PLANT(17, 2, 1); k <= o->nzg - 2; k++)
        for (int i = 1; i <= o->nx - 2; i++) s = s + u(i, jc, k);
    return s / (double)((o->nx - 2) * (o->nzg - 2));
}
double orc_mean_mass_flow_V(const orc_t *o, const double *V) {
    const double *y = o->y;
    double q = 0.0, nrm = 0.0;
    for (int k = 2; k <= o->nzg - 1; k++)
        for (int j = 2; j <= o->ny; j++)
            for (int i = 2; i <= o->nxg - 1; i++) {
                q = q + (v(i, j, k) + v(i, j - 1, k)) * 0.5 * (Yf(j) - Yf(j - 1));
                nrm = nrm + (1.0 + 1.0) * 0.5 * (Yf(j) - Yf(j - 1));
            }
    return q / nrm;
}

/* time_integration.f90:179-227 */
double orc_compute_dt(orc_t *o) {
    const double *xg = o->xg, *yg = o->yg, *zg = o->zg, *U = o->U, *V = o->V, *W = o->W;
    double lU = 0.0, lV = 0.0, lW = 0.0;
    for (int i = 2; i <= o->nxg - 1; i++)
        for (int j = 2; j <= o->nyg - 1; j++)
            for (int k = 2; k <= o->nzg - 1; k++) {
                lU = fmax(lU, (XG(i + 1) - XG(i)) / fabs(u(i, j, k)));
                lV = fmax(lV, (YG(j + 1) - YG(j)) / fabs(v(i, j, k)));
                lW = fmax(lW, (ZG(k + 1) - ZG(k)) / fabs(w(i, j, k)));
            }
    double cfl = o->p.cfl, nu = o->p.nu;
    double dc = fmin(fmin(cfl * lU, cfl * lV), cfl * lW);
    double dv = fmin(fmin(cfl * sq(o->dxmin) / nu, cfl * sq(o->dymin) / nu), cfl * sq(o->dzmin) / nu);
    double dt = fmin(dc, dv);
    dt = fmin(dt, 5e-3);
    if (cfl < 0) dt = -cfl;
    o->dt = dt;
    return dt;
}

/* ======================================================================
 * subgrid.f90:891-949 compute_Sij and 564-619 constant Smagorinsky,
 * completed with the ghost sequence of subgrid.f90:521-536
 * ====================================================================== */
/* compute_Sij (subgrid.f90:891-949): Sij at cell centres (the off-diagonal
   components averaged from the edges) and |S| = sqrt(2 Sij Sij).
   S6 holds Sij_(2:nxg,2:nyg,2:nzg,6); Smag is a full (nxg,nyg,nzg) array
   of which the interior 2..nxg-1 x 2..nyg-1 x 2..nzg-1 is written. */
#define SIJ(i, j, k, c) S6[(size_t)(c - 1) * ns + I3(s1, s2, (i)-1, (j)-1, (k)-1)]
static void compute_sij(const orc_t *o, const double *U, const double *V, const double *W, double *S6,
                        double *Smag) {
    const int nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    const int s1 = nxg - 1, s2 = nyg - 1, s3 = nzg - 1;
    const size_t ns = (size_t)s1 * s2 * s3;
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                SIJ(i, j, k, 1) = (u(i, j, k) - u(i - 1, j, k)) / (X(i) - X(i - 1));
                SIJ(i, j, k, 2) = (v(i, j, k) - v(i, j - 1, k)) / (Yf(j) - Yf(j - 1));
                SIJ(i, j, k, 3) = (w(i, j, k) - w(i, j, k - 1)) / (Z(k) - Z(k - 1));
            }
#pragma omp parallel for schedule(static)
    for (int k = 2; k <= nzg; k++)
        for (int j = 2; j <= nyg; j++)
            for (int i = 2; i <= nxg; i++) {
                SIJ(i, j, k, 4) = 0.5 * ((u(i - 1, j, k) - u(i - 1, j - 1, k)) / (YG(j) - YG(j - 1)) + (v(i, j - 1, k) - v(i - 1, j - 1, k)) / (XG(i) - XG(i - 1)));
                SIJ(i, j, k, 5) = 0.5 * ((u(i - 1, j, k) - u(i - 1, j, k - 1)) / (ZG(k) - ZG(k - 1)) + (w(i, j, k - 1) - w(i - 1, j, k - 1)) / (XG(i) - XG(i - 1)));
                SIJ(i, j, k, 6) = 0.5 * ((v(i, j - 1, k) - v(i, j - 1, k - 1)) / (ZG(k) - ZG(k - 1)) + (w(i, j, k - 1) - w(i, j - 1, k - 1)) / (YG(j) - YG(j - 1)));
            }
    /* in-place edge -> centre averaging reads only not-yet-updated values
       (loop order k, j, i ascending), so a serial sweep reproduces it */
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                SIJ(i, j, k, 4) = 0.25 * (SIJ(i, j, k, 4) + SIJ(i + 1, j, k, 4) + SIJ(i, j + 1, k, 4) + SIJ(i + 1, j + 1, k, 4));
                SIJ(i, j, k, 5) = 0.25 * (SIJ(i, j, k, 5) + SIJ(i + 1, j, k, 5) + SIJ(i, j, k + 1, 5) + SIJ(i + 1, j, k + 1, 5));
                SIJ(i, j, k, 6) = 0.25 * (SIJ(i, j, k, 6) + SIJ(i, j + 1, k, 6) + SIJ(i, j, k + 1, 6) + SIJ(i, j + 1, k + 1, 6));
            }
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) {
                double sm = 0.0;
                for (int c = 1; c <= 6; c++) {
                    double e = SIJ(i, j, k, c);
                    sm = sm + (c <= 3 ? 2.0 : 4.0) * e * e;
                }
                Smag[IC(i, j, k)] = sqrt(sm);
            }
}

static void orc_dsm(orc_t *o);

void orc_eddy_viscosity(orc_t *o) {
    const int nxg = o->nxg, nyg = o->nyg, nzg = o->nzg, ny = o->ny;
    double *nut = o->nu_t;
    if (o->p.les_model == 2) { orc_dsm(o); return; }
    if (o->p.les_model != 1) { /* subgrid.f90:41-45 */
        memset(nut, 0, o->nC * sizeof(double));
        memset(o->avg_nu_t, 0, o->nyg * sizeof(double));
        return;
    }
    const double *y = o->y, *yg = o->yg;
    const int s1 = nxg - 1, s2 = nyg - 1, s3 = nzg - 1;
    const size_t ns = (size_t)s1 * s2 * s3;
    double *S6 = dalloc(6 * ns), *Smag = dalloc(o->nC);
    compute_sij(o, o->U, o->V, o->W, S6, Smag);
    /* subgrid.f90:583-602 */
    double utau_t = sqrt(o->dPdx);
    for (int j = 2; j <= nyg - 1; j++) {
        double f;
        if (j < nyg / 2) f = 1.0 - exp(-(YG(j) - Yf(1)) / 25.0 / (o->p.nu / utau_t));
        else f = 1.0 - exp(-(Yf(ny) - YG(j)) / 25.0 / (o->p.nu / utau_t));
        if (o->p.dirichlet_nu_t == 0) f = 1.0;
        double Delta = pow(o->dx * o->dz * (Yf(j) - Yf(j - 1)), 1.0 / 3.0);
        double coef = sq(o->p.cs * Delta * f);
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) nut[IC(i, j, k)] = coef * Smag[IC(i, j, k)];
    }
    /* subgrid.f90:605-617 */
    for (int j = 2; j <= nyg - 1; j++) {
        double sum = 0.0;
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) sum += nut[IC(i, j, k)];
        o->avg_nu_t[j - 1] = fmax(sum / (double)((nxg - 2) * (nzg - 2)), 0.0);
    }
    o->avg_nu_t[0] = fmax(0.0, o->avg_nu_t[0]);
    if (o->p.dirichlet_nu_t == 1) { o->avg_nu_t[0] = -o->avg_nu_t[1]; o->avg_nu_t[nyg - 1] = -o->avg_nu_t[nyg - 2]; }
    else { o->avg_nu_t[0] = o->avg_nu_t[1]; o->avg_nu_t[nyg - 1] = o->avg_nu_t[nyg - 2]; }
    /* ghost sequence of subgrid.f90:521-536 */
    periodic_x(o, nut, nxg, nyg, nzg, 0);
    periodic_z(nut, nxg, nyg, nzg, 0);
    for (size_t q = 0; q < o->nC; q++) nut[q] = fmax(nut[q], 0.0);
    if (o->p.dirichlet_nu_t == 1) dirichlet_y(nut, nxg, nyg, nzg, 0);
    else neumann_y(nut, nxg, nyg, nzg);
    free(S6);
    free(Smag);
}

/* ======================================================================
 * Dynamic Smagorinsky, Lilly's form (subgrid.f90:376-538) with the
 * filters filter_xzy (subgrid.f90:134-192): Simpson weights in x and z
 * (1/36 1/9 1/36; 1/9 4/9 1/9) and (1/6 2/3 1/6) in y, one-sided
 * (2/3 1/3) on the first and last row of the filtered y range.  Arrays are
 * full (n1,n2,n3) with 1-based indices; the reference's sections map to
 * index ranges given explicitly.
 * ====================================================================== */
/* Tf(I,J,K) for I in [i0,i1], K in [k0,k1], J in [j0,j1] (the filtered y
   range: J = j0 and j1 are the one-sided rows), accumulated in the
   reference's order (x offset outer, z offset inner) */
static void filter_xzy_range(const double *T, int n1, int n2, double *Tf, int i0, int i1, int j0, int j1,
                             int k0, int k1) {
    static const double fil[3][3] = {{1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0},
                                     {1.0 / 9.0, 4.0 / 9.0, 1.0 / 9.0},
                                     {1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0}};
#define TT(i, j, k) T[I3(n1, n2, i, j, k)]
#pragma omp parallel for schedule(static)
    for (int fk = k0; fk <= k1; fk++)
        for (int fj = j0; fj <= j1; fj++)
            for (int fi = i0; fi <= i1; fi++) {
                double acc = 0.0;
                for (int a = -1; a <= 1; a++)
                    for (int c = -1; c <= 1; c++) {
                        double t;
                        if (fj == j0)
                            t = 2.0 / 3.0 * TT(fi + a, fj, fk + c) + 1.0 / 3.0 * TT(fi + a, fj + 1, fk + c);
                        else if (fj == j1)
                            t = 1.0 / 3.0 * TT(fi + a, fj - 1, fk + c) + 2.0 / 3.0 * TT(fi + a, fj, fk + c);
                        else
                            t = 1.0 / 6.0 * TT(fi + a, fj - 1, fk + c) + 2.0 / 3.0 * TT(fi + a, fj, fk + c) +
                                1.0 / 6.0 * TT(fi + a, fj + 1, fk + c);
                        acc = acc + t * fil[a + 1][c + 1];
                    }
                Tf[I3(n1, n2, fi, fj, fk)] = acc;
            }
#undef TT
}

static void orc_dsm(orc_t *o) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *U = o->U, *V = o->V, *W = o->W;
    const size_t nC = o->nC;
    const int s1 = nxg - 1, s2 = nyg - 1, s3 = nzg - 1;
    const size_t ns = (size_t)s1 * s2 * s3;
    double *uc = dalloc(nC), *vc = dalloc(nC), *wc = dalloc(nC); /* term, term_1, term_2 */
    double *tb = dalloc(6 * nC), *L = dalloc(6 * nC), *hv = dalloc(3 * nC), *M = dalloc(6 * nC);
    double *S6 = dalloc(6 * ns), *Smag = dalloc(nC), *tmp = dalloc(nC), *tf = dalloc(nC);
    /* Part 1: velocities at the cell centres (interpolate_x/y/z, faces to
       centres), periodic in x (term) and z (term_2) */
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++)
            for (int i = 2; i <= nxg - 1; i++) uc[IC(i, j, k)] = 0.5 * (u(i - 1, j, k) + u(i, j, k));
    for (int k = 1; k <= nzg; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 1; i <= nxg; i++) vc[IC(i, j, k)] = 0.5 * (v(i, j - 1, k) + v(i, j, k));
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 1; j <= nyg; j++)
            for (int i = 1; i <= nxg; i++) wc[IC(i, j, k)] = 0.5 * (w(i, j, k - 1) + w(i, j, k));
    periodic_x(o, uc, nxg, nyg, nzg, 0);
    periodic_z(wc, nxg, nyg, nzg, 0);
    /* products u_i u_j at the centres (2:nxg, 2:nyg-1, 2:nzg), periodic */
    for (int c = 0; c < 6; c++) {
        double *T = tb + (size_t)c * nC;
        const double *A = c == 0 || c == 3 || c == 4 ? uc : (c == 1 || c == 5 ? vc : wc);
        const double *B = c == 0 ? uc : (c == 1 || c == 3 ? vc : wc);
        for (int k = 2; k <= nzg; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg; i++) T[IC(i, j, k)] = A[IC(i, j, k)] * B[IC(i, j, k)];
        periodic_x(o, T, nxg, nyg, nzg, 0);
        periodic_z(T, nxg, nyg, nzg, 0);
        /* hat(ui uj) */
        filter_xzy_range(T, nxg, nyg, L + (size_t)c * nC, 2, nxg - 1, 2, nyg - 1, 2, nzg - 1);
    }
    /* hat(ui) */
    filter_xzy_range(uc, nxg, nyg, hv, 2, nxg - 1, 2, nyg - 1, 2, nzg - 1);
    filter_xzy_range(vc, nxg, nyg, hv + nC, 2, nxg - 1, 2, nyg - 1, 2, nzg - 1);
    filter_xzy_range(wc, nxg, nyg, hv + 2 * nC, 2, nxg - 1, 2, nyg - 1, 2, nzg - 1);
    for (int c = 0; c < 6; c++) {
        const double *A = hv + (size_t)(c == 0 || c == 3 || c == 4 ? 0 : (c == 1 || c == 5 ? 1 : 2)) * nC;
        const double *B = hv + (size_t)(c == 0 ? 0 : (c == 1 || c == 3 ? 1 : 2)) * nC;
        double *Lc = L + (size_t)c * nC;
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) Lc[IC(i, j, k)] = Lc[IC(i, j, k)] - A[IC(i, j, k)] * B[IC(i, j, k)];
    }
    /* Part 2: filtered velocities with their periodic ghosts */
    double *Uf = dalloc(o->nU), *Vf = dalloc(o->nV), *Wf = dalloc(o->nW);
    filter_xzy_range(U, nx, nyg, Uf, 2, nx - 1, 1, nyg, 2, nzg - 1);
    filter_xzy_range(V, nxg, ny, Vf, 2, nxg - 1, 1, ny, 2, nzg - 1);
    filter_xzy_range(W, nxg, nyg, Wf, 2, nxg - 1, 1, nyg, 2, nz - 1);
    periodic_x(o, Uf, nx, nyg, nzg, 1);
    periodic_z(Uf, nx, nyg, nzg, 0);
    periodic_x(o, Vf, nxg, ny, nzg, 0);
    periodic_z(Vf, nxg, ny, nzg, 0);
    periodic_x(o, Wf, nxg, nyg, nz, 0);
    periodic_z(Wf, nxg, nyg, nz, 1);
    /* M = fil_size^2 |hat S| hat(Sij) - hat(|S| Sij), fil_size = 2 */
    compute_sij(o, Uf, Vf, Wf, S6, Smag);
#define SIJ(i, j, k, c) S6[(size_t)(c - 1) * ns + I3(s1, s2, (i)-1, (j)-1, (k)-1)]
    const double fs2 = pow(2.0, 2.0);
    for (int c = 1; c <= 6; c++)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++)
                    M[(size_t)(c - 1) * nC + IC(i, j, k)] = fs2 * Smag[IC(i, j, k)] * SIJ(i, j, k, c);
    compute_sij(o, U, V, W, S6, Smag);
    for (int c = 1; c <= 6; c++) {
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) tmp[IC(i, j, k)] = Smag[IC(i, j, k)] * SIJ(i, j, k, c);
        periodic_x(o, tmp, nxg, nyg, nzg, 0);
        periodic_z(tmp, nxg, nyg, nzg, 0);
        filter_xzy_range(tmp, nxg, nyg, tf, 2, nxg - 1, 2, nyg - 1, 2, nzg - 1);
        double *Mc = M + (size_t)(c - 1) * nC;
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) Mc[IC(i, j, k)] = Mc[IC(i, j, k)] - tf[IC(i, j, k)];
    }
#undef SIJ
    /* Part 3: plane averages <Lij Mij> / <Mij Mij>, nu_t = -0.5 avg |S| */
    double *nut = o->nu_t, *avg = o->avg_nu_t;
    memset(nut, 0, nC * sizeof(double));
#define LL(c) L[(size_t)((c)-1) * nC + q]
#define MM(c) M[(size_t)((c)-1) * nC + q]
    for (int j = 2; j <= nyg - 1; j++) {
        double lm = 0.0, mm = 0.0;
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) {
                const size_t q = IC(i, j, k);
                lm += LL(1) * MM(1) + LL(2) * MM(2) + LL(3) * MM(3) + 2.0 * (LL(4) * MM(4) + LL(5) * MM(5) + LL(6) * MM(6));
            }
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) {
                const size_t q = IC(i, j, k);
                mm += MM(1) * MM(1) + MM(2) * MM(2) + MM(3) * MM(3) + 2.0 * (MM(4) * MM(4) + MM(5) * MM(5) + MM(6) * MM(6));
            }
        avg[j - 1] = lm / mm;
    }
#undef LL
#undef MM
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) nut[IC(i, j, k)] = -0.5 * avg[j - 1] * Smag[IC(i, j, k)];
    periodic_x(o, nut, nxg, nyg, nzg, 0);
    periodic_z(nut, nxg, nyg, nzg, 0);
    for (size_t q = 0; q < nC; q++) nut[q] = fmax(nut[q], 0.0);
    if (o->p.dirichlet_nu_t == 1) {
        dirichlet_y(nut, nxg, nyg, nzg, 0);
        avg[0] = -avg[1];
        avg[nyg - 1] = -avg[nyg - 2];
    } else {
        neumann_y(nut, nxg, nyg, nzg);
        avg[0] = avg[1];
        avg[nyg - 1] = avg[nyg - 2];
    }
    free(uc); free(vc); free(wc); free(tb); free(L); free(hv); free(M);
    free(S6); free(Smag); free(tmp); free(tf); free(Uf); free(Vf); free(Wf);
}

/* ======================================================================
 * eqwm.f90:20-223
 * ====================================================================== */
void orc_eqwm(double umag_les, double mu_lam, double rho_eq, double h_wm, int ncv,
              double stretch, double tol, int max_iter, double tau_in, double *tau_out,
              int *iter_out) {
    int nfa = ncv + 1;
    double *y_cv = dalloc(ncv + 1), *y_fa = dalloc(nfa), *mu_fa = dalloc(nfa);
    double *aa = dalloc(ncv), *bb = dalloc(ncv), *cc = dalloc(ncv), *b = dalloc(ncv);
    double *u_wm = dalloc(ncv), *cp = dalloc(ncv), *dp = dalloc(ncv);
    const double kappa = 0.41, Aplus = 17.0;
    double tau_wall_lam = mu_lam * umag_les / h_wm;
    /* grid (:20-50); the maximum is taken over the ncv set centroids
       (the reference's y_cv(ncv+1) is never assigned) */
    y_fa[0] = 0.0;
    double inc = 1.0;
    for (int ifa = 2; ifa <= nfa; ifa++) {
        y_fa[ifa - 1] = y_fa[ifa - 2] + inc;
        y_cv[ifa - 2] = 0.5 * (y_fa[ifa - 1] + y_fa[ifa - 2]);
        inc = inc * stretch;
    }
    double ymax = y_cv[0];
    for (int q = 1; q < ncv; q++) if (y_cv[q] > ymax) ymax = y_cv[q];
    for (int ifa = 0; ifa < nfa; ifa++) y_fa[ifa] = y_fa[ifa] / ymax * h_wm;
    for (int icv = 0; icv < ncv; icv++) y_cv[icv] = y_cv[icv] / ymax * h_wm;

    int iter = 0, done = 0;
    double tau_wall = tau_in;
    while (!done) {
        double tau_prev = tau_wall;
        for (int q = 0; q < nfa; q++) mu_fa[q] = mu_lam * 1.0;
        for (int ifa = 2; ifa <= nfa; ifa++) {
            double nu_eq = mu_lam / rho_eq;
            double utau_eq = sqrt(tau_wall / rho_eq);
            double D = 1.0 - exp(-1.0 * y_fa[ifa - 1] * utau_eq / nu_eq / Aplus);
            D = D * D;
            double mut = rho_eq * kappa * y_fa[ifa - 1] * utau_eq * D;
            mu_fa[ifa - 1] = mu_fa[ifa - 1] + mut;
        }
        /* :152-193 tridiagonal assembly */
        bb[ncv - 1] = 1.0; b[ncv - 1] = umag_les; aa[ncv - 1] = 0.0; cc[ncv - 1] = 0.0;
        for (int icv = 2; icv <= ncv - 1; icv++) {
            double sup = mu_fa[icv] / (y_cv[icv] - y_cv[icv - 1]);
            double sub = mu_fa[icv - 1] / (y_cv[icv - 1] - y_cv[icv - 2]);
            bb[icv - 1] = -1.0 * (sup + sub);
            cc[icv - 1] = sup;
            aa[icv - 1] = sub;
            b[icv - 1] = 0.0;
        }
        b[0] = 0.0;
        double sup = mu_fa[1] / (y_cv[1] - y_cv[0]);
        bb[0] = -1.0 * (sup + mu_fa[0] / (y_cv[0] - y_fa[0]));
        cc[0] = sup;
        aa[0] = 0.0;
        /* thomas_algorithm (:52-88) */
        cp[0] = cc[0] / bb[0];
        dp[0] = b[0] / bb[0];
        for (int q = 1; q < ncv; q++) {
            double m = bb[q] - cp[q - 1] * aa[q];
            cp[q] = cc[q] / m;
            dp[q] = (b[q] - dp[q - 1] * aa[q]) / m;
        }
        u_wm[ncv - 1] = dp[ncv - 1];
        for (int q = ncv - 2; q >= 0; q--) u_wm[q] = dp[q] - cp[q] * u_wm[q + 1];
        tau_wall = mu_fa[0] * (u_wm[0] - 0.0) / (y_cv[0] - y_fa[0]);
        if (fabs((tau_wall - tau_prev) / tau_wall_lam) < tol) done = 1;
        if (!done) iter = iter + 1;
        if (iter > max_iter) done = 1;
    }
    *tau_out = tau_wall;
    *iter_out = iter;
    free(y_cv); free(y_fa); free(mu_fa); free(aa); free(bb); free(cc); free(b);
    free(u_wm); free(cp); free(dp);
}

/* wallmodel.f90:591-685 plus compute_pseudo_pressure_bc_for_robin_bc
   (:1342-1378), which leaves Dyy unchanged because alpha_y = 0 */
void orc_wall_model(orc_t *o, int *n_nan, int *n_iter) {
    if (o->p.iwall_model != 3) { if (n_nan) *n_nan = 0; if (n_iter) *n_iter = 0; return; }
    const int nx = o->nx, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg, nz = o->nz;
    const double *U = o->U, *W = o->W, *nut = o->nu_t;
    const int jref = o->p.jref_eqwm, jtop = nyg - jref + 1;
    double *t1 = dalloc((size_t)nxg * nzg), *t2 = dalloc((size_t)nx * nzg), *t4 = dalloc((size_t)nx * nzg);
    for (int wall = 0; wall < 2; wall++) {
        int jw = wall == 0 ? jref : jtop;
        double *dst = wall == 0 ? t2 : t4;
        /* interpolate_z_2nd(z, W(:,jw,:), zg, temp_1) */
        for (int k = 1; k <= nzg; k++) {
            int jj = o->iz_jj[k - 1];
            const double *wz = o->iz_w + 3 * (k - 1);
            for (int i = 1; i <= nxg; i++)
                t1[(size_t)(k - 1) * nxg + (i - 1)] = wz[0] * w(i, jw, jj - 1) + wz[1] * w(i, jw, jj) + wz[2] * w(i, jw, jj + 1);
        }
        /* interpolate_x_2nd(xg, temp_1, x, temp_2) */
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nx; i++) {
                int jj = o->ix_jj[i - 1];
                const double *wx = o->ix_w + 3 * (i - 1);
                const double *row = t1 + (size_t)(k - 1) * nxg;
                dst[(size_t)(k - 1) * nx + (i - 1)] = wx[0] * row[jj - 2] + wx[1] * row[jj - 1] + wx[2] * row[jj];
            }
    }
    int nn = 0, ni = 0;
    const double nu = o->p.nu, h = o->yg[jref - 1];
    for (int i = 1; i <= nx; i++)
        for (int k = 1; k <= nzg; k++)
            for (int wall = 0; wall < 2; wall++) {
                int jw = wall == 0 ? jref : jtop;
                double wt = (wall == 0 ? t2 : t4)[(size_t)(k - 1) * nx + (i - 1)];
                double ul = sqrt(sq(u(i, jw, k)) + sq(wt));
                double tau; int it;
                orc_eqwm(ul, nu, 1.0, h, o->p.eqwm_ncv, o->p.eqwm_stretch, o->p.eqwm_tol,
                         o->p.eqwm_max_iter, o->tau_init[AX(i, wall + 1, k)], &tau, &it);
                if (isnan(tau)) { nn++; tau = o->p.tau_mean_eqwm; }
                if (it == o->p.eqwm_max_iter) { tau = o->p.tau_mean_eqwm; ni++; }
                o->tau_init[AX(i, wall + 1, k)] = tau;
                double tx = tau * u(i, jw, k) / ul;
                double tz = tau * wt / ul;
                double nik;
                if (wall == 0)
                    nik = 0.5 * (0.5 * (nut[IC(i, 1, k)] + nut[IC(i + 1, 1, k)]) + 0.5 * (nut[IC(i, 2, k)] + nut[IC(i + 1, 2, k)]));
                else
                    nik = 0.5 * (0.5 * (nut[IC(i, nyg, k)] + nut[IC(i + 1, nyg, k)]) + 0.5 * (nut[IC(i, nyg - 1, k)] + nut[IC(i + 1, nyg - 1, k)]));
                o->alpha_x[AX(i, wall + 1, k)] = tx / (nu + nik);
                o->alpha_z_temp[AX(i, wall + 1, k)] = tz / (nu + nik);
            }
    /* :667-674 */
    double *az = o->alpha_z, *azt = o->alpha_z_temp;
    for (int i = 1; i <= nxg - 2; i++)
        for (int k = 1; k <= nz - 1; k++)
            for (int wl = 1; wl <= 2; wl++)
                az[AZ(i, wl, k)] = 0.25 * (azt[AX(i, wl, k)] + azt[AX(i + 1, wl, k)] + azt[AX(i, wl, k + 1)] + azt[AX(i + 1, wl, k + 1)]);
    for (int k = 1; k <= nz; k++)
        for (int wl = 1; wl <= 2; wl++) {
            az[AZ(nxg, wl, k)] = az[AZ(3, wl, k)];
            az[AZ(nxg - 1, wl, k)] = az[AZ(2, wl, k)];
        }
    for (int i = 1; i <= nxg; i++)
        for (int wl = 1; wl <= 2; wl++) az[AZ(i, wl, nz)] = az[AZ(i, wl, 2)];
    free(t1); free(t2); free(t4);
    if (n_nan) *n_nan = nn;
    if (n_iter) *n_iter = ni;
}

/* ======================================================================
 * time_integration.f90:89-173
 * ====================================================================== */
void orc_rk3_step(orc_t *o) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    double to = o->t;
    memcpy(o->Uo, o->U, o->nU * sizeof(double));
    memcpy(o->Vo, o->V, o->nV * sizeof(double));
    memcpy(o->Wo, o->W, o->nW * sizeof(double));
    double *U = o->U, *V = o->V, *W = o->W;
    for (int s = 1; s <= 3; s++) {
        orc_eddy_viscosity(o);
        if (s == 1) {
            orc_wall_model(o, NULL, NULL);
            orc_eddy_viscosity(o);
        }
        orc_compute_rhs(o, U, V, W, o->Fu[s - 1], o->Fv[s - 1], o->Fw[s - 1]);
        const double dt = o->dt;
        double a1 = o->rk_coef[s - 1][0], a2 = o->rk_coef[s - 1][1], a3 = o->rk_coef[s - 1][2];
#define RKC(F, q) (s == 1 ? a1 * F[0][q] : (s == 2 ? (a1 * F[0][q] + a2 * F[1][q]) : (a1 * F[0][q] + a2 * F[1][q] + a3 * F[2][q])))
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nx - 1; i++) {
                    size_t q = IRU(i, j, k);
                    U[IU(i, j, k)] = o->Uo[IU(i, j, k)] + (s == 1 ? dt * a1 * o->Fu[0][q] : dt * RKC(o->Fu, q));
                }
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    size_t q = IRV(i, j, k);
                    V[IV(i, j, k)] = o->Vo[IV(i, j, k)] + (s == 1 ? dt * a1 * o->Fv[0][q] : dt * RKC(o->Fv, q));
                }
#pragma omp parallel for schedule(static)
        for (int k = 2; k <= nz - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) {
                    size_t q = IRW(i, j, k);
                    W[IW(i, j, k)] = o->Wo[IW(i, j, k)] + (s == 1 ? dt * a1 * o->Fw[0][q] : dt * RKC(o->Fw, q));
                }
#undef RKC
        o->t = to + o->rk_t[s - 1] * dt;
        orc_apply_boundary_conditions(o);
        orc_projection_step(o, s);
        orc_apply_boundary_conditions(o);
    }
    if (o->p.x_mass_cte == 1) { /* :153-157 */
        double d = o->Qflow_x_0 - orc_mean_mass_flow_U(o, U);
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nx - 1; i++) U[IU(i, j, k)] = U[IU(i, j, k)] + d;
        o->dPdx = d / o->dt;
        orc_apply_boundary_conditions(o);
    } else if (o->p.x_mass_cte == 2) { /* :158-163 */
        double d = o->p.u_centerline - orc_mean_centerline_U(o, U);
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nx - 1; i++) U[IU(i, j, k)] = U[IU(i, j, k)] + d;
        o->dPdx = d / o->dt;
        orc_apply_boundary_conditions(o);
    }
    if (o->p.y_mass_cte == 1) { /* :166-171 */
        double d = PLANT(18, orc_mean_mass_flow_V(o, V) - o->Qflow_y_0, o->Qflow_y_0 - orc_mean_mass_flow_V(o, V));
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= ny - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) V[IV(i, j, k)] = V[IV(i, j, k)] + d;
        o->dPdy = d / o->dt;
        orc_apply_boundary_conditions(o);
    }
}

/* ======================================================================
 * statistics.f90:43-155 (the instantaneous plane statistics)
 * ====================================================================== */
void orc_statistics(const orc_t *o, double *out) {
    const int nx = o->nx, ny = o->ny, nz = o->nz, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *U = o->U, *V = o->V, *W = o->W, *w0 = o->w0, *w1 = o->w1;
    double *Um = out, *Vm = out + ny, *Wm = out + 2 * ny, *U2 = out + 3 * ny, *V2 = out + 4 * ny,
           *W2 = out + 5 * ny, *UV = out + 6 * ny;
    for (int jj = 1; jj <= ny; jj++) {
        double su = 0, su2 = 0, sw = 0, sw2 = 0, sv = 0, sv2 = 0, suv = 0;
        /* term_1 = interpolate_y(U,2): U at y faces; sum over (1:nx-2, 1:nzg-2) */
        for (int k = PLANT(10, 2, 1); k <= nzg - 2; k++)
            for (int i = 1; i <= nx - 2; i++) {
                double t = W0(jj) * u(i, jj, k) + W1(jj) * u(i, jj + 1, k);
                su += t; su2 += sq(t);
            }
        for (int k = 1; k <= nz - 2; k++)
            for (int i = 1; i <= nxg - 2; i++) {
                double t = W0(jj) * w(i, jj, k) + W1(jj) * w(i, jj + 1, k);
                sw += t; sw2 += PLANT(11, t, sq(t));
            }
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= PLANT(15, nxg - 2, nxg - 1); i++) { sv += v(i, jj, k); sv2 += sq(v(i, jj, k)); }
        /* UVmean: V(2:nxg-1) * term(1:nxg-2), term = interpolate_x(term_1) */
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) {
                int ii = PLANT(13, i, i - 1);
                double ta = W0(jj) * u(ii, jj, k) + W1(jj) * u(ii, jj + 1, k);
                double tb = ii + 1 <= nx ? W0(jj) * u(ii + 1, jj, k) + W1(jj) * u(ii + 1, jj + 1, k) : 0.0;
                suv += v(i, jj, k) * (0.5 * (ta + tb));
            }
        Um[jj - 1] = su / (double)((nx - 2) * (nzg - 2));
        Vm[jj - 1] = sv / (double)((nxg - 2) * (nzg - 2));
        Wm[jj - 1] = sw / (double)((nxg - 2) * (nz - 2));
        U2[jj - 1] = su2 / (double)((nx - 2) * (nzg - 2));
        V2[jj - 1] = sv2 / (double)((nxg - 2) * (nzg - 2));
        W2[jj - 1] = sw2 / (double)((nxg - 2) * (nz - 2));
        UV[jj - 1] = suv / (double)((nxg - 2) * (nzg - 2));
    }
    double U1 = 0, U2s = 0, Un1 = 0, Un2 = 0;
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nx - 1; i++) {
            U1 += u(i, 1, k); U2s += u(i, 2, k); Un1 += u(i, nyg - 1, k); Un2 += u(i, nyg, k);
        }
    double den = (double)((nx - 2) * (nzg - 2));
    U1 /= den; U2s /= den; Un1 /= den; Un2 /= den;
    const double *yg = o->yg, *y = o->y, *nut = o->nu_t;
    double dUb = (U2s - U1) / (YG(2) - YG(1));
    double dUt = (Un2 - Un1) / (YG(nyg) - YG(nyg - 1));
    double nb = 0, ntp = 0;
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 1; j <= 2; j++)
            for (int i = 2; i <= nxg - 1; i++) nb += nut[IC(i, j, k)];
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = nyg - 1; j <= nyg; j++)
            for (int i = 2; i <= nxg - 1; i++) ntp += nut[IC(i, j, k)];
    nb = nb / (double)((nxg - 2) * (nzg - 2)) / 2.0;
    ntp = ntp / (double)((nxg - 2) * (nzg - 2)) / 2.0;
    double Ly = Yf(ny) - Yf(1);
    double utau = sqrt((PLANT(12, UV[0] - UV[ny - 1], UV[ny - 1] - UV[0]) - (o->p.nu + ntp) * dUt + (o->p.nu + nb) * dUb) / Ly);
    double *tail = out + 7 * ny;
    tail[0] = utau;
    tail[1] = utau * (Yf(ny) - Yf(1)) / 2.0 / o->p.nu;
    tail[2] = orc_mean_mass_flow_U(o, U);
    tail[3] = orc_mean_mass_flow_V(o, V);
    (void)nz;
}
