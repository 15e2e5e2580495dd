/*
 * channel_flow.h -- C-ABI of the MI355X (gfx950) time-step library for the
 * channel-flow solver.
 *
 * Every entry point takes plain pointers and sizes, returns 0 on success and
 * a nonzero code on failure; cf_last_error() then returns the message.  The
 * Fortran bindings (channel-flow_amd/fortran/) turn a nonzero code into
 * `Stop 'message'`, which is the reference's error convention.
 *
 * Arrays exchanged with the host use the reference layout exactly
 * (global.f90:82-84; Fortran order, x fastest, ghosts included):
 *   U (nx ,nyg,nzg)   V (nxg,ny ,nzg)   W (nxg,nyg,nz )   P, nu_t (nxg,nyg,nzg)
 *   rhs_u (2:nx-1,2:nyg-1,2:nzg-1)  rhs_v (2:nxg-1,2:ny-1,2:nzg-1)
 *   rhs_w (2:nxg-1,2:nyg-1,2:nz-1)
 *   alpha_x (nx,2,nzg)  alpha_z (nxg,2,nz)  tau_eqwm_initial (nx,2,nzg)
 * with nx,ny,nz the face counts nx_global.. of the input file and
 * nxg = nx+1, nyg = ny+1, nzg = nz+1 (one rank owns the whole domain).
 *
 * Scalars vector (7 doubles): [t, dt, dPdx, dPdz, Qflow_x_0, dPdy, Qflow_y_0].
 */
#ifndef CHANNEL_FLOW_H
#define CHANNEL_FLOW_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cf_ctx cf_ctx;

/* Run configuration: the fields of read_input_parameters
 * (input_output.f90:21-81) that reach the time step, plus the constants the
 * reference hard-codes (Smagorinsky Cs subgrid.f90:580; EQWM ncv, stretch
 * wallmodel.f90:621, tolerance and iteration cap eqwm.f90:109-110). */
typedef struct cf_params {
    int nx, ny, nz;          /* nx_global, ny_global, nz_global (faces)          */
    int precision;           /* 0 = FP64, 1 = mixed (FP32 stencils/FFT storage) */
    double cfl;              /* CFL; < 0 means fixed dt = -CFL                   */
    double nu;
    double dPdx, dPdz;       /* initial forcing                                   */
    double omega_z;          /* Rossby_plus*utau/2 (input_output.f90:84)         */
    int x_mass_cte;          /* 0 constant dPdx, 1 constant bulk flow
                                (mass_flow.f90:103-132), 2 constant centreline
                                velocity (mass_flow.f90:192-211)                 */
    int y_mass_cte;          /* 1 constant bulk flow in y (mass_flow.f90:147-176) */
    int les_model;           /* 0 none, 1 constant-coefficient Smagorinsky,
                                2 dynamic Smagorinsky (subgrid.f90:376-538)     */
    int iwall_model;         /* 0 no-slip, 3 equilibrium wall model               */
    int dirichlet_nu_t;
    int jref_eqwm;
    double cs;               /* Smagorinsky coefficient                           */
    int eqwm_ncv;            /* wall-model ODE cells (reference: 100)             */
    int eqwm_max_iter;       /* Picard iteration cap (reference: 100)            */
    double eqwm_stretch;     /* wall-model grid stretching (reference: 1.10)     */
    double eqwm_tol;         /* Picard tolerance (reference: 1e-7)               */
    double tau_mean_eqwm;    /* fallback stress = dPdx_ref                        */
    int wall_surrogate;      /* 0 ODE only, 1 learned MLP with ODE fallback       */
    int device;              /* HIP device ordinal                                */
    double u_centerline;     /* x_mass_cte = 2: target centreline velocity
                                Ucenterline_x (initialization.f90:453-455)       */
} cf_params;

const char *cf_last_error(void);

/* initialize (initialization.f90:23-536): grids, operators, device buffers.
 * x (nx), y (ny), z (nz) are the face coordinates of the input file. */
int cf_create(const cf_params *p, const double *x, const double *y, const double *z,
              cf_ctx **out);
int cf_destroy(cf_ctx *c);

/* host mirrors <-> device state */
int cf_set_state(cf_ctx *c, const double *U, const double *V, const double *W);
int cf_get_state(cf_ctx *c, double *U, double *V, double *W);
int cf_get_pressure(cf_ctx *c, double *P);      /* projection.f90:129-137 */
/* compute_pressure (pressure.f90:21-341): the actual pressure of the current
   state, written to P (nxg,nyg,nzg) on the host.  Uses the current nu_t for
   the plane-averaged eddy viscosity of the wall conditions. */
int cf_compute_pressure(cf_ctx *c, double *P);
int cf_set_nu_t(cf_ctx *c, const double *nu_t);
int cf_get_nu_t(cf_ctx *c, double *nu_t);
/* avg_nu_t(1,nyg,1) of global.f90:180 after the last eddy-viscosity
   evaluation: the dynamic model's <LM>/<MM> per row (subgrid.f90:505-535) or
   the clipped plane mean of nu_t (subgrid.f90:605-617); wall rows by the nu_t
   ghost rule; zeros without a model.  nyg values. */
int cf_get_avg_nu_t(cf_ctx *c, double *avg);
int cf_get_wall(cf_ctx *c, double *alpha_x, double *alpha_z, double *tau_init);
int cf_set_tau_init(cf_ctx *c, const double *tau_init);
int cf_set_scalars(cf_ctx *c, const double *s);
int cf_get_scalars(cf_ctx *c, double *s);

/* initialization.f90:460: Qflow_x_0 from the current U */
int cf_init_mass_flow(cf_ctx *c);

/* apply_boundary_conditions (boundary_conditions.f90:20-75) */
int cf_apply_boundary_conditions(cf_ctx *c);
/* compute_dt (time_integration.f90:179-227) */
int cf_compute_dt(cf_ctx *c, double *dt);
/* compute_time_step_RK3 (time_integration.f90:89-173) */
int cf_rk3_step(cf_ctx *c);
/* main.f90:67-74 for nsteps: compute_dt + compute_time_step_RK3, all on the
 * device with no host synchronisation inside (HIP graph when use_graph). */
int cf_run_steps(cf_ctx *c, int nsteps, int use_graph);
/* compute_projection_step (projection.f90:20-28); rk_step selects whether
 * the pressure is saved (rk_step == 1) */
int cf_projection_step(cf_ctx *c, int rk_step);
/* compute_rhs_u/v/w (equations.f90:24,186,351) of host fields, with the
 * context's nu_t, dPdx, dPdz */
int cf_compute_rhs(cf_ctx *c, const double *U, const double *V, const double *W, double *rhs_u,
                   double *rhs_v, double *rhs_w);
/* compute_eddy_viscosity (subgrid.f90:23-47) of the current state */
int cf_compute_eddy_viscosity(cf_ctx *c);
/* compute_wall_model (wallmodel.f90:24-82) of the current state */
int cf_compute_wall_model(cf_ctx *c, int *n_nan, int *n_iter);
/* compute_mean_mass_flow_U (mass_flow.f90:25-52) of the current state */
int cf_mean_mass_flow_u(cf_ctx *c, double *qflow);
/* compute_mean_mass_flow_V (mass_flow.f90:61-88) of the current state */
int cf_mean_mass_flow_v(cf_ctx *c, double *qflow);
/* the centreline mean of compute_dPx_for_constant_Ucenterline
 * (mass_flow.f90:200-207): U(1:nx-2, ny/2+1, 1:nzg-2) over (nx-2)(nzg-2) */
int cf_mean_centerline_u(cf_ctx *c, double *umean);
/* check_divergence (projection.f90:234-256) plus max |div| */
int cf_check_divergence(cf_ctx *c, double *max_signed, double *max_abs);
/* compute_statistics (statistics.f90:22-171), FP64, 7*ny + 4 values:
 * Umean,Vmean,Wmean,U2mean,V2mean,W2mean,UVmean (ny each), then utau, Retau,
 * Qflow_x, Qflow_y.  Returns nonzero ("Error: NaNs!") when U, V or W holds a
 * NaN (statistics.f90:165-167); the values are written either way. */
int cf_statistics(cf_ctx *c, double *out);

/* solve_u_eqwm (eqwm.f90:90-97) for a batch of n wall points on the device.
 * Host arrays; rho = 1.  Returns tau and the iteration count per point. */
int cf_eqwm_batch(int n, const double *umag, const double *nu, const double *h,
                  const double *tau_in, int ncv, double stretch, double tol, int max_iter,
                  double *tau, int *iter);

/* cf_eqwm_batch without any device allocation or release.  No counterpart in
 * the reference (which calls solve_u_eqwm point by point): it is for callers
 * that cannot allocate device memory, such as serving loops and harnesses
 * that forbid hipMalloc.  The staging of the n points lives in a fixed
 * device array of CF_EQWM_NOALLOC_MAX points per GPU, held until the process
 * ends; larger n is rejected with an error.  Runs on the current device.
 * Calls that overlap on one GPU share that array and race. */
#define CF_EQWM_NOALLOC_MAX 524288
int cf_eqwm_batch_noalloc(int n, const double *umag, const double *nu, const double *h,
                          const double *tau_in, int ncv, double stretch, double tol, int max_iter,
                          double *tau, int *iter);

/* ---- learned wall model ----
 * A 4 -> 64 -> 64 -> 64 -> 1 tanh MLP that maps the wall-model inputs
 * (u_par, h, nu, dP/dx) to the wall stress tau_w.  It stands in for the
 * per-point ODE solve solve_u_eqwm (eqwm.f90:90-97, called from
 * wallmodel.f90:629-650) and has no counterpart in the reference itself:
 * the reference always solves the ODE.  The two 64x64 layers run on BF16
 * matrix cores with FP32 accumulation; master weights are FP32.  Inputs
 * outside the box of the training features fall back to the ODE solver. */
typedef struct cf_mlp cf_mlp;
int cf_mlp_create(unsigned seed, cf_mlp **out);
int cf_mlp_destroy(cf_mlp *m);
/* number of FP32 parameters: W1[4][64] b1[64] W2[64][64] b2 W3[64][64] b3 W4[64] b4
 * (weights stored [input][output]) */
int cf_mlp_nparams(void);
int cf_mlp_get_params(cf_mlp *m, float *w);
int cf_mlp_set_params(cf_mlp *m, const float *w);
/* nrm[16]: feature means [0..3], feature std [4..7], ln tau mean [8], std [9];
 * hull[8]: per-feature min [0..3] and max [4..7] in standardised units */
int cf_mlp_get_norm(cf_mlp *m, float *nrm, float *hull);
int cf_mlp_set_norm(cf_mlp *m, const float *nrm, const float *hull);
/* Fit the normalisation and hull to the data, then `epochs` passes of Adam
 * over mini-batches of `batch` samples (mean squared error of standardised
 * ln tau), learning rate decaying exponentially from lr to lr_final.
 * feat: n x 4 raw (u_par, h, nu, dP/dx); tau: n.  loss_history: epochs
 * values (loss of the last batch of each epoch), or null. */
int cf_mlp_train(cf_mlp *m, long long n, const double *feat, const double *tau, int epochs, int batch,
                 float lr, float lr_final, float *loss_history);
/* gradient of the batch loss with respect to every parameter (no update) */
int cf_mlp_gradients(cf_mlp *m, int n, const double *feat, const double *tau, float *grad, float *loss);
/* tau for n raw feature rows; inside[q] = 1 when row q lies in the hull (may be null) */
int cf_mlp_predict(cf_mlp *m, long long n, const double *feat, double *tau, int *inside);
/* Use the model inside cf_compute_wall_model / the RK3 step of a context
 * (iwall_model = 3); null detaches it.  The model must outlive the use. */
int cf_set_wall_surrogate(cf_ctx *c, cf_mlp *m);
/* wall points that fell back to the ODE solver in the last wall-model call */
int cf_wall_fallbacks(cf_ctx *c);

/* Time-window statistics.  The reference writes instantaneous plane means
 * every nstats steps (statistics.f90:22-171); these add the same seven
 * plane sums (U, V, W, their squares, UV at the y faces) into an FP64 device
 * accumulator without a host round trip, and return their average over the
 * accumulated samples (7*ny values, as cf_statistics) and the sample count.
 * Reynolds stresses follow as <UU> - <U><U>, <UV> - <U><V>. */
int cf_stats_accumulate(cf_ctx *c);
int cf_stats_window(cf_ctx *c, double *out, double *count);
int cf_stats_reset(cf_ctx *c);

/* Device bytes that a context of these parameters needs (exact; the
 * serving arena admits cases by this figure).  No device is touched. */
long long cf_footprint(const cf_params *p, const double *x, const double *y, const double *z);

/* ---- batches of one size class ----
 * n cases that share every option but nu, dPdx, dPdz, u_centerline and
 * tau_mean_eqwm (same grid x, y, z), with their device buffers in one
 * allocation at a fixed stride, stepped together: each kernel of the RK3
 * step is launched once for all n cases (the case index is the launch
 * grid's second dimension), so small grids fill the GPU.  The reference has
 * no counterpart (one case is one run of main.f90); each member computes
 * the same bits as the same case alone.  The batched step takes both
 * precisions, the Smagorinsky model and the equilibrium wall model; not the
 * dynamic Smagorinsky model, y mass flow or the learned wall model.
 * cf_batch_member gives member m as an ordinary context for everything but
 * the step (set_state, get_state, statistics, ...); it is owned by the batch
 * and is stepped only by cf_batch_run_steps. */
typedef struct cf_batch cf_batch;
int cf_batch_create(int n, const cf_params *p, const double *x, const double *y, const double *z,
                    cf_batch **out);
int cf_batch_member(cf_batch *b, int m, cf_ctx **ctx);
int cf_batch_run_steps(cf_batch *b, int nsteps, int use_graph);
int cf_batch_destroy(cf_batch *b);

/* ---- serving ----
 * Many cases of different grids, Reynolds numbers and models on one GPU
 * with continuous batching.  The reference has no counterpart: there one
 * case is one run of main.f90 (main.f90:61-87).
 * hbm_budget: bytes reserved once (one allocation) for all running cases;
 * a case is admitted, in submission order once its arrival time has passed,
 * when a free slot and a contiguous free region of cf_footprint bytes exist.
 * max_active: concurrent cases (one HIP stream each); depth: RK3 steps kept
 * in flight per case. */
typedef struct cf_server cf_server;
int cf_server_create(int device, long long hbm_budget, int max_active, int depth, cf_server **out);
/* cancels what is left, stops the scheduler and frees the arena */
int cf_server_destroy(cf_server *s);
/* queue one case: raw initial field (projected at admission, as
 * initialization.f90:156-160), nsteps RK3 steps, arrival time in seconds
 * after cf_server_start.  Arrays are copied. */
int cf_server_submit(cf_server *s, const cf_params *p, const double *x, const double *y, const double *z,
                     const double *U, const double *V, const double *W, int nsteps, double arrival,
                     long long *id);
/* batching (before cf_server_start): cases of one size class (every option
 * but nu, dPdx, dPdz, u_centerline and tau_mean_eqwm shared, the same grid;
 * see cf_batch_create for the classes taken) are stepped together, up to
 * max_members per group, in stretches of at most chunk steps; cases join a
 * running group between stretches.  Only cases of at most max_cells unique
 * cells are batched (0: no limit); larger ones run one per stream, as every
 * case does with max_members 0 (the default). */
int cf_server_set_batching(cf_server *s, int max_members, int chunk, long long max_cells);
/* start the clock, the scheduler and the admission worker (cases may be
 * submitted before and after) */
int cf_server_start(cf_server *s);
int cf_server_cancel(cf_server *s, long long id);
/* wait until every submitted case is done, cancelled or failed; 1 on timeout
 * (timeout_s <= 0 waits without limit) */
int cf_server_wait(cf_server *s, double timeout_s);
/* state: 0 pending, 1 preparing, 2 running, 3 done, 4 cancelled, 5 failed */
int cf_server_status(cf_server *s, long long id, int *state, int *steps_done);
/* final field of a finished case and its 7 scalars (as cf_get_scalars) */
int cf_server_result(cf_server *s, long long id, double *U, double *V, double *W, double *scalars);
/* info (up to 14): arrival, admitted, ready, finished (s after start), steps
 * done, steps asked, footprint bytes, step ms p50, p99, max, state, unique
 * cells, precision, iwall_model */
int cf_server_case_info(cf_server *s, long long id, double *info, int n);
/* per-step GPU durations (ms) of a case; returns how many were written */
int cf_server_step_times(cf_server *s, long long id, float *ms, int n);
/* out (up to 8): last finish time, peak arena bytes in use, peak running
 * cases, cases that waited for memory at the head of the queue, arena bytes,
 * cases submitted, batch groups formed, most members stepped together */
int cf_server_summary(cf_server *s, double *out, int n);

/* Per-kernel timing with HIP events recorded on the context's stream. */
int cf_timing_enable(cf_ctx *c, int on);
/* writes "name avg_ms count\n" lines; returns 0 */
int cf_timing_report(cf_ctx *c, char *buf, int len);
int cf_timing_reset(cf_ctx *c);
/* bytes of device memory held by the context */
long long cf_device_bytes(cf_ctx *c);
/* synchronise the context's stream */
int cf_sync(cf_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
