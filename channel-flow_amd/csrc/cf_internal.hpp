// Internal context types of the library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/channel_flow.h"
#include "cf_eqwm.hpp"
#include "cf_kernels.hpp"
#include "cf_mlp.hpp"

namespace cf {

void set_error(const std::string &m);

// tiny single-thread kernels (cf_stencil.hip)
void launch_reset_counters(Scalars *sc, hipStream_t st);
void launch_dt_final_only(Scalars *sc, double cfl, hipStream_t st);
void launch_advance_time(Scalars *sc, hipStream_t st);
void launch_ubar_shift(Scalars *sc, double *ubar, int nyg, int neumann, hipStream_t st);

// Host copy of the grids and operators of initialization.f90.
struct Host {
    cf_params prm;
    Dims d;
    std::vector<double> x, y, z, xg, yg, zg;                // 0-based as given / derived
    std::vector<double> x1, y1, z1, xg1, yg1, zg1;          // 1-based copies
    std::vector<double> rdxf, rdxc, rdyf, rdyc, rdzf, rdzc;  // 1-based reciprocals
    std::vector<double> dxc, dyc, dzc, dyg;                 // 1-based spacings
    std::vector<double> w0, w1, kxx, kzz, dd, dl, du, delta;
    std::vector<int> jjz, jjx;
    std::vector<double> wz, wx;
    double dxmin = 0, dymin = 0, dzmin = 0, norm_u = 0, norm_v = 0, dt_vis = 0;
    Host(const cf_params &p, const double *x_, const double *y_, const double *z_);
    std::string validate() const;
};

// A device region handed to a context: buffers are carved from it in
// order (256-byte aligned).  dry: count the footprint without touching the
// device.
struct Slab {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    bool dry = false;
};

struct CtxBase {
    Host h;
    Slab *slab = nullptr;          // optional: allocate from this region
    hipStream_t stream = nullptr;  // optional: a caller's stream (set before init)
    bool stream_ext = false;
    explicit CtxBase(const Host &hh) : h(hh) {}
    virtual ~CtxBase() {}
    virtual int init() = 0;
    virtual int set_state(const double *U, const double *V, const double *W) = 0;
    virtual int get_state(double *U, double *V, double *W) = 0;
    virtual int get_pressure(double *P) = 0;
    virtual int compute_pressure(double *P) = 0;
    virtual int set_nu_t(const double *nt) = 0;
    virtual int get_nu_t(double *nt) = 0;
    virtual int get_avg_nu_t(double *avg) = 0;
    virtual int get_wall(double *ax, double *az, double *ti) = 0;
    virtual int set_tau_init(const double *ti) = 0;
    virtual int set_scalars(const double *s) = 0;
    virtual int get_scalars(double *s) = 0;
    virtual int get_raw_scalars(Scalars *s) = 0;
    virtual int init_mass_flow() = 0;
    virtual int apply_bc() = 0;
    virtual int compute_dt(double *dt) = 0;
    virtual int rk3_step() = 0;
    virtual int run_steps(int n, bool graph) = 0;
    virtual int projection_step(int rk_step) = 0;
    virtual int compute_rhs(const double *U, const double *V, const double *W, double *ru, double *rv,
                            double *rw) = 0;
    virtual int compute_eddy_viscosity() = 0;
    virtual int compute_wall_model(int *n_nan, int *n_iter) = 0;
    virtual int mean_mass_flow_u(double *q) = 0;
    virtual int mean_mass_flow_v(double *q) = 0;
    virtual int mean_centerline_u(double *q) = 0;
    virtual int check_divergence(double *ms, double *ma) = 0;
    virtual int statistics(double *out) = 0;
    virtual int stats_accumulate() = 0;
    virtual int stats_window(double *out, double *count) = 0;
    virtual int stats_reset() = 0;
    virtual void set_timing(bool on) = 0;
    virtual void timing_reset() = 0;
    virtual std::string timing_report() = 0;
    virtual size_t device_bytes() const = 0;
    virtual int sync() = 0;
    virtual hipStream_t get_stream() const = 0;
    virtual int attach_mlp(Mlp *m) = 0;
    virtual int wall_fallbacks() = 0;
    virtual int prepare_graphs() = 0;
    // A batch member (cf_batch) is stepped only by batch_run on the first
    // member: each step kernel is launched once for all members, whose slabs
    // sit at the byte stride cs.
    bool batch_member = false;
    virtual int batch_run(const std::vector<CtxBase *> &members, int64_t cs, int nsteps, bool graph) = 0;
    // ping-pong parity of the state buffers (members of a batch share it)
    virtual int parity() const = 0;
    virtual void set_parity(int p) = 0;
};

// The size class of a case for batched stepping: its options less the
// per-case ones (nu, forcing, centreline target, fallback stress) and its
// grids, as bytes; empty when the batched step does not take the case.
std::string batch_class_key(const Host &h);

// a context for one case: device buffers from `slab` when given (else
// hipMalloc), work on `stream` when given (else its own); null on failure
std::unique_ptr<CtxBase> make_ctx(const Host &h, Slab *slab, hipStream_t stream);
// device bytes a context of these parameters takes from a slab
size_t ctx_footprint(const Host &h);

template <typename S>
struct Ctx : CtxBase {
    explicit Ctx(const Host &hh) : CtxBase(hh) {}
    ~Ctx() override;

    int dev = -1;
    std::vector<void *> owned;
    size_t bytes = 0;
    size_t n3[3] = {0, 0, 0};
    S *st[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int cur = 0;
    S *G[3] = {nullptr, nullptr, nullptr};
    S *NT = nullptr;
    bool nut_user = false;
    Cx<S> *spec = nullptr;
    double *cscr = nullptr;  // y-solve scratch (lane form only)
    DsmWork dsm{};           // dynamic Smagorinsky work arrays (les_model 2 only)
    S *phi1 = nullptr, *phi23 = nullptr;
    double *part = nullptr, *stats_out = nullptr;
    double *mfpart = nullptr;  // per-workgroup mass-flow partials of the last substep
    double *dtfp = nullptr;    // per-workgroup compute_dt maxima of the last projection
    bool dt_fresh = false;     // dt_next holds compute_dt of the current state
    // dt taken from the last projection: adaptive dt (a fixed dt needs no
    // reduction) and V untouched after that projection (no y mass-flow
    // correction)
    bool fuse_dt() const { return h.prm.cfl >= 0 && h.prm.y_mass_cte == 0; }
    void step_head();
    double *dtscr = nullptr;
    double *ubar = nullptr;            // mixed path: FP64 profile U = ubar(j) + U'
    std::vector<double> ubar_h;
    int set_profile(const double *U);
    int sync_profile();
    Scalars *sc = nullptr;
    double *m_rdxf, *m_rdxc, *m_rdyf, *m_rdyc, *m_rdzf, *m_rdzc, *m_w0, *m_w1, *m_dyg;
    double *m_dxc, *m_dyc, *m_dzc, *m_x, *m_y, *m_z, *m_xg, *m_yg, *m_zg;
    double *m_kxx, *m_kzz, *m_dd, *m_dl, *m_du, *m_delta, *m_wz, *m_wx;
    int *m_jjz, *m_jjx;
    Metrics met{};
    WallTables wt{};
    FFTPlan pM{}, pZ{};
    Cx<S> *twM = nullptr, *twN = nullptr, *twZ = nullptr;
    double *alpha_x = nullptr, *alpha_z = nullptr, *azt = nullptr, *tau_init = nullptr;
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    int bn = 1;            // cases the step kernels launch for (batch_run)
    int64_t bcs = 0;       // their byte stride
    bool batch_running = false;
    // optional learned wall-model hook: fills predicted stresses and a mask
    std::function<void(Ctx<S> *, S *const *, const double **, const int **)> surrogate;

    // timing
    bool timing = false;
    struct Rec { const char *name; hipEvent_t a, b; };
    std::vector<Rec> pending, done_ev;
    std::vector<hipEvent_t> evpool, ev_free;
    std::map<std::string, std::pair<double, long>> tstat;
    void tic(const char *name);
    void toc();
    hipEvent_t ev_get();
    void collect();

    bool prec_mixed() const { return std::is_same<S, float>::value; }
    void invalidate_graphs() {
        for (auto &g : graphs)
            if (g) { hipGraphExecDestroy(g); g = nullptr; }
    }
    void ghosts(S *const *F, int mask);
    // gcopy: the projection kernels also write the periodic and wall ghosts
    void poisson(S *const *F, S *phi, const Scalars *corr = nullptr, const DtParts *dtp = nullptr, int gcopy = 0);
    void eddy_viscosity(S *const *F);
    void wall_model(S *const *F);
    void dt_launch();
    void step_launch();
    int put(S *dst, const double *src, size_t n);
    int get(double *dst, const S *src, size_t n);

    int init() override;
    int set_state(const double *U, const double *V, const double *W) override;
    int get_state(double *U, double *V, double *W) override;
    int get_pressure(double *P) override;
    int compute_pressure(double *P) override;
    int set_nu_t(const double *nt) override;
    int get_nu_t(double *nt) override;
    int get_avg_nu_t(double *avg) override;
    int get_wall(double *ax, double *az, double *ti) override;
    int set_tau_init(const double *ti) override;
    int set_scalars(const double *s) override;
    int get_scalars(double *s) override;
    int get_raw_scalars(Scalars *s) override;
    int init_mass_flow() override;
    int apply_bc() override;
    int compute_dt(double *dt) override;
    int rk3_step() override;
    int run_steps(int n, bool graph) override;
    int projection_step(int rk_step) override;
    int compute_rhs(const double *U, const double *V, const double *W, double *ru, double *rv,
                    double *rw) override;
    int compute_eddy_viscosity() override;
    int compute_wall_model(int *n_nan, int *n_iter) override;
    int mean_mass_flow_u(double *q) override;
    int mean_mass_flow_v(double *q) override;
    int mean_centerline_u(double *q) override;
    int check_divergence(double *ms, double *ma) override;
    int statistics(double *out) override;
    int stats_accumulate() override;
    int stats_window(double *out, double *count) override;
    int stats_reset() override;
    double *stats_acc = nullptr;
    void set_timing(bool on) override { collect(); timing = on; }
    void timing_reset() override { collect(); tstat.clear(); }
    std::string timing_report() override;
    size_t device_bytes() const override { return bytes; }
    int sync() override {
        if (hipStreamSynchronize(stream) != hipSuccess) { set_error("stream synchronize failed"); return -1; }
        return 0;
    }
    hipStream_t get_stream() const override { return stream; }
    double *wall_u = nullptr, *tau_pred = nullptr;
    int *use_pred = nullptr;
    int attach_mlp(Mlp *m) override;
    int wall_fallbacks() override;
    int prepare_graphs() override;
    int batch_run(const std::vector<CtxBase *> &members, int64_t cs, int nsteps, bool graph) override;
    int parity() const override { return cur; }
    void set_parity(int p) override { cur = p; }
};

}  // namespace cf
