// Learned wall-model surrogate (MLP) -- device state and host methods.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "cf_common.hpp"

struct cf_mlp;

namespace cf {

constexpr int MLP_H = 64;

struct Mlp {
    hipStream_t stream = nullptr;
    std::vector<void *> owned;
    float *P = nullptr, *m1 = nullptr, *m2 = nullptr;      // FP32 master weights, Adam moments
    unsigned short *W2b = nullptr, *W3b = nullptr;          // BF16 copies of the 64x64 layers
    float *nrm_d = nullptr, *hull_d = nullptr, *loss_d = nullptr;
    float nrm_h[16] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 1};     // feature mean[4], std[4], ln tau mean, std
    float hull_h[8] = {-1e30f, -1e30f, -1e30f, -1e30f, 1e30f, 1e30f, 1e30f, 1e30f};
    float *X = nullptr, *Yd = nullptr, *yo = nullptr, *gpart = nullptr, *lpart = nullptr, *gbuf = nullptr;
    size_t cap_data = 0, cap_waves = 0, ndata = 0;
    int adam_t = 0;

    ~Mlp();
    int init(unsigned seed);
    int set_params(const float *h);
    int get_params(float *h);
    int ensure(size_t ndata, int nwaves);
    int fit_normalisation(const double *feat_raw, const double *tau, size_t n);
    int upload(const double *feat_raw, const double *tau, size_t n);
    int step(size_t off, int bs, float lr, bool update, float *grad_out, float *loss_out);
    void wall_predict(int n, const double *u_les, double h, double nu, const double *dpdx_dev,
                      double *tau_pred, int *use_pred, Scalars *sc, hipStream_t st);
};

// the model behind a C-ABI handle (cf_mlp.hip)
Mlp *mlp_of(struct ::cf_mlp *p);

}  // namespace cf
