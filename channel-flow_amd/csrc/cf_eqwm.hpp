// Wall-model (EQWM) device interfaces.
#pragma once

#include "cf_kernels.hpp"

namespace cf {

constexpr int EQWM_MAXNCV = 128;
constexpr int EQWM_WG = 128;  // lanes per block (one wall point each; only the grid is in LDS)

// 3-point Lagrange stencils of interpolate_z_2nd / interpolate_x_2nd
// (interpolation.f90:114-252) for the wall planes: 1-based centre index and
// three weights per target point.
struct WallTables {
    const int *jjz;     // nzg entries
    const double *wz;   // 3*nzg
    const int *jjx;     // nx entries
    const double *wx;   // 3*nx
};

size_t eqwm_lds_bytes(int ncv, int wg);
int eqwm_block(int ncv);

template <typename S>
void launch_wall_model(const Dims &d, const S *U, const S *W, const S *NT, const WallTables &t,
                       double *tau_init, double *alpha_x, double *azt, double *az, Scalars *sc,
                       EqwmParams p, double nu, double h, int jref, double tau_mean,
                       const double *tau_pred, const int *use_pred, const double *ubar,
                       hipStream_t st);
template <typename S>
void launch_wall_inputs(const Dims &d, const S *U, const S *W, const WallTables &t, int jref,
                        const double *ubar, double *out, hipStream_t st);

}  // namespace cf
