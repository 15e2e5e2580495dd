// Shared definitions for the channel-flow HIP kernels (gfx950).
//
// Layout in HBM: every velocity-like array keeps the reference's Fortran
// layout, x fastest, then y, then z, with the reference's ghost extents
// (global.f90:82-84, initialization.f90:121-124):
//   U (nx ,nyg,nzg)   V (nxg,ny ,nzg)   W (nxg,nyg,nz )   nu_t (nxg,nyg,nzg)
// so host mirrors and device buffers share one indexing scheme.  Kernels
// use 1-based (i,j,k) like the reference.  1-D metric arrays are stored
// 1-based too (index 0 unused).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cf {

struct Dims {
    int nx, ny, nz;       // faces incl. periodic duplicates (x, z) and walls (y)
    int nxg, nyg, nzg;    // centres + ghosts
    int nxp, nzp, nyp;    // unique periodic points in x, z; interior centres in y
    int mh;               // x half spectrum length nxp/2+1
};

// 1-D grid metrics, all 1-based (index 0 unused), double precision.
struct Metrics {
    const double *rdxf;   // 1/(x(i)-x(i-1))      i = 2..nx
    const double *rdxc;   // 1/(xg(i+1)-xg(i))    i = 1..nxg-1
    const double *rdyf;   // 1/(y(j)-y(j-1))      j = 2..ny
    const double *rdyc;   // 1/(yg(j+1)-yg(j))    j = 1..nyg-1
    const double *rdzf;   // 1/(z(k)-z(k-1))      k = 2..nz
    const double *rdzc;   // 1/(zg(k+1)-zg(k))    k = 1..nzg-1
    const double *w0, *w1;  // y-face weights (initialization.f90:387-388)
    const double *dyg;    // yg(j)-yg(j-1)        j = 2..nyg
    const double *dxc, *dyc, *dzc;  // xg(i+1)-xg(i) etc. i = 1..nxg-1 (compute_dt's spacings)
};

// Device-resident scalars shared by all kernels of a step (no host sync).
struct Scalars {
    double dt;        // current time step
    double dPdx;      // forcing carried between steps (time_integration.f90:156)
    double dPdz;
    double dPdy;
    double t;
    double qflow_x0;  // target bulk flow (initialization.f90:460)
    double qflow_y0;
    double corr;      // last mass-flow correction (U increment)
    double norm_u;    // trapezoid normalisation (mass_flow.f90:41), host-computed
    double norm_v;
    double dt_conv[3];
    double qbar_u;    // bulk-flow contribution of the stored profile ubar (mixed path)
    double corr_v;    // last y mass-flow correction (V increment, mass_flow.f90:174)
    double dt_next;   // compute_dt of the state at the end of the last step
    double ucl;       // centreline target Ucenterline_x (x_mass_cte = 2)
    double nu;        // kinematic viscosity of the case (read by step kernels of a batch)
    double dt_vis;    // viscous limit of compute_dt (time_integration.f90:205-212)
    double tau_mean;  // wall-model fallback stress of the case (tau_mean_eqwm)
    double pad[3];    // [0] last bulk (or centreline) mean of U, [1] last bulk mean of V
    int n_nan_eqwm;
    int n_iter_eqwm;
    int n_fallback_mlp;
    int pad_i;
};

// Partial maxima of compute_dt's spacing/|u| (time_integration.f90:191-196)
// taken by the kernels that write the final velocities of a step: one value
// per workgroup (null: not taken).  U's values are those of the stored U plus
// ubar(j) (mixed path) plus corr when add_corr (the profile shift that
// follows the step).
struct DtParts {
    double *pu, *pv, *pw;
    const double *ubar;
    const Scalars *sc;
    int add_corr;
};

// A batch of cases of one size class keeps every per-case buffer at one byte
// stride cs from the first case's: the step kernels take the case index from
// blockIdx.y and move each per-case pointer by its offset (cs = 0: one case).
// Grid metrics, transform tables and the y operator are the class's own and
// are read from the first case.
template <typename P>
__device__ __forceinline__ P *case_ptr(P *p, int64_t cs) {
    return (p && cs) ? reinterpret_cast<P *>(reinterpret_cast<char *>(const_cast<void *>(
                                                 static_cast<const void *>(p))) +
                                             (int64_t)blockIdx.y * cs)
                     : p;
}

__host__ __device__ inline int64_t i3(int n1, int n2, int i, int j, int k) {
    return (int64_t)(i - 1) + (int64_t)n1 * ((int64_t)(j - 1) + (int64_t)n2 * (int64_t)(k - 1));
}

template <typename T> struct Cx { T x, y; };

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) { return {a.x - b.x, a.y - b.y}; }

// Launch geometry helper: blocks are remapped so that consecutive tiles land
// on the same XCD (workgroups are dispatched round-robin over 8 XCDs), which
// keeps stencil neighbours inside one XCD's L2.
// Bijective: XCD x receives blocks x, x+8, x+16, ... and owns the contiguous
// tile range [x*q + min(x,r), ...) of length q + (x<r), q = n/8, r = n%8.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int nxcd = 8;
    int q = nblocks / nxcd, r = nblocks % nxcd;
    int xcd = bid % nxcd;
    int slot = bid / nxcd;
    return xcd * q + (xcd < r ? xcd : r) + slot;
}

// Ghost copies of one unique value v at (i, j, k) of an (n1, n2, *) field,
// as apply_boundary_conditions (boundary_conditions.f90:20-75, one rank)
// leaves them with Dirichlet walls: gi, gk are the ghost column and plane
// whose periodic image (i, k) is (0: none), gj the wall ghost row mirroring
// row j (0: none), which takes -v.  Every combination other than (i, j, k)
// itself is written, so each ghost cell is written once, by its source.
template <typename S>
__device__ inline void put_ghost_copies(S *F, int n1, int n2, int i, int j, int k, S v, int gi, int gj, int gk) {
    for (int a = 0; a < 2; a++) {
        const int ii = a ? gi : i;
        if (!ii) continue;
        for (int c = 0; c < 2; c++) {
            const int kk = c ? gk : k;
            if (!kk) continue;
            for (int b = 0; b < 2; b++) {
                const int jj = b ? gj : j;
                if (!jj || (a == 0 && b == 0 && c == 0)) continue;
                F[i3(n1, n2, ii, jj, kk)] = b ? (S)(-v) : v;
            }
        }
    }
}
// the ghost column or plane whose periodic image is unique index i: faces
// (1..n, unique 2..n-1): 1 <- n-1, n <- 2; centres (1..n, unique 2..n-2):
// 1 <- n-2, n-1 <- 2, n <- 3
__host__ __device__ inline int ghost_of_face(int i, int n) { return i == n - 1 ? 1 : (i == 2 ? n : 0); }
__host__ __device__ inline int ghost_of_ctr(int i, int n) {
    return i == n - 2 ? 1 : (i == 2 ? n - 1 : (i == 3 ? n : 0));
}

}  // namespace cf
