// Dynamic Smagorinsky model, Lilly's form (subgrid.f90:376-538), on gfx950.
//
//   Lij = hat(ui uj) - hat(ui) hat(uj)                 (cell centres)
//   Mij = 4 |hat S| hat(S)ij - hat(|S| Sij)            (filter width ratio 2)
//   C(y) = <Lij Mij> / <Mij Mij>   (plane sums, off-diagonals twice)
//   nu_t = max(-C(y)/2 |S|, 0), wall ghosts odd (Dirichlet) or even
//
// hat() is filter_xzy (subgrid.f90:134-192): Simpson (1/6 2/3 1/6) in x, z
// and y, the y weights one-sided (2/3 1/3) on the first and last row of the
// filtered range.  The filter is linear, so it is applied as an x-z pass per
// row followed by the three-row y combination, which lets one thread march
// up a column reusing the x-z sums of the rows it has passed.
//
// Everything is FP64, also when the fields are stored as FP32 deviations
// (the mixed path): Lij is a difference of nearly equal filtered products.
//
// Kernels, in launch order:
//   k_dsm_prep       centre velocities, |S| and |S|Sij of the unfiltered
//                    field at every centre (ghost columns take the values of
//                    their periodic images)                     -> Q[10]
//   k_dsm_filter_vel filtered U, V, W on their own grids, all rows  -> Fu
//   k_dsm_lm         per cell Lij, hat Sij (from Fu), Mij; LM and MM summed
//                    over each wave's 64 cells of a row         -> part
//   k_dsm_avg        one block per row: fixed-order sum of the wave partials,
//                    C(y) = LM/MM                               -> avg
//   k_dsm_nut        nu_t at every centre, wall ghosts, avg ghosts
//
// The reference sums the plane over centres 2..nxg-1 x 2..nzg-1, which holds
// the periodic images nxg-1 (of 2) and nzg-1 (of 2): cells in the first
// unique column or plane count twice (four times on both).
#include <algorithm>

#include "cf_kernels.hpp"

namespace cf {

namespace {

constexpr int DSM_WG = 256;
constexpr int DSM_LM_ROWS = 16;  // rows per thread in k_dsm_lm (one halo row each side)

// periodic image of a centre index (ghosts 1, n-1, n of a centred axis of n
// points: 1 <- n-2, n-1 <- 2, n <- 3) and of a face index (1 <- n-1, n <- 2)
__device__ __forceinline__ int img_c(int i, int n) { return i == 1 ? n - 2 : (i >= n - 1 ? i - (n - 3) : i); }
__device__ __forceinline__ int img_f(int i, int n) { return i == 1 ? n - 1 : (i == n ? 2 : i); }

// compute_Sij (subgrid.f90:891-949) at centre (i, j, k): three diagonal
// strains at the centre, the off-diagonal edge values averaged over the four
// edges around the centre, |S| = sqrt(2 Sij Sij)
template <class FU, class FV, class FW>
__device__ __forceinline__ double strain6(const Metrics &m, int i, int j, int k, const FU &u, const FV &v,
                                          const FW &w, double s[6]) {
    const double rxc0 = m.rdxc[i - 1], rxc1 = m.rdxc[i];
    const double ryc0 = m.rdyc[j - 1], ryc1 = m.rdyc[j];
    const double rzc0 = m.rdzc[k - 1], rzc1 = m.rdzc[k];
    s[0] = (u(i, j, k) - u(i - 1, j, k)) * m.rdxf[i];
    s[1] = (v(i, j, k) - v(i, j - 1, k)) * m.rdyf[j];
    s[2] = (w(i, j, k) - w(i, j, k - 1)) * m.rdzf[k];
    auto e4 = [&](int I, int J, int K, double ryc, double rxc) {
        return 0.5 * ((u(I - 1, J, K) - u(I - 1, J - 1, K)) * ryc + (v(I, J - 1, K) - v(I - 1, J - 1, K)) * rxc);
    };
    auto e5 = [&](int I, int J, int K, double rzc, double rxc) {
        return 0.5 * ((u(I - 1, J, K) - u(I - 1, J, K - 1)) * rzc + (w(I, J, K - 1) - w(I - 1, J, K - 1)) * rxc);
    };
    auto e6 = [&](int I, int J, int K, double rzc, double ryc) {
        return 0.5 * ((v(I, J - 1, K) - v(I, J - 1, K - 1)) * rzc + (w(I, J, K - 1) - w(I, J - 1, K - 1)) * ryc);
    };
    s[3] = 0.25 * (e4(i, j, k, ryc0, rxc0) + e4(i + 1, j, k, ryc0, rxc1) + e4(i, j + 1, k, ryc1, rxc0) +
                   e4(i + 1, j + 1, k, ryc1, rxc1));
    s[4] = 0.25 * (e5(i, j, k, rzc0, rxc0) + e5(i + 1, j, k, rzc0, rxc1) + e5(i, j, k + 1, rzc1, rxc0) +
                   e5(i + 1, j, k + 1, rzc1, rxc1));
    s[5] = 0.25 * (e6(i, j, k, rzc0, ryc0) + e6(i, j + 1, k, rzc0, ryc1) + e6(i, j, k + 1, rzc1, ryc0) +
                   e6(i, j + 1, k + 1, rzc1, ryc1));
    double q = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) q = q + (c < 3 ? 2.0 : 4.0) * s[c] * s[c];
    return sqrt(q);
}

// x-z Simpson weights fil(a, c), a, c in {-1, 0, 1}
__device__ __forceinline__ double fil_xz(int a, int c) {
    return (a == 0 && c == 0) ? 4.0 / 9.0 : ((a == 0 || c == 0) ? 1.0 / 9.0 : 1.0 / 36.0);
}

// ---------------------------------------------------------------------------
// Q layout: 10 centre arrays (nxg, nyg, nzg), rows 2..nyg-1 written:
//   0..2 centre velocities (interpolate_x/y/z, subgrid.f90:391-393)
//   3..8 |S| Sij, 9 |S|                     (compute_Sij of the raw field)
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_prep(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V, const S *__restrict__ W,
           const double *__restrict__ ubar, double *__restrict__ Q, int64_t nC) {
    const int64_t ncell = (int64_t)d.nxg * (d.nyg - 2) * d.nzg;
    const int64_t t = (int64_t)blockIdx.x * DSM_WG + threadIdx.x;
    if (t >= ncell) return;
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny;
    const int i = 1 + (int)(t % nxg);
    const int64_t r = t / nxg;
    const int j = 2 + (int)(r % (nyg - 2));
    const int k = 1 + (int)(r / (nyg - 2));
    const int ii = img_c(i, nxg), kk = img_c(k, d.nzg);
    auto u = [&](int I, int J, int K) {
        return (double)U[i3(nx, nyg, I, J, K)] + (ubar ? ubar[J] : 0.0);
    };
    auto v = [&](int I, int J, int K) { return (double)V[i3(nxg, ny, I, J, K)]; };
    auto w = [&](int I, int J, int K) { return (double)W[i3(nxg, nyg, I, J, K)]; };
    double s[6];
    const double sm = strain6(m, ii, j, kk, u, v, w, s);
    const int64_t q = i3(nxg, nyg, i, j, k);
    Q[q] = 0.5 * (u(ii - 1, j, kk) + u(ii, j, kk));
    Q[nC + q] = 0.5 * (v(ii, j - 1, kk) + v(ii, j, kk));
    Q[2 * nC + q] = 0.5 * (w(ii, j, kk - 1) + w(ii, j, kk));
#pragma unroll
    for (int c = 0; c < 6; c++) Q[(3 + c) * nC + q] = sm * s[c];
    Q[9 * nC + q] = sm;
}

// ---------------------------------------------------------------------------
// Filtered velocities (subgrid.f90:428-443) on the full grids of U, V, W:
// blockIdx.y selects the component.  Each point takes the filter at its
// periodic image; all rows, one-sided y weights on the first and last.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_filter_vel(Dims d, const S *__restrict__ U, const S *__restrict__ V, const S *__restrict__ W,
                 const double *__restrict__ ubar, double *__restrict__ Fu) {
    const int comp = blockIdx.y;
    const int n1 = comp == 0 ? d.nx : d.nxg, n2 = comp == 1 ? d.ny : d.nyg, n3 = comp == 2 ? d.nz : d.nzg;
    const int64_t n = (int64_t)n1 * n2 * n3;
    const int64_t t = (int64_t)blockIdx.x * DSM_WG + threadIdx.x;
    if (t >= n) return;
    const int I = 1 + (int)(t % n1);
    const int64_t r = t / n1;
    const int J = 1 + (int)(r % n2);
    const int K = 1 + (int)(r / n2);
    const int ii = comp == 0 ? img_f(I, n1) : img_c(I, n1);
    const int kk = comp == 2 ? img_f(K, n3) : img_c(K, n3);
    const S *F = comp == 0 ? U : (comp == 1 ? V : W);
    const bool addu = comp == 0 && ubar != nullptr;
    auto f = [&](int a, int b, int c) {
        return (double)F[i3(n1, n2, a, b, c)] + (addu ? ubar[b] : 0.0);
    };
    double acc = 0.0;
#pragma unroll
    for (int a = -1; a <= 1; a++)
#pragma unroll
        for (int c = -1; c <= 1; c++) {
            double tt;
            if (J == 1) tt = 2.0 / 3.0 * f(ii + a, J, kk + c) + 1.0 / 3.0 * f(ii + a, J + 1, kk + c);
            else if (J == n2) tt = 1.0 / 3.0 * f(ii + a, J - 1, kk + c) + 2.0 / 3.0 * f(ii + a, J, kk + c);
            else
                tt = 1.0 / 6.0 * f(ii + a, J - 1, kk + c) + 2.0 / 3.0 * f(ii + a, J, kk + c) +
                     1.0 / 6.0 * f(ii + a, J + 1, kk + c);
            acc = acc + tt * fil_xz(a, c);
        }
    const int64_t off = comp == 0 ? 0 : (comp == 1 ? (int64_t)d.nx * d.nyg * d.nzg
                                                   : (int64_t)d.nx * d.nyg * d.nzg + (int64_t)d.nxg * d.ny * d.nzg);
    Fu[off + t] = acc;
}

// x-z filtered sums of one row r at column (i, k): 3 velocities, 6 products
// (uu vv ww uv uw vw), 6 |S|Sij
__device__ __forceinline__ void dsm_row(const double *__restrict__ Q, int64_t nC, int nxg, int nyg, int i, int r,
                                        int k, double F[15]) {
#pragma unroll
    for (int c = 0; c < 15; c++) F[c] = 0.0;
#pragma unroll
    for (int a = -1; a <= 1; a++)
#pragma unroll
        for (int c = -1; c <= 1; c++) {
            const double f = fil_xz(a, c);
            const int64_t q = i3(nxg, nyg, i + a, r, k + c);
            const double u = Q[q], v = Q[nC + q], w = Q[2 * nC + q];
            F[0] += f * u;
            F[1] += f * v;
            F[2] += f * w;
            F[3] += f * (u * u);
            F[4] += f * (v * v);
            F[5] += f * (w * w);
            F[6] += f * (u * v);
            F[7] += f * (u * w);
            F[8] += f * (v * w);
#pragma unroll
            for (int s = 0; s < 6; s++) F[9 + s] += f * Q[(3 + s) * nC + q];
        }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// ---------------------------------------------------------------------------
// Lij, Mij and their plane products.  Block: 64 columns in x by 4 in z; one
// thread marches DSM_LM_ROWS rows of its column (blockIdx.y selects the row
// chunk).  part[(cw * nyp + j - 2) * 2 + {0, 1}] = wave sums of LM, MM, with
// cw = (4 bz + wave) * tiles_x + bx.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(DSM_WG)
k_dsm_lm(Dims d, Metrics m, const double *__restrict__ Q, const double *__restrict__ Fu, double *__restrict__ part,
         int tiles_x) {
    const int nxg = d.nxg, nyg = d.nyg, nzg = d.nzg, nx = d.nx, ny = d.ny;
    const int64_t nC = (int64_t)nxg * nyg * nzg;
    const int bx = blockIdx.x % tiles_x, bz = blockIdx.x / tiles_x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = 2 + bx * 64 + lane, k = 2 + bz * 4 + wv;
    const bool act = i <= nxg - 2 && k <= nzg - 2;
    const int cw = (4 * bz + wv) * tiles_x + bx;
    const int nyp = nyg - 2;
    const int j0 = 2 + blockIdx.y * DSM_LM_ROWS;
    const int j1 = min(j0 + DSM_LM_ROWS - 1, nyg - 1);
    const double wgt = (i == 2 ? 2.0 : 1.0) * (k == 2 ? 2.0 : 1.0);
    const double *Uf = Fu, *Vf = Fu + (int64_t)nx * nyg * nzg, *Wf = Vf + (int64_t)nxg * ny * nzg;
    auto uf = [&](int I, int J, int K) { return Uf[i3(nx, nyg, I, J, K)]; };
    auto vf = [&](int I, int J, int K) { return Vf[i3(nxg, ny, I, J, K)]; };
    auto wf = [&](int I, int J, int K) { return Wf[i3(nxg, nyg, I, J, K)]; };
    double Fm[15], F0[15], Fp[15];
#pragma unroll
    for (int c = 0; c < 15; c++) { Fm[c] = 0.0; Fp[c] = 0.0; }
    if (act) {
        if (j0 > 2) dsm_row(Q, nC, nxg, nyg, i, j0 - 1, k, Fm);
        dsm_row(Q, nC, nxg, nyg, i, j0, k, F0);
    }
    for (int j = j0; j <= j1; j++) {
        double lm = 0.0, mm = 0.0;
        if (act) {
            if (j < nyg - 1) dsm_row(Q, nC, nxg, nyg, i, j + 1, k, Fp);
            const double am = j == 2 ? 0.0 : (j == nyg - 1 ? 1.0 / 3.0 : 1.0 / 6.0);
            const double a0 = 2.0 / 3.0;
            const double ap = j == nyg - 1 ? 0.0 : (j == 2 ? 1.0 / 3.0 : 1.0 / 6.0);
            double h[15];
#pragma unroll
            for (int c = 0; c < 15; c++) h[c] = am * Fm[c] + a0 * F0[c] + ap * Fp[c];
            double L[6];
            L[0] = h[3] - h[0] * h[0];
            L[1] = h[4] - h[1] * h[1];
            L[2] = h[5] - h[2] * h[2];
            L[3] = h[6] - h[0] * h[1];
            L[4] = h[7] - h[0] * h[2];
            L[5] = h[8] - h[1] * h[2];
            double sf[6];
            const double smf = strain6(m, i, j, k, uf, vf, wf, sf);
            double M[6];
#pragma unroll
            for (int c = 0; c < 6; c++) M[c] = 4.0 * smf * sf[c] - h[9 + c];
            lm = L[0] * M[0] + L[1] * M[1] + L[2] * M[2] + 2.0 * (L[3] * M[3] + L[4] * M[4] + L[5] * M[5]);
            mm = M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + 2.0 * (M[3] * M[3] + M[4] * M[4] + M[5] * M[5]);
            lm *= wgt;
            mm *= wgt;
#pragma unroll
            for (int c = 0; c < 15; c++) { Fm[c] = F0[c]; F0[c] = Fp[c]; }
        }
        lm = wave_sum(lm);
        mm = wave_sum(mm);
        if (lane == 0) {
            const int64_t p = ((int64_t)cw * nyp + (j - 2)) * 2;
            part[p] = lm;
            part[p + 1] = mm;
        }
    }
}

// one block per row: C(y) = sum LM / sum MM over the wave partials
__global__ void __launch_bounds__(DSM_WG)
k_dsm_avg(int ncw, int nyp, const double *__restrict__ part, double *__restrict__ avg) {
    __shared__ double sh[2][DSM_WG / 64];
    const int jr = blockIdx.x;  // row j = jr + 2
    double lm = 0.0, mm = 0.0;
    for (int c = threadIdx.x; c < ncw; c += DSM_WG) {
        const int64_t p = ((int64_t)c * nyp + jr) * 2;
        lm += part[p];
        mm += part[p + 1];
    }
    lm = wave_sum(lm);
    mm = wave_sum(mm);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sh[0][wv] = lm; sh[1][wv] = mm; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < DSM_WG / 64; q++) { a += sh[0][q]; b += sh[1][q]; }
        avg[jr + 1] = a / b;  // avg is 0-based over rows 1..nyg
    }
}

// nu_t at every centre: -C(y)/2 |S| clipped at zero; wall ghost rows odd
// (Dirichlet) or even (Neumann) after clipping; the same rule for C(y)
template <typename S>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_nut(Dims d, const double *__restrict__ Q, double *__restrict__ avg, S *__restrict__ NT, int dirichlet) {
    const int nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    const int64_t nC = (int64_t)nxg * nyg * nzg;
    const int64_t t = (int64_t)blockIdx.x * DSM_WG + threadIdx.x;
    if (t >= nC) return;
    const int i = 1 + (int)(t % nxg);
    const int64_t r = t / nxg;
    const int j = 1 + (int)(r % nyg);
    const int k = 1 + (int)(r / nyg);
    const int jj = j == 1 ? 2 : (j == nyg ? nyg - 1 : j);
    const double sgn = (jj != j && dirichlet) ? -1.0 : 1.0;
    const double val = fmax(-0.5 * avg[jj - 1] * Q[9 * nC + i3(nxg, nyg, i, jj, k)], 0.0);
    NT[t] = (S)(sgn * val);
    if (t == 0) {
        avg[0] = dirichlet ? -avg[1] : avg[1];
        avg[nyg - 1] = dirichlet ? -avg[nyg - 2] : avg[nyg - 2];
    }
}

inline int dsm_tiles_x(const Dims &d) { return (d.nxg - 3 + 63) / 64; }
inline int dsm_tiles_z(const Dims &d) { return (d.nzg - 3 + 3) / 4; }

}  // namespace

size_t dsm_part_elems(const Dims &d) {
    return (size_t)dsm_tiles_x(d) * dsm_tiles_z(d) * 4 * (d.nyg - 2) * 2;
}

template <typename S>
void launch_dsm(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W, const double *ubar,
                const DsmWork &w, S *NT, int dirichlet, hipStream_t st) {
    const int64_t nC = (int64_t)d.nxg * d.nyg * d.nzg;
    const int64_t nprep = (int64_t)d.nxg * (d.nyg - 2) * d.nzg;
    hipLaunchKernelGGL((k_dsm_prep<S>), dim3((unsigned)((nprep + DSM_WG - 1) / DSM_WG)), dim3(DSM_WG), 0, st, d, m,
                       U, V, W, ubar, w.Q, nC);
    const int64_t nmax = std::max((int64_t)d.nx * d.nyg * d.nzg,
                                  std::max((int64_t)d.nxg * d.ny * d.nzg, (int64_t)d.nxg * d.nyg * d.nz));
    hipLaunchKernelGGL((k_dsm_filter_vel<S>), dim3((unsigned)((nmax + DSM_WG - 1) / DSM_WG), 3), dim3(DSM_WG), 0, st,
                       d, U, V, W, ubar, w.Fu);
    const int tx = dsm_tiles_x(d), tz = dsm_tiles_z(d);
    const int nyp = d.nyg - 2;
    const int chunks = (nyp + DSM_LM_ROWS - 1) / DSM_LM_ROWS;
    hipLaunchKernelGGL(k_dsm_lm, dim3(tx * tz, chunks), dim3(DSM_WG), 0, st, d, m, w.Q, w.Fu, w.part, tx);
    hipLaunchKernelGGL(k_dsm_avg, dim3(nyp), dim3(DSM_WG), 0, st, tx * tz * 4, nyp, w.part, w.avg);
    hipLaunchKernelGGL((k_dsm_nut<S>), dim3((unsigned)((nC + DSM_WG - 1) / DSM_WG)), dim3(DSM_WG), 0, st, d, w.Q,
                       w.avg, NT, dirichlet);
}

template void launch_dsm<double>(const Dims &, const Metrics &, const double *, const double *, const double *,
                                 const double *, const DsmWork &, double *, int, hipStream_t);
template void launch_dsm<float>(const Dims &, const Metrics &, const float *, const float *, const float *,
                                const double *, const DsmWork &, float *, int, hipStream_t);

}  // namespace cf
