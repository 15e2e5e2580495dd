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
// Block 64 (x) by 4 (y) threads, grid (x tiles, y tiles, nzg).
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_prep(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V, const S *__restrict__ W,
           const double *__restrict__ ubar, double *__restrict__ Q, int64_t nC) {
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny;
    const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 1 + blockIdx.z;
    if (i > nxg || j > nyg - 1) return;
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
// Filtered velocities (subgrid.f90:428-443) on the full grid of one of U, V,
// W (COMP 0, 1, 2).  Each point takes the filter at its periodic image; all
// rows, one-sided y weights on the first and last.  Block 64 (x) by 4 (y),
// grid (x tiles, y tiles, n3).
// ---------------------------------------------------------------------------
template <typename S, int COMP>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_filter_vel(Dims d, const S *__restrict__ F, const double *__restrict__ ubar, double *__restrict__ Fo) {
    const int n1 = COMP == 0 ? d.nx : d.nxg, n2 = COMP == 1 ? d.ny : d.nyg, n3 = COMP == 2 ? d.nz : d.nzg;
    const int I = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = 1 + blockIdx.z;
    if (I > n1 || J > n2) return;
    const int ii = COMP == 0 ? img_f(I, n1) : img_c(I, n1);
    const int kk = COMP == 2 ? img_f(K, n3) : img_c(K, n3);
    auto f = [&](int a, int b, int c) {
        return (double)F[i3(n1, n2, a, b, c)] + ((COMP == 0 && ubar) ? ubar[b] : 0.0);
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
    Fo[i3(n1, n2, I, J, K)] = acc;
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// tile of k_dsm_lm: 64 columns in x by 4 in z, with a one-cell halo
constexpr int LM_TX = 64, LM_TZ = 4, LM_HX = LM_TX + 2, LM_HZ = LM_TZ + 2, LM_NP = LM_HX * LM_HZ;
constexpr int LM_NQ = 9;  // staged quantities: centre velocities and |S|Sij

// ---------------------------------------------------------------------------
// Lij, Mij and their plane products.  Block: 64 columns in x by 4 in z (one
// wave per z column row); the block marches DSM_LM_ROWS rows (blockIdx.y
// selects the row chunk).  Each row of the nine staged quantities is loaded
// once per block into LDS with its halo; every thread forms its x-z filtered
// sums (products included) from LDS and keeps the sums of the previous two
// rows for the y combination.
// part[(cw * nyp + j - 2) * 2 + {0, 1}] = wave sums of LM, MM, with
// cw = (4 bz + wave) * tiles_x + bx.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(DSM_WG) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_dsm_lm(Dims d, Metrics m, const double *__restrict__ Q, const double *__restrict__ Fu, double *__restrict__ part,
         int tiles_x) {
    __shared__ double sq[LM_NQ][LM_NP];
    const int nxg = d.nxg, nyg = d.nyg, nzg = d.nzg, nx = d.nx, ny = d.ny;
    const int64_t nC = (int64_t)nxg * nyg * nzg;
    const int bx = blockIdx.x % tiles_x, bz = blockIdx.x / tiles_x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = 2 + bx * LM_TX + lane, k = 2 + bz * LM_TZ + wv;
    const int ih = 1 + bx * LM_TX, kh = 1 + bz * LM_TZ;  // halo origin
    const bool act = i <= nxg - 2 && k <= nzg - 2;
    const int cw = (LM_TZ * bz + wv) * tiles_x + bx;
    const int nyp = nyg - 2;
    const int j0 = 2 + blockIdx.y * DSM_LM_ROWS;
    const int j1 = min(j0 + DSM_LM_ROWS - 1, nyg - 1);
    const double wgt = (i == 2 ? 2.0 : 1.0) * (k == 2 ? 2.0 : 1.0);
    const double *Uf = Fu, *Vf = Fu + (int64_t)nx * nyg * nzg, *Wf = Vf + (int64_t)nxg * ny * nzg;
    auto uf = [&](int I, int J, int K) { return Uf[i3(nx, nyg, I, J, K)]; };
    auto vf = [&](int I, int J, int K) { return Vf[i3(nxg, ny, I, J, K)]; };
    auto wf = [&](int I, int J, int K) { return Wf[i3(nxg, nyg, I, J, K)]; };
    // stage row r (block-uniform) into LDS; points outside the arrays read 0
    // and are never used by an active thread
    auto stage = [&](int r) {
        __syncthreads();
        for (int p = threadIdx.x; p < LM_NQ * LM_NP; p += DSM_WG) {
            const int qa = p / LM_NP, pt = p - qa * LM_NP;
            const int pz = pt / LM_HX, px = pt - pz * LM_HX;
            const int I = ih + px, K = kh + pz;
            sq[qa][pt] = (I <= nxg && K <= nzg) ? Q[qa * nC + i3(nxg, nyg, I, r, K)] : 0.0;
        }
        __syncthreads();
    };
    // x-z filtered sums of the staged row: 3 velocities, 6 products
    // (uu vv ww uv uw vw), 6 |S|Sij
    auto rowsum = [&](double F[15]) {
#pragma unroll
        for (int c = 0; c < 15; c++) F[c] = 0.0;
#pragma unroll
        for (int a = -1; a <= 1; a++)
#pragma unroll
            for (int c = -1; c <= 1; c++) {
                const double f = fil_xz(a, c);
                const int pt = (wv + 1 + c) * LM_HX + lane + 1 + a;
                const double u = sq[0][pt], v = sq[1][pt], w = sq[2][pt];
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
                for (int s = 0; s < 6; s++) F[9 + s] += f * sq[3 + s][pt];
            }
    };
    double Fm[15], F0[15], Fp[15];
#pragma unroll
    for (int c = 0; c < 15; c++) { Fm[c] = 0.0; Fp[c] = 0.0; }
    if (j0 > 2) {
        stage(j0 - 1);
        rowsum(Fm);
    }
    stage(j0);
    rowsum(F0);
    for (int j = j0; j <= j1; j++) {
        if (j < nyg - 1) {
            stage(j + 1);
            rowsum(Fp);
        }
        double lm = 0.0, mm = 0.0;
        if (act) {
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
        }
#pragma unroll
        for (int c = 0; c < 15; c++) { Fm[c] = F0[c]; F0[c] = Fp[c]; }
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
    const int nxg = d.nxg, nyg = d.nyg;
    const int64_t nC = (int64_t)nxg * nyg * d.nzg;
    const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 1 + blockIdx.z;
    if (i > nxg || j > nyg) return;
    const int jj = j == 1 ? 2 : (j == nyg ? nyg - 1 : j);
    const double sgn = (jj != j && dirichlet) ? -1.0 : 1.0;
    const double val = fmax(-0.5 * avg[jj - 1] * Q[9 * nC + i3(nxg, nyg, i, jj, k)], 0.0);
    NT[i3(nxg, nyg, i, j, k)] = (S)(sgn * val);
    if (i == 1 && j == 1 && k == 1) {
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
    auto g3 = [](int n1, int n2, int n3) { return dim3((n1 + 63) / 64, (n2 + 3) / 4, n3); };
    hipLaunchKernelGGL((k_dsm_prep<S>), g3(d.nxg, d.nyg - 2, d.nzg), dim3(DSM_WG), 0, st, d, m, U, V, W, ubar, w.Q,
                       nC);
    double *Uf = w.Fu, *Vf = Uf + (int64_t)d.nx * d.nyg * d.nzg, *Wf = Vf + (int64_t)d.nxg * d.ny * d.nzg;
    hipLaunchKernelGGL((k_dsm_filter_vel<S, 0>), g3(d.nx, d.nyg, d.nzg), dim3(DSM_WG), 0, st, d, U, ubar, Uf);
    hipLaunchKernelGGL((k_dsm_filter_vel<S, 1>), g3(d.nxg, d.ny, d.nzg), dim3(DSM_WG), 0, st, d, V, ubar, Vf);
    hipLaunchKernelGGL((k_dsm_filter_vel<S, 2>), g3(d.nxg, d.nyg, d.nz), dim3(DSM_WG), 0, st, d, W, ubar, Wf);
    const int tx = dsm_tiles_x(d), tz = dsm_tiles_z(d);
    const int nyp = d.nyg - 2;
    const int chunks = (nyp + DSM_LM_ROWS - 1) / DSM_LM_ROWS;
    hipLaunchKernelGGL(k_dsm_lm, dim3(tx * tz, chunks), dim3(DSM_WG), 0, st, d, m, w.Q, w.Fu, w.part, tx);
    hipLaunchKernelGGL(k_dsm_avg, dim3(nyp), dim3(DSM_WG), 0, st, tx * tz * 4, nyp, w.part, w.avg);
    hipLaunchKernelGGL((k_dsm_nut<S>), g3(d.nxg, d.nyg, d.nzg), dim3(DSM_WG), 0, st, d, w.Q, w.avg, NT, dirichlet);
}

template void launch_dsm<double>(const Dims &, const Metrics &, const double *, const double *, const double *,
                                 const double *, const DsmWork &, double *, int, hipStream_t);
template void launch_dsm<float>(const Dims &, const Metrics &, const float *, const float *, const float *,
                                const double *, const DsmWork &, float *, int, hipStream_t);

}  // namespace cf
