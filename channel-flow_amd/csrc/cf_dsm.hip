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
//                    their periodic images)                     -> Q[0..9]
//   k_dsm_filter_vel filtered U, V, W on their own grids, all rows  -> Fu
//   k_dsm_fstrain    4 |hat S| hat(S)ij from Fu                 -> Q[10..15]
//   k_dsm_lm         per cell Lij and Mij from the LDS-staged rows of
//                    Q[0..8]; LM and MM summed over each wave's 64 cells of
//                    a row                                      -> part
//   k_dsm_avg        one block per row: fixed-order sum of the wave partials,
//                    C(y) = LM/MM                               -> avg
//   k_dsm_nut        nu_t at every centre, wall ghosts, avg ghosts
//
// The reference sums the plane over centres 2..nxg-1 x 2..nzg-1, which holds
// the periodic images nxg-1 (of 2) and nzg-1 (of 2): cells in the first
// unique column or plane count twice (four times on both).
#include <utility>

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
// Block 64 (x) by 4 (y) threads over tiles (x, y, k), k fastest, so that
// the blocks of one XCD sweep neighbouring k planes of one (x, y) tile.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_prep(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V, const S *__restrict__ W,
           const double *__restrict__ ubar, double *__restrict__ Q, int64_t nC, int tiles_x, int ntiles) {
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny;
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int k = 1 + t % d.nzg, r = t / d.nzg;
    const int i = 1 + (r % tiles_x) * 64 + (threadIdx.x & 63);
    const int j = 2 + (r / tiles_x) * 4 + (threadIdx.x >> 6);
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
// W (COMP 0, 1, 2), all rows, one-sided y weights on the first and last.
// The filter factors into (1/6 2/3 1/6) per axis: each thread owns one
// (x, y) point, forms the x-y filtered value of every plane it passes and
// marches FV_KCH unique z planes, so each output costs nine loads.  Ghost
// columns and planes take the values of their periodic images.  Tiles
// (x, y, k chunk), k chunk fastest.
// ---------------------------------------------------------------------------
constexpr int FV_KCH = 16;

template <typename S, int COMP>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_filter_vel(Dims d, const S *__restrict__ F, const double *__restrict__ ubar, double *__restrict__ Fo,
                 int tiles_x, int nchunks, int ntiles) {
    const int n1 = COMP == 0 ? d.nx : d.nxg, n2 = COMP == 1 ? d.ny : d.nyg, n3 = COMP == 2 ? d.nz : d.nzg;
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int ch = t % nchunks, r = t / nchunks;
    const int I = 1 + (r % tiles_x) * 64 + (threadIdx.x & 63);
    const int J = 1 + (r / tiles_x) * 4 + (threadIdx.x >> 6);
    if (I > n1 || J > n2) return;
    const int ii = COMP == 0 ? img_f(I, n1) : img_c(I, n1);
    const int nu = COMP == 2 ? n3 - 1 : n3 - 2;  // last unique plane
    const int k0 = 2 + ch * FV_KCH, k1 = min(k0 + FV_KCH - 1, nu);
    const double wm = J == 1 ? 0.0 : (J == n2 ? 1.0 / 3.0 : 1.0 / 6.0);
    const double wp = J == n2 ? 0.0 : (J == 1 ? 1.0 / 3.0 : 1.0 / 6.0);
    const int jm = J > 1 ? J - 1 : J, jp = J < n2 ? J + 1 : J;
    auto f = [&](int a, int b, int c) {
        return (double)F[i3(n1, n2, a, b, c)] + ((COMP == 0 && ubar) ? ubar[b] : 0.0);
    };
    auto X = [&](int K) {
        double acc = 0.0;
#pragma unroll
        for (int a = -1; a <= 1; a++) {
            const double y = wm * f(ii + a, jm, K) + 2.0 / 3.0 * f(ii + a, J, K) + wp * f(ii + a, jp, K);
            acc = acc + (a == 0 ? 2.0 / 3.0 : 1.0 / 6.0) * y;
        }
        return acc;
    };
    double xm = X(k0 - 1), x0 = X(k0);
    for (int K = k0; K <= k1; K++) {
        const double xp = X(K + 1);
        const double out = 1.0 / 6.0 * xm + 2.0 / 3.0 * x0 + 1.0 / 6.0 * xp;
        Fo[i3(n1, n2, I, J, K)] = out;
        if (COMP == 2) {  // faces: 1 <- n3-1, n3 <- 2
            if (K == 2) Fo[i3(n1, n2, I, J, n3)] = out;
            if (K == n3 - 1) Fo[i3(n1, n2, I, J, 1)] = out;
        } else {  // centres: 1 <- n3-2, n3-1 <- 2, n3 <- 3
            if (K == 2) Fo[i3(n1, n2, I, J, n3 - 1)] = out;
            if (K == 3) Fo[i3(n1, n2, I, J, n3)] = out;
            if (K == n3 - 2) Fo[i3(n1, n2, I, J, 1)] = out;
        }
        xm = x0;
        x0 = xp;
    }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// ---------------------------------------------------------------------------
// Strain of the filtered field (compute_Sij of Uf, Vf, Wf, subgrid.f90:446)
// at the unique interior centres: T[c] = 4 |hat S| hat(S)c, the first term
// of Mij (fil_size**2 = 4), stored in Q arrays 10..15.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(DSM_WG)
k_dsm_fstrain(Dims d, Metrics m, const double *__restrict__ Fu, double *__restrict__ Q, int64_t nC, int tiles_x,
              int ntiles) {
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny, nzp = d.nzg - 3;
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int k = 2 + t % nzp, r = t / nzp;
    const int i = 2 + (r % tiles_x) * 64 + (threadIdx.x & 63);
    const int j = 2 + (r / tiles_x) * 4 + (threadIdx.x >> 6);
    if (i > nxg - 2 || j > nyg - 1) return;
    const double *Uf = Fu, *Vf = Fu + (int64_t)nx * nyg * d.nzg, *Wf = Vf + (int64_t)nxg * ny * d.nzg;
    auto uf = [&](int I, int J, int K) { return Uf[i3(nx, nyg, I, J, K)]; };
    auto vf = [&](int I, int J, int K) { return Vf[i3(nxg, ny, I, J, K)]; };
    auto wf = [&](int I, int J, int K) { return Wf[i3(nxg, nyg, I, J, K)]; };
    double sf[6];
    const double smf = strain6(m, i, j, k, uf, vf, wf, sf);
    const int64_t q = i3(nxg, nyg, i, j, k);
#pragma unroll
    for (int c = 0; c < 6; c++) Q[(10 + c) * nC + q] = 4.0 * smf * sf[c];
}

// tile of k_dsm_lm: 64 columns in x by 4 in z, with a one-cell halo
constexpr int LM_TX = 64, LM_TZ = 4, LM_HX = LM_TX + 2, LM_HZ = LM_TZ + 2, LM_NP = LM_HX * LM_HZ;
constexpr int LM_NQ = 9;  // staged quantities: centre velocities and |S|Sij

// ---------------------------------------------------------------------------
// Lij, Mij and their plane products.  Block: 64 columns in x by 4 in z (one
// wave per z column row); the block marches DSM_LM_ROWS rows of one row
// chunk.  Each row of the nine staged quantities is loaded once per block
// into LDS with its halo; the global loads of the next row are issued into
// registers before the current row is computed and written to LDS after it.
// Every thread forms its x-z filtered sums (products included) from LDS and
// folds them into the streaming y combination below.
// part[(cw * nyp + j - 2) * 2 + {0, 1}] = wave sums of LM, MM, with
// cw = (4 bz + wave) * tiles_x + bx.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(DSM_WG)
k_dsm_lm(Dims d, const double *__restrict__ Q, double *__restrict__ part, int tiles_x, int tiles_z, int nchunks) {
    __shared__ double sq[LM_NQ][LM_NP];
    const int nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    const int64_t nC = (int64_t)nxg * nyg * nzg;
    // tiles (z, row chunk, x), z fastest: neighbouring z tiles share halo planes
    const int t = xcd_remap(blockIdx.x, tiles_x * tiles_z * nchunks);
    const int bz = t % tiles_z, ch = (t / tiles_z) % nchunks, bx = t / (tiles_z * nchunks);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i = 2 + bx * LM_TX + lane, k = 2 + bz * LM_TZ + wv;
    const int ih = 1 + bx * LM_TX, kh = 1 + bz * LM_TZ;  // halo origin
    const bool act = i <= nxg - 2 && k <= nzg - 2;
    const int cw = (LM_TZ * bz + wv) * tiles_x + bx;
    const int nyp = nyg - 2;
    const int j0 = 2 + ch * DSM_LM_ROWS;
    const int j1 = min(j0 + DSM_LM_ROWS - 1, nyg - 1);
    const double wgt = (i == 2 ? 2.0 : 1.0) * (k == 2 ? 2.0 : 1.0);
    // stage row r (block-uniform) into LDS; points outside the arrays read 0
    // and are never used by an active thread
    constexpr int NPF = (LM_NQ * LM_NP + DSM_WG - 1) / DSM_WG;
    double pf[NPF];
    // global loads of row r into registers (block-uniform r)
    auto fetch = [&](int r) {
#pragma unroll
        for (int u = 0; u < NPF; u++) {
            const int p = threadIdx.x + u * DSM_WG;
            pf[u] = 0.0;
            if (p < LM_NQ * LM_NP) {
                const int qa = p / LM_NP, pt = p - qa * LM_NP;
                const int pz = pt / LM_HX, px = pt - pz * LM_HX;
                const int I = ih + px, K = kh + pz;
                if (I <= nxg && K <= nzg) pf[u] = Q[qa * nC + i3(nxg, nyg, I, r, K)];
            }
        }
    };
    auto commit = [&]() {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NPF; u++) {
            const int p = threadIdx.x + u * DSM_WG;
            if (p < LM_NQ * LM_NP) {
                const int qa = p / LM_NP, pt = p - qa * LM_NP;
                sq[qa][pt] = pf[u];
            }
        }
        __syncthreads();
    };
    // Streaming y combination: hat(j) = am(j) F(j-1) + 2/3 F(j) + ap(j) F(j+1)
    // with F the x-z filtered row sums.  P1 holds the partial hat of the row
    // below the staged one, P2 that of the staged row; when row r is staged,
    // each component of F(r) completes hat(r-1), which is consumed at once,
    // component by component, so only the two partial vectors stay live.
    // Component order: u v w, then per pair (uu vv ww uv uw vw) the product
    // and the matching |S|Sij.
    auto am = [&](int jj) { return jj == 2 ? 0.0 : (jj == nyg - 1 ? 1.0 / 3.0 : 1.0 / 6.0); };
    auto ap = [&](int jj) { return jj == nyg - 1 ? 0.0 : (jj == 2 ? 1.0 / 3.0 : 1.0 / 6.0); };
    double P1[15], P2[15];
#pragma unroll
    for (int c = 0; c < 15; c++) { P1[c] = 0.0; P2[c] = 0.0; }
    const int rs = j0 > 2 ? j0 - 1 : j0;
    const int rlast = min(j1 + 1, nyg - 1);  // last row to stage
    fetch(rs);
    commit();
    for (int r = rs; r <= j1 + 1; r++) {
        const bool have = r <= nyg - 1;  // block-uniform
        if (r + 1 <= rlast) fetch(r + 1);
        const int jd = r - 1;            // row completed by this one
        const bool use = act && jd >= j0 && jd <= j1;
        const double wap = use ? ap(jd) : 0.0, wam = am(r + 1);
        double T[6];
        const int64_t qd = i3(nxg, nyg, i, jd, k);
#pragma unroll
        for (int c = 0; c < 6; c++) T[c] = use ? Q[(10 + c) * nC + qd] : 0.0;
        double hv[3], lm = 0.0, mm = 0.0;
        const int pc = (wv + 1) * LM_HX + lane + 1;  // own point in the staged tile
        auto nb = [&](int e, int q) { return sq[e][pc + (q % 3 - 1) * LM_HX + q / 3 - 1]; };
        // component c of F(r): sum over the 3x3 neighbours of fil * g(q)
        auto fsum = [&](auto g) {
            double acc = 0.0;
            if (have && act) {
#pragma unroll
                for (int q = 0; q < 9; q++) acc += fil_xz(q / 3 - 1, q % 3 - 1) * g(q);
            }
            return acc;
        };
        auto push = [&](int c, double f) {  // fold F_c(r) in; returns the completed hat(r-1)_c
            const double done = P1[c] + wap * f;
            P1[c] = P2[c] + 2.0 / 3.0 * f;
            P2[c] = wam * f;
            return done;
        };
        hv[0] = push(0, fsum([&](int q) { return nb(0, q); }));
        hv[1] = push(1, fsum([&](int q) { return nb(1, q); }));
        hv[2] = push(2, fsum([&](int q) { return nb(2, q); }));
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const int ia = c == 0 || c == 3 || c == 4 ? 0 : (c == 1 || c == 5 ? 1 : 2);
            const int ib = c == 0 ? 0 : (c == 1 || c == 3 ? 1 : 2);
            const double hp = push(3 + c, fsum([&](int q) { return nb(ia, q) * nb(ib, q); }));
            const double hs = push(9 + c, fsum([&](int q) { return nb(3 + c, q); }));
            const double L = hp - hv[ia] * hv[ib];
            const double M = T[c] - hs;
            const double wc = c < 3 ? 1.0 : 2.0;
            lm += wc * (L * M);
            mm += wc * (M * M);
        }
        if (r + 1 <= rlast) commit();
        if (jd >= j0 && jd <= j1) {  // block-uniform
            lm = use ? lm * wgt : 0.0;
            mm = use ? mm * wgt : 0.0;
            lm = wave_sum(lm);
            mm = wave_sum(mm);
            if (lane == 0) {
                const int64_t p = ((int64_t)cw * nyp + (jd - 2)) * 2;
                part[p] = lm;
                part[p + 1] = mm;
            }
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
k_dsm_nut(Dims d, const double *__restrict__ Q, double *__restrict__ avg, S *__restrict__ NT, int dirichlet,
          int tiles_x, int ntiles) {
    const int nxg = d.nxg, nyg = d.nyg;
    const int64_t nC = (int64_t)nxg * nyg * d.nzg;
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int k = 1 + t % d.nzg, r = t / d.nzg;
    const int i = 1 + (r % tiles_x) * 64 + (threadIdx.x & 63);
    const int j = 1 + (r / tiles_x) * 4 + (threadIdx.x >> 6);
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
    const int txc = (d.nxg + 63) / 64;
    {
        const int nt = txc * ((d.nyg - 2 + 3) / 4) * d.nzg;
        hipLaunchKernelGGL((k_dsm_prep<S>), dim3(nt), dim3(DSM_WG), 0, st, d, m, U, V, W, ubar, w.Q, nC, txc, nt);
    }
    double *Uf = w.Fu, *Vf = Uf + (int64_t)d.nx * d.nyg * d.nzg, *Wf = Vf + (int64_t)d.nxg * d.ny * d.nzg;
    const int nkc = (d.nzp + FV_KCH - 1) / FV_KCH;
    auto fv = [&](int n1, int n2) {
        const int tx = (n1 + 63) / 64, nt = tx * ((n2 + 3) / 4) * nkc;
        return std::make_pair(tx, nt);
    };
    {
        auto g = fv(d.nx, d.nyg);
        hipLaunchKernelGGL((k_dsm_filter_vel<S, 0>), dim3(g.second), dim3(DSM_WG), 0, st, d, U, ubar, Uf, g.first,
                           nkc, g.second);
        g = fv(d.nxg, d.ny);
        hipLaunchKernelGGL((k_dsm_filter_vel<S, 1>), dim3(g.second), dim3(DSM_WG), 0, st, d, V, ubar, Vf, g.first,
                           nkc, g.second);
        g = fv(d.nxg, d.nyg);
        hipLaunchKernelGGL((k_dsm_filter_vel<S, 2>), dim3(g.second), dim3(DSM_WG), 0, st, d, W, ubar, Wf, g.first,
                           nkc, g.second);
    }
    {
        const int tx = (d.nxg - 3 + 63) / 64, nt = tx * ((d.nyg - 2 + 3) / 4) * (d.nzg - 3);
        hipLaunchKernelGGL(k_dsm_fstrain, dim3(nt), dim3(DSM_WG), 0, st, d, m, w.Fu, w.Q, nC, tx, nt);
    }
    const int tx = dsm_tiles_x(d), tz = dsm_tiles_z(d);
    const int nyp = d.nyg - 2;
    const int chunks = (nyp + DSM_LM_ROWS - 1) / DSM_LM_ROWS;
    hipLaunchKernelGGL(k_dsm_lm, dim3(tx * tz * chunks), dim3(DSM_WG), 0, st, d, w.Q, w.part, tx, tz, chunks);
    hipLaunchKernelGGL(k_dsm_avg, dim3(nyp), dim3(DSM_WG), 0, st, tx * tz * 4, nyp, w.part, w.avg);
    {
        const int nt = txc * ((d.nyg + 3) / 4) * d.nzg;
        hipLaunchKernelGGL((k_dsm_nut<S>), dim3(nt), dim3(DSM_WG), 0, st, d, w.Q, w.avg, NT, dirichlet, txc, nt);
    }
}

template void launch_dsm<double>(const Dims &, const Metrics &, const double *, const double *, const double *,
                                 const double *, const DsmWork &, double *, int, hipStream_t);
template void launch_dsm<float>(const Dims &, const Metrics &, const float *, const float *, const float *,
                                const double *, const DsmWork &, float *, int, hipStream_t);

}  // namespace cf
