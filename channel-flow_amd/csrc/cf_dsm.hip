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
//   k_dsm_strain       centre velocities, |S| and |S|Sij of the field at
//                      every centre (ghost columns and planes are copies of
//                      their periodic images)              -> Q[0..9]
//   k_dsm_filter_vel   filtered U, V, W on their own grids, all rows -> Fu
//   k_dsm_lm           per cell Lij and Mij from the LDS-staged rows of
//                      Q[0..8] and of Fu (4 |hat S| hat(S)ij formed in
//                      place); LM and MM summed over each wave's 64 cells
//                      of a row                            -> part
//   k_dsm_avg          one block per row: fixed-order sum of the wave
//                      partials, C(y) = LM/MM              -> avg
//   k_dsm_nut          nu_t at every centre, wall ghosts, avg ghosts
//
// The reference sums the plane over centres 2..nxg-1 x 2..nzg-1, which holds
// the periodic images nxg-1 (of 2) and nzg-1 (of 2): cells in the first
// unique column or plane count twice (four times on both).
#include <algorithm>

#include "cf_kernels.hpp"

namespace cf {

namespace {

constexpr int DSM_WG = 256;
// The marching kernels walk chunks of up to *_CHUNK planes or rows (strain
// and filter marches in z, the LM march in y); on small grids the chunks
// shorten until there are DSM_MIN_BLOCKS workgroups (down to 2 rows or
// planes per chunk).
constexpr int SK_CHUNK = 32, FV_CHUNK = 16, LM_CHUNK = 32, DSM_MIN_BLOCKS = 512;

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
// Filtered velocities (subgrid.f90:428-443) on the full grid of one of U, V,
// W (COMP 0, 1, 2), all rows, one-sided y weights on the first and last.
// The filter factors into (1/6 2/3 1/6) per axis.  A block owns 64 unique
// columns by 4 rows and marches a chunk of kch unique z planes: each plane's
// 66 x 6 tile (one-cell x-y halo, actual neighbours) is staged in LDS with
// the next plane's loads in flight; each thread forms the x-y filtered value
// of its point from LDS and combines three planes in z in registers.  The
// value of a unique column or plane is also written at its periodic ghost
// copies.  Tiles (x, y, k chunk), k chunk fastest.
// ---------------------------------------------------------------------------
constexpr int FV_HX = 66, FV_HY = 6, FV_NP = FV_HX * FV_HY;

template <typename S, int COMP>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_filter_vel(Dims d, const S *__restrict__ F, const double *__restrict__ ubar, double *__restrict__ Fo,
                 int tiles_x, int nchunks, int ntiles, int kch) {
    __shared__ double tile[2][FV_NP];  // planes K and K+1 (slot K & 1)
    const int n1 = COMP == 0 ? d.nx : d.nxg, n2 = COMP == 1 ? d.ny : d.nyg, n3 = COMP == 2 ? d.nz : d.nzg;
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int ch = t % nchunks, r = t / nchunks;
    const int bx = r % tiles_x, by = r / tiles_x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ih = 1 + bx * 64, jh = by * 4;  // halo origin (column, row)
    const int I = ih + 1 + lane, J = jh + 1 + wv;
    const int iu = COMP == 0 ? n1 - 1 : n1 - 2;  // last unique column
    const bool act = I <= iu && J <= n2;
    const int nu = COMP == 2 ? n3 - 1 : n3 - 2;  // last unique plane
    const int k0 = 2 + ch * kch, k1 = min(k0 + kch - 1, nu);
    const double wm = J == 1 ? 0.0 : (J == n2 ? 1.0 / 3.0 : 1.0 / 6.0);
    const double wp = J == n2 ? 0.0 : (J == 1 ? 1.0 / 3.0 : 1.0 / 6.0);
    const int rm = (J > 1 ? J - 1 : J) - jh, r0 = J - jh, rp = (J < n2 ? J + 1 : J) - jh;  // tile rows
    constexpr int NPF = (FV_NP + DSM_WG - 1) / DSM_WG;
    double pf[NPF];
    auto fetch = [&](int K) {
#pragma unroll
        for (int u = 0; u < NPF; u++) {
            const int p = threadIdx.x + u * DSM_WG;
            pf[u] = 0.0;
            if (p < FV_NP) {
                const int py = p / FV_HX, px = p - py * FV_HX;
                const int Ia = ih + px, Ja = jh + py;
                if (Ia <= n1 && Ja >= 1 && Ja <= n2)
                    pf[u] = (double)F[i3(n1, n2, Ia, Ja, K)] + ((COMP == 0 && ubar) ? ubar[Ja] : 0.0);
            }
        }
    };
    // both slots are read between two barriers at most once each, so one
    // barrier after each write suffices
    auto commit = [&](int K) {
#pragma unroll
        for (int u = 0; u < NPF; u++) {
            const int p = threadIdx.x + u * DSM_WG;
            if (p < FV_NP) tile[K & 1][p] = pf[u];
        }
        __syncthreads();
    };
    auto X = [&](int K) {
        const double *tp = tile[K & 1];
        double acc = 0.0;
#pragma unroll
        for (int a = -1; a <= 1; a++) {
            const int c = lane + 1 + a;
            const double y = wm * tp[rm * FV_HX + c] + 2.0 / 3.0 * tp[r0 * FV_HX + c] + wp * tp[rp * FV_HX + c];
            acc = acc + (a == 0 ? 2.0 / 3.0 : 1.0 / 6.0) * y;
        }
        return acc;
    };
    // the column and its periodic ghost copy (faces: 1 <- n1-1, n1 <- 2;
    // centres: 1 <- n1-2, n1-1 <- 2, n1 <- 3)
    const int ig = COMP == 0 ? (I == 2 ? n1 : (I == n1 - 1 ? 1 : 0))
                             : (I == 2 ? n1 - 1 : (I == 3 ? n1 : (I == n1 - 2 ? 1 : 0)));
    fetch(k0 - 1);
    commit(k0 - 1);
    fetch(k0);
    double xm = act ? X(k0 - 1) : 0.0;
    commit(k0);
    fetch(k0 + 1);
    double x0 = act ? X(k0) : 0.0;
    for (int K = k0; K <= k1; K++) {
        commit(K + 1);
        if (K + 2 <= k1 + 1) fetch(K + 2);
        if (act) {
            const double xp = X(K + 1);
            const double out = 1.0 / 6.0 * xm + 2.0 / 3.0 * x0 + 1.0 / 6.0 * xp;
            // the plane and its periodic ghost copy (faces: 1 <- n3-1,
            // n3 <- 2; centres: 1 <- n3-2, n3-1 <- 2, n3 <- 3)
            const int kg = COMP == 2 ? (K == 2 ? n3 : (K == n3 - 1 ? 1 : 0))
                                     : (K == 2 ? n3 - 1 : (K == 3 ? n3 : (K == n3 - 2 ? 1 : 0)));
            Fo[i3(n1, n2, I, J, K)] = out;
            if (ig) Fo[i3(n1, n2, ig, J, K)] = out;
            if (kg) {
                Fo[i3(n1, n2, I, J, kg)] = out;
                if (ig) Fo[i3(n1, n2, ig, J, kg)] = out;
            }
            xm = x0;
            x0 = xp;
        }
    }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// ---------------------------------------------------------------------------
// compute_Sij (subgrid.f90:891-949) on a 2.5-D march: a block owns 64 x 4
// unique centres in (x, y) and walks a chunk of `planes` z planes; U, V, W
// (FP32 deviations + ubar, or FP64) sit in an LDS ring of three planes (z-1,
// z, z+1) with a one-cell x-y halo, the next plane's loads issued into
// registers while the current plane is computed.  Every input value is read
// from HBM about once (plus the halo).  Centre velocities, |S|Sij and |S| go
// to Q[0..9] at the centre and at its periodic ghost copies (columns 1,
// nxg-1, nxg and planes 1, nzg-1, nzg).
// Tiles (x, y, z chunk), z chunk fastest.
// Q layout: 10 centre arrays (nxg, nyg, nzg), rows 2..nyg-1 written:
//   0..2 centre velocities (interpolate_x/y/z, subgrid.f90:391-393)
//   3..8 |S| Sij, 9 |S|
// ---------------------------------------------------------------------------
constexpr int SK_HX = 66, SK_HY = 6, SK_NP = SK_HX * SK_HY;

template <typename T>
__global__ void __launch_bounds__(DSM_WG)
k_dsm_strain(Dims d, Metrics m, const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
             const double *__restrict__ ubar, double *__restrict__ Q, int64_t nC, int tiles_x, int nchunks,
             int ntiles, int planes) {
    __shared__ double ring[3][3][SK_NP];  // [array][plane slot][point]
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny, nz = d.nz, nzg = d.nzg;
    const int t = xcd_remap(blockIdx.x, ntiles);
    const int ch = t % nchunks, r = t / nchunks;
    const int bx = r % tiles_x, by = r / tiles_x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ih = 1 + bx * 64, jh = 1 + by * 4;  // halo origin
    const int i = ih + 1 + lane, j = jh + 1 + wv;
    const bool act = i <= nxg - 2 && j <= nyg - 1;
    const int k0 = 2 + ch * planes, k1 = min(k0 + planes - 1, nzg - 2);
    // cooperative load of plane K of the three arrays: 3 * SK_NP values
    constexpr int NPF = (3 * SK_NP + DSM_WG - 1) / DSM_WG;
    double pf[NPF];
    auto fetch = [&](int K) {
#pragma unroll
        for (int u = 0; u < NPF; u++) {
            const int p = threadIdx.x + u * DSM_WG;
            pf[u] = 0.0;
            if (p < 3 * SK_NP) {
                const int a = p / SK_NP, pt = p - a * SK_NP;
                const int py = pt / SK_HX, px = pt - py * SK_HX;
                const int I = ih + px, J = jh + py;
                if (a == 0) {
                    if (I <= nx && J <= nyg && K <= nzg)
                        pf[u] = (double)U[i3(nx, nyg, I, J, K)] + (ubar ? ubar[J] : 0.0);
                } else if (a == 1) {
                    if (I <= nxg && J <= ny && K <= nzg) pf[u] = (double)V[i3(nxg, ny, I, J, K)];
                } else {
                    if (I <= nxg && J <= nyg && K <= nz) pf[u] = (double)W[i3(nxg, nyg, I, J, K)];
                }
            }
        }
    };
    auto commit = [&](int K) {
        __syncthreads();
        const int sl = K % 3;
#pragma unroll
        for (int u = 0; u < NPF; u++) {
            const int p = threadIdx.x + u * DSM_WG;
            if (p < 3 * SK_NP) {
                const int a = p / SK_NP, pt = p - a * SK_NP;
                ring[a][sl][pt] = pf[u];
            }
        }
        __syncthreads();
    };
    auto at = [&](int a, int I, int J, int K) { return ring[a][K % 3][(J - jh) * SK_HX + (I - ih)]; };
    auto u = [&](int I, int J, int K) { return at(0, I, J, K); };
    auto v = [&](int I, int J, int K) { return at(1, I, J, K); };
    auto w = [&](int I, int J, int K) { return at(2, I, J, K); };
    fetch(k0 - 1);
    commit(k0 - 1);
    fetch(k0);
    commit(k0);
    fetch(k0 + 1);
    for (int k = k0; k <= k1; k++) {
        commit(k + 1);  // ring holds k-1, k, k+1
        if (k + 2 <= k1 + 1) fetch(k + 2);
        if (act) {
            double s[6];
            const double sm = strain6(m, i, j, k, u, v, w, s);
            const double cu = 0.5 * (u(i - 1, j, k) + u(i, j, k));
            const double cv = 0.5 * (v(i, j - 1, k) + v(i, j, k));
            const double cw = 0.5 * (w(i, j, k - 1) + w(i, j, k));
            double val[10] = {cu, cv, cw, sm * s[0], sm * s[1], sm * s[2], sm * s[3], sm * s[4], sm * s[5], sm};
            {
                const int64_t q = i3(nxg, nyg, i, j, k);
#pragma unroll
                for (int e = 0; e < 10; e++) Q[e * nC + q] = val[e];
            }
            // periodic ghost copies of the centre (rare: three columns and
            // three planes)
            const int ig = i == 2 ? nxg - 1 : (i == 3 ? nxg : (i == nxg - 2 ? 1 : 0));
            const int kg = k == 2 ? nzg - 1 : (k == 3 ? nzg : (k == nzg - 2 ? 1 : 0));
            if (ig | kg) {
#pragma unroll 1
                for (int g = 0; g < 3; g++) {
                    const int I = g == 1 ? i : ig, K = g == 0 ? k : kg;
                    if (I == 0 || K == 0) continue;
                    double *q = Q + i3(nxg, nyg, I, j, K);
#pragma unroll
                    for (int e = 0; e < 10; e++) q[e * nC] = val[e];
                }
            }
        }
    }
}

// tile of k_dsm_lm: 64 columns in x by 4 in z, with a one-cell halo
constexpr int LM_TX = 64, LM_TZ = 4, LM_HX = LM_TX + 2, LM_HZ = LM_TZ + 2, LM_NP = LM_HX * LM_HZ;
constexpr int LM_NQ = 9;  // staged quantities: centre velocities and |S|Sij

// ---------------------------------------------------------------------------
// Lij, Mij and their plane products.  Block: 64 columns in x by 4 in z (one
// wave per z column row); the block marches the `rows` rows of one row
// chunk.  Each row of the nine staged quantities, and of the filtered U, V,
// W, is loaded once per block into LDS with its halo; the global loads of
// the next row are issued into registers before the current row is computed
// and written to LDS after it.  Every thread forms its x-z filtered sums
// (products included) from LDS and folds them into the streaming y
// combination below.  The filtered fields sit in a ring of three rows, from
// which compute_Sij gives 4 |hat S| hat(S)ij of the completed row in place
// (fil_size**2 = 4, subgrid.f90:449-451).
// part[(cw * nyp + j - 2) * 2 + {0, 1}] = wave sums of LM, MM, with
// cw = (4 bz + wave) * tiles_x + bx.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(DSM_WG)
k_dsm_lm(Dims d, Metrics m, const double *__restrict__ Q, const double *__restrict__ Uf,
         const double *__restrict__ Vf, const double *__restrict__ Wf, double *__restrict__ part, int tiles_x,
         int tiles_z, int nchunks, int rows) {
    __shared__ double sq[LM_NQ][LM_NP];
    __shared__ double fr[3][3][LM_NP];  // filtered U, V, W: ring of rows r-2, r-1, r
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
    const int j0 = 2 + ch * rows;
    const int j1 = min(j0 + rows - 1, nyg - 1);
    const double wgt = (i == 2 ? 2.0 : 1.0) * (k == 2 ? 2.0 : 1.0);
    // stage row r (block-uniform) into LDS; points outside the arrays read 0
    // and are never used by an active thread
    constexpr int NPF = (LM_NQ * LM_NP + DSM_WG - 1) / DSM_WG;
    constexpr int NPG = (3 * LM_NP + DSM_WG - 1) / DSM_WG;
    double pf[NPF], pg[NPG];
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    // global loads of row r into registers (block-uniform r and flags): the
    // staged quantities when wq, the filtered fields when wf
    auto fetch = [&](int r, bool wq, bool wf) {
        if (wq) {
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
        }
        if (wf) {
#pragma unroll
            for (int u = 0; u < NPG; u++) {
                const int p = threadIdx.x + u * DSM_WG;
                pg[u] = 0.0;
                if (p < 3 * LM_NP) {
                    const int a = p / LM_NP, pt = p - a * LM_NP;
                    const int pz = pt / LM_HX, px = pt - pz * LM_HX;
                    const int I = ih + px, K = kh + pz;
                    if (a == 0) {
                        if (I <= nx && K <= nzg) pg[u] = Uf[i3(nx, nyg, I, r, K)];
                    } else if (a == 1) {
                        if (I <= nxg && r <= ny && K <= nzg) pg[u] = Vf[i3(nxg, ny, I, r, K)];
                    } else {
                        if (I <= nxg && K <= nz) pg[u] = Wf[i3(nxg, nyg, I, r, K)];
                    }
                }
            }
        }
    };
    auto commit = [&](int r, bool wq, bool wf) {
        __syncthreads();
        if (wq) {
#pragma unroll
            for (int u = 0; u < NPF; u++) {
                const int p = threadIdx.x + u * DSM_WG;
                if (p < LM_NQ * LM_NP) {
                    const int qa = p / LM_NP, pt = p - qa * LM_NP;
                    sq[qa][pt] = pf[u];
                }
            }
        }
        if (wf) {
            const int sl = r % 3;
#pragma unroll
            for (int u = 0; u < NPG; u++) {
                const int p = threadIdx.x + u * DSM_WG;
                if (p < 3 * LM_NP) {
                    const int a = p / LM_NP, pt = p - a * LM_NP;
                    fr[a][sl][pt] = pg[u];
                }
            }
        }
        __syncthreads();
    };
    // filtered fields at (I, J, K) from the ring (J in r-2..r)
    auto fat = [&](int a, int I, int J, int K) { return fr[a][J % 3][(K - kh) * LM_HX + (I - ih)]; };
    auto uf = [&](int I, int J, int K) { return fat(0, I, J, K); };
    auto vf = [&](int I, int J, int K) { return fat(1, I, J, K); };
    auto wf_ = [&](int I, int J, int K) { return fat(2, I, J, K); };
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
    const int rlast = min(j1 + 1, nyg - 1);  // last row of Q to stage
    const int rflast = j1 + 1;               // last filtered row to stage (<= nyg)
    if (j0 == 2) {                           // filtered row 1 (the wall ghost row)
        fetch(1, false, true);
        commit(1, false, true);
    }
    fetch(rs, true, true);
    commit(rs, true, true);
    for (int r = rs; r <= j1 + 1; r++) {
        const bool have = r <= nyg - 1;  // block-uniform
        const bool nq = r + 1 <= rlast, nf = r + 1 <= rflast;
        if (nf) fetch(r + 1, nq, true);
        const int jd = r - 1;            // row completed by this one
        const bool use = act && jd >= j0 && jd <= j1;
        const double wap = use ? ap(jd) : 0.0, wam = am(r + 1);
        // 4 |hat S| hat(S)ij at (i, jd, k) from the filtered rows jd-1..jd+1
        double T[6];
        if (use) {
            const double sm = strain6(m, i, jd, k, uf, vf, wf_, T);
#pragma unroll
            for (int c = 0; c < 6; c++) T[c] = 4.0 * sm * T[c];
        } else {
#pragma unroll
            for (int c = 0; c < 6; c++) T[c] = 0.0;
        }
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
        if (nf) commit(r + 1, nq, true);
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
// chunk length of a march over n planes or rows with `tiles` tiles per plane
inline int dsm_chunk(int n, int tiles, int maxc) {
    const int64_t len = std::min<int64_t>(maxc, (int64_t)n * tiles / DSM_MIN_BLOCKS);
    return (int)std::max<int64_t>(2, len);
}
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
    const int stx = (d.nxg - 3 + 63) / 64, sty = (d.nyg - 2 + 3) / 4;
    const int spl = dsm_chunk(d.nzg - 3, stx * sty, SK_CHUNK), skc = (d.nzg - 3 + spl - 1) / spl, snt = stx * sty * skc;
    hipLaunchKernelGGL((k_dsm_strain<S>), dim3(snt), dim3(DSM_WG), 0, st, d, m, U, V, W, ubar, w.Q, nC, stx,
                       skc, snt, spl);
    double *Uf = w.Fu, *Vf = Uf + (int64_t)d.nx * d.nyg * d.nzg, *Wf = Vf + (int64_t)d.nxg * d.ny * d.nzg;
    auto fv = [&](auto kern, const S *F, double *Fo, int n1, int n2) {
        const int tx = (d.nxg - 3 + 63) / 64, tiles = tx * ((n2 + 3) / 4);  // nxg - 3 unique columns
        const int kch = dsm_chunk(d.nzp, tiles, FV_CHUNK), nkc = (d.nzp + kch - 1) / kch;
        hipLaunchKernelGGL(kern, dim3(tiles * nkc), dim3(DSM_WG), 0, st, d, F, ubar, Fo, tx, nkc, tiles * nkc, kch);
    };
    fv(k_dsm_filter_vel<S, 0>, U, Uf, d.nx, d.nyg);
    fv(k_dsm_filter_vel<S, 1>, V, Vf, d.nxg, d.ny);
    fv(k_dsm_filter_vel<S, 2>, W, Wf, d.nxg, d.nyg);
    const int tx = dsm_tiles_x(d), tz = dsm_tiles_z(d);
    const int nyp = d.nyg - 2;
    const int rows = dsm_chunk(nyp, tx * tz, LM_CHUNK), chunks = (nyp + rows - 1) / rows;
    hipLaunchKernelGGL(k_dsm_lm, dim3(tx * tz * chunks), dim3(DSM_WG), 0, st, d, m, w.Q, (const double *)Uf,
                       (const double *)Vf, (const double *)Wf, w.part, tx, tz, chunks, rows);
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
