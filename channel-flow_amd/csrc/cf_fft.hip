// Pressure-Poisson chain of the fractional step (projection.f90:40-139):
//   div(u*)  ->  R2C FFT in x  ->  C2C FFT in z  ->  tridiagonal solve in y
//   per (kx,kz)  ->  inverse z  ->  C2R inverse x (+ U projection).
//
// The reference takes a full complex 2-D DFT of every y-plane (FFTW,
// unnormalised FORWARD/BACKWARD) and solves all nxp*nzp modes; because the
// input is real and kxx(N-k) == kxx(k) exactly, the half spectrum
// kx = 0..nxp/2 carries the same information, so only it is stored and
// solved (initialization.f90:281-296 gives the symmetry).
//
// Spectrum layout in HBM: spec[(j' * nzp + kz) * mh + kx], j' = j-2, complex,
// kx fastest.  x-FFT rows, z-FFT column batches and Thomas lanes all read it
// with unit stride across threads.
//
// The FFT engine is an LDS Stockham autosort transform (radix 4, 2, 3) over a
// batch of lines per workgroup; each pass stages its butterflies through
// registers so a single LDS buffer suffices.
#include <cstdlib>

#include "cf_kernels.hpp"

namespace cf {

template <typename T>
__device__ __forceinline__ void dft2(Cx<T> *a) {
    Cx<T> t0 = cadd(a[0], a[1]), t1 = csub(a[0], a[1]);
    a[0] = t0; a[1] = t1;
}
template <typename T>
__device__ __forceinline__ void dft4(Cx<T> *a, int sign) {
    Cx<T> s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
    Cx<T> s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
    // W4 = sign * i
    Cx<T> wd = sign < 0 ? Cx<T>{d13.y, -d13.x} : Cx<T>{-d13.y, d13.x};
    a[0] = cadd(s02, s13);
    a[2] = csub(s02, s13);
    a[1] = cadd(d02, wd);
    a[3] = csub(d02, wd);
}
template <typename T>
__device__ __forceinline__ void dft3(Cx<T> *a, int sign) {
    const T h = (T)0.5, s3 = (T)0.86602540378443864676372317075294;
    Cx<T> t1 = cadd(a[1], a[2]), t2 = csub(a[1], a[2]);
    Cx<T> m = {a[0].x - h * t1.x, a[0].y - h * t1.y};
    // sign * i * s3 * t2
    Cx<T> r = sign < 0 ? Cx<T>{s3 * t2.y, -s3 * t2.x} : Cx<T>{-s3 * t2.y, s3 * t2.x};
    a[0] = cadd(a[0], t1);
    a[1] = cadd(m, r);
    a[2] = csub(m, r);
}

// One Stockham pass over `nlines` lines of length n held in LDS (line stride
// ls).  MAXIT bounds the butterflies per thread: callers guarantee
// nlines*n <= 8*WG.
template <int R, int WG, typename T>
__device__ __forceinline__ void fft_pass(Cx<T> *buf, int ls, int nlines, int n, int Ns,
                                         const Cx<T> *__restrict__ tw, int sign) {
    constexpr int MAXIT = (8 + R - 1) / R;
    const int nb = n / R, items = nb * nlines;
    const int stw = n / (Ns * R);
    Cx<T> v[MAXIT][R];
#pragma unroll
    for (int c = 0; c < MAXIT; c++) {
        const int it = threadIdx.x + c * WG;
        if (it < items) {
            const int line = it / nb, jj = it - line * nb;
            const int k = jj % Ns;
            const Cx<T> *src = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; r++) {
                Cx<T> a = src[jj + r * nb];
                if (r) {
                    Cx<T> w = tw[r * k * stw];
                    if (sign > 0) w.y = -w.y;
                    a = cmul(a, w);
                }
                v[c][r] = a;
            }
            if (R == 2) dft2(v[c]);
            else if (R == 4) dft4(v[c], sign);
            else dft3(v[c], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXIT; c++) {
        const int it = threadIdx.x + c * WG;
        if (it < items) {
            const int line = it / nb, jj = it - line * nb;
            const int k = jj % Ns;
            const int base = (jj / Ns) * Ns * R + k;
            Cx<T> *dst = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; r++) dst[base + r * Ns] = v[c][r];
        }
    }
    __syncthreads();
}

template <int WG, typename T>
__device__ void lds_fft(Cx<T> *buf, int ls, int nlines, const FFTPlan &p, const Cx<T> *tw,
                        int sign) {
    int Ns = 1;
    for (int s = 0; s < p.nr; s++) {
        const int R = p.radix[s];
        if (R == 4) fft_pass<4, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        else if (R == 2) fft_pass<2, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        else fft_pass<3, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        Ns *= R;
    }
}

// ---------------------------------------------------------------------------
// div(u*) fused with the forward real FFT along x.
// compute_pseudo_pressure_rhs (projection.f90:45-53) for i = 2..nxp+1 then
// the x part of the 2-D DFT (projection.f90:84-86), as a length-M complex FFT
// of z[m] = div[2m] + i div[2m+1] (M = nxp/2) and the standard split.
// ---------------------------------------------------------------------------
template <typename S, typename T>
__global__ void __launch_bounds__(256)
k_div_fftx(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V,
           const S *__restrict__ W, Cx<T> *__restrict__ spec, FFTPlan pM,
           const Cx<T> *__restrict__ twM, const Cx<T> *__restrict__ twN, int rb) {
    extern __shared__ unsigned char smem_raw[];
    Cx<T> *lds = reinterpret_cast<Cx<T> *>(smem_raw);
    const int M = pM.n, N = 2 * M, ls = M;
    const int nrows = d.nyp * d.nzp;
    const int row0 = blockIdx.x * rb;
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny;
    T *ldr = reinterpret_cast<T *>(lds);
    for (int e = threadIdx.x; e < rb * N; e += 256) {
        const int lr = e / N, ip = e - lr * N;
        const int r = row0 + lr;
        T dv = (T)0;
        if (r < nrows) {
            const int j = 2 + r / d.nzp, k = 2 + r % d.nzp, i = 2 + ip;
            dv = ((T)U[i3(nx, nyg, i, j, k)] - (T)U[i3(nx, nyg, i - 1, j, k)]) * (T)m.rdxf[i] +
                 ((T)V[i3(nxg, ny, i, j, k)] - (T)V[i3(nxg, ny, i, j - 1, k)]) * (T)m.rdyf[j] +
                 ((T)W[i3(nxg, nyg, i, j, k)] - (T)W[i3(nxg, nyg, i, j, k - 1)]) * (T)m.rdzf[k];
        }
        ldr[lr * 2 * ls + ip] = dv;
    }
    __syncthreads();
    lds_fft<256>(lds, ls, rb, pM, twM, -1);
    const int mh = M + 1;
    for (int e = threadIdx.x; e < rb * mh; e += 256) {
        const int lr = e / mh, kx = e - lr * mh;
        const int r = row0 + lr;
        if (r >= nrows) continue;
        const Cx<T> *z = lds + lr * ls;
        const Cx<T> a = z[kx % M];
        const Cx<T> bc = z[(M - kx) % M];
        const Cx<T> b = {bc.x, -bc.y};  // conj(Z[M-k])
        const Cx<T> E = {(T)0.5 * (a.x + b.x), (T)0.5 * (a.y + b.y)};
        const Cx<T> dd = csub(a, b);
        const Cx<T> O = {(T)0.5 * dd.y, -(T)0.5 * dd.x};  // (a-b)/(2i)
        const Cx<T> X = cadd(E, cmul(twN[kx], O));
        spec[(int64_t)r * mh + kx] = X;
    }
}

// ---------------------------------------------------------------------------
// C2C FFT along z over column batches of the spectrum (in place).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(512)
k_fftz(Dims d, Cx<T> *__restrict__ spec, FFTPlan pZ, const Cx<T> *__restrict__ twZ, int bcols,
       int sign) {
    extern __shared__ unsigned char smem_raw[];
    Cx<T> *lds = reinterpret_cast<Cx<T> *>(smem_raw);
    const int n = pZ.n, ls = n + 1, mh = d.mh;
    const int nch = (mh + bcols - 1) / bcols;
    const int jp = blockIdx.x / nch, ch = blockIdx.x - jp * nch;
    const int kx0 = ch * bcols;
    Cx<T> *base = spec + (int64_t)jp * n * mh;
    for (int e = threadIdx.x; e < bcols * n; e += 512) {
        const int kz = e / bcols, b = e - kz * bcols;
        const int kx = kx0 + b;
        lds[b * ls + kz] = kx < mh ? base[(int64_t)kz * mh + kx] : Cx<T>{(T)0, (T)0};
    }
    __syncthreads();
    lds_fft<512>(lds, ls, bcols, pZ, twZ, sign);
    for (int e = threadIdx.x; e < bcols * n; e += 512) {
        const int kz = e / bcols, b = e - kz * bcols;
        const int kx = kx0 + b;
        if (kx < mh) base[(int64_t)kz * mh + kx] = lds[b * ls + kz];
    }
}

// ---------------------------------------------------------------------------
// Tridiagonal solve in y per (kx,kz) (projection.f90:90-113).
//   D(j) = Dyy(j,j) + kxx(kx) + kzz(kz),  DL, DU from Dyy,  D(2) *= 3/2 for
//   the (0,0) mode.  The matrix is diagonally dominant (kxx, kzz <= 0), so
//   Gaussian elimination without pivoting (Thomas) is used; LAPACK zgtsv's
//   row interchanges, when they happen, change results at round-off only.
// One lane per mode; the forward sweep keeps c' in a scratch array so the
// back substitution streams it back.  Arithmetic is FP64 in every mode.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256)
k_thomas(Dims d, Cx<T> *__restrict__ spec, double *__restrict__ cscr,
         const double *__restrict__ kxx, const double *__restrict__ kzz,
         const double *__restrict__ dd, const double *__restrict__ dl,
         const double *__restrict__ du) {
    const int mh = d.mh, nzp = d.nzp, nyp = d.nyp;
    const int nmodes = mh * nzp;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nmodes) return;
    const int kz = q / mh, kx = q - kz * mh;
    const double kx2 = kxx[kx], kz2 = kzz[kz];
    const int64_t jstride = (int64_t)nzp * mh;
    Cx<T> *col = spec + q;
    // forward elimination; D(j) = (Dyy(j,j) + kxx) + kzz as in the reference
    double D0 = (dd[0] + kx2) + kz2;
    if (kx == 0 && kz == 0) D0 = 1.5 * D0;
    double r = 1.0 / D0;
    double cp = du[0] * r;
    Cx<T> b0 = col[0];
    double dpx = (double)b0.x * r, dpy = (double)b0.y * r;
    col[0] = Cx<T>{(T)dpx, (T)dpy};
    cscr[q] = cp;
#pragma unroll 4
    for (int jj = 1; jj < nyp; jj++) {
        const double a = dl[jj - 1];
        const double mden = ((dd[jj] + kx2) + kz2) - a * cp;
        r = 1.0 / mden;
        const Cx<T> b = col[jj * jstride];
        dpx = ((double)b.x - a * dpx) * r;
        dpy = ((double)b.y - a * dpy) * r;
        col[jj * jstride] = Cx<T>{(T)dpx, (T)dpy};
        cp = (jj < nyp - 1) ? du[jj] * r : 0.0;
        cscr[(int64_t)jj * nmodes + q] = cp;
    }
    // back substitution
    double xr = dpx, xi = dpy;
#pragma unroll 4
    for (int jj = nyp - 2; jj >= 0; jj--) {
        const double c = cscr[(int64_t)jj * nmodes + q];
        const Cx<T> dv = col[jj * jstride];
        xr = (double)dv.x - c * xr;
        xi = (double)dv.y - c * xi;
        col[jj * jstride] = Cx<T>{(T)xr, (T)xi};
    }
}

// ---------------------------------------------------------------------------
// Inverse real FFT along x (projection.f90:116-120: BACKWARD, divide by
// nxp*nzp, real part), writes phi and projects U on the same row
// (projection.f90:208-211), phi(nx) being the periodic copy of phi(2).
// ---------------------------------------------------------------------------
template <typename S, typename T, typename P>
__global__ void __launch_bounds__(256)
k_ifftx_proju(Dims d, Metrics m, const Cx<T> *__restrict__ spec, P *__restrict__ phi,
              S *__restrict__ U, FFTPlan pM, const Cx<T> *__restrict__ twM,
              const Cx<T> *__restrict__ twN, int rb, double norm) {
    extern __shared__ unsigned char smem_raw[];
    Cx<T> *lds = reinterpret_cast<Cx<T> *>(smem_raw);
    const int M = pM.n, N = 2 * M, mh = M + 1, ls = M;
    const int nrows = d.nyp * d.nzp;
    const int row0 = blockIdx.x * rb;
    Cx<T> *xr = lds + rb * ls;  // raw spectrum rows (rb * mh)
    for (int e = threadIdx.x; e < rb * mh; e += 256) {
        const int lr = e / mh, kx = e - lr * mh;
        const int r = row0 + lr;
        xr[e] = r < nrows ? spec[(int64_t)r * mh + kx] : Cx<T>{(T)0, (T)0};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < rb * M; e += 256) {
        const int lr = e / M, kx = e - lr * M;
        const Cx<T> a = xr[lr * mh + kx];
        const Cx<T> bc = xr[lr * mh + (M - kx)];
        const Cx<T> b = {bc.x, -bc.y};
        const Cx<T> s = cadd(a, b), df = csub(a, b);
        Cx<T> wc = twN[kx];
        wc.y = -wc.y;  // exp(+2 pi i k/N)
        const Cx<T> t = cmul(df, wc);
        lds[lr * ls + kx] = Cx<T>{s.x - t.y, s.y + t.x};  // s + i t
    }
    __syncthreads();
    lds_fft<256>(lds, ls, rb, pM, twM, +1);
    const T *zr = reinterpret_cast<const T *>(lds);
    const int nx = d.nx, nyg = d.nyg;
    for (int e = threadIdx.x; e < rb * N; e += 256) {
        const int lr = e / N, ip = e - lr * N;
        const int r = row0 + lr;
        if (r >= nrows) continue;
        const int jp = r / d.nzp, kp = r - jp * d.nzp;
        const double ph = (double)zr[lr * 2 * ls + ip] / norm;
        phi[((int64_t)kp * d.nyp + jp) * d.nxp + ip] = (P)ph;
        const int ipn = ip + 1 == N ? 0 : ip + 1;
        const double phn = (double)zr[lr * 2 * ls + ipn] / norm;
        const int i = 2 + ip, j = 2 + jp, k = 2 + kp;
        const int64_t qu = i3(nx, nyg, i, j, k);
        U[qu] = (S)((double)U[qu] - (phn - ph) * m.rdxc[i]);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// rows per workgroup for the x transforms: at most 2048 complex elements (so
// a radix-2 pass needs <= 4 butterflies per thread of 256) and at most 64 KiB
// of LDS for `elems_per_row` complex values of `esize` bytes per row.
static int env_int(const char *name, int dflt) {
    const char *v = getenv(name);
    return v ? atoi(v) : dflt;
}

int fft_rows_per_wg(int M, int elems_per_row, int esize) {
    int rb = 2048 / M;
    if (rb > 16) rb = 16;
    // tuning knob: cap on rows per workgroup (more workgroups per CU)
    // (defaults from a sweep at 384x256x384: fewer rows per workgroup raise
    // occupancy, especially for the inverse with its larger LDS footprint)
    const bool inv = elems_per_row > M;
    const int cap = env_int(inv ? "CF_FFT_ROWS_INV" : "CF_FFT_ROWS_FWD", inv ? 6 : 8);
    if (cap > 0 && rb > cap) rb = cap;
    const int lim = 65536 / (elems_per_row * esize);
    if (rb > lim) rb = lim;
    if (rb < 1) rb = 1;
    return rb;
}
// columns per workgroup for the z transform: at most 4096 complex elements
// (512 threads) and 64 KiB of LDS with the n+1 line padding.
int fft_cols_per_wg(int n, int esize) {
    int b = 4096 / n;
    if (b > 16) b = 16;
    const int cap = env_int("CF_FFT_COLS", 0);
    if (cap > 0 && b > cap) b = cap;
    const int lim = 65536 / ((n + 1) * esize);
    if (b > lim) b = lim;
    if (b < 1) b = 1;
    return b;
}

template <typename S, typename T>
void launch_div_fftx(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                     Cx<T> *spec, const FFTPlan &pM, const Cx<T> *twM, const Cx<T> *twN,
                     hipStream_t st) {
    const int rb = fft_rows_per_wg(pM.n, pM.n, (int)sizeof(Cx<T>));
    const int nrows = d.nyp * d.nzp;
    const size_t shm = (size_t)rb * pM.n * sizeof(Cx<T>);
    hipLaunchKernelGGL((k_div_fftx<S, T>), dim3((nrows + rb - 1) / rb), dim3(256), shm, st, d, m, U, V,
                       W, spec, pM, twM, twN, rb);
}

template <typename T>
void launch_fftz(const Dims &d, Cx<T> *spec, const FFTPlan &pZ, const Cx<T> *twZ, int sign,
                 hipStream_t st) {
    const int b = fft_cols_per_wg(pZ.n, (int)sizeof(Cx<T>));
    const int nch = (d.mh + b - 1) / b;
    const size_t shm = (size_t)b * (pZ.n + 1) * sizeof(Cx<T>);
    hipLaunchKernelGGL((k_fftz<T>), dim3(d.nyp * nch), dim3(512), shm, st, d, spec, pZ, twZ, b, sign);
}

template <typename T>
void launch_thomas(const Dims &d, Cx<T> *spec, double *cscr, const double *kxx, const double *kzz,
                   const double *dd, const double *dl, const double *du, hipStream_t st) {
    const int nmodes = d.mh * d.nzp;
    hipLaunchKernelGGL((k_thomas<T>), dim3((nmodes + 255) / 256), dim3(256), 0, st, d, spec, cscr, kxx,
                       kzz, dd, dl, du);
}

template <typename S, typename T, typename P>
void launch_ifftx_proju(const Dims &d, const Metrics &m, const Cx<T> *spec, P *phi, S *U,
                        const FFTPlan &pM, const Cx<T> *twM, const Cx<T> *twN, double norm,
                        hipStream_t st) {
    const int rb = fft_rows_per_wg(pM.n, 2 * pM.n + 1, (int)sizeof(Cx<T>));
    const int nrows = d.nyp * d.nzp;
    const size_t shm = (size_t)rb * (2 * pM.n + 1) * sizeof(Cx<T>);
    hipLaunchKernelGGL((k_ifftx_proju<S, T, P>), dim3((nrows + rb - 1) / rb), dim3(256), shm, st, d, m,
                       spec, phi, U, pM, twM, twN, rb, norm);
}

template void launch_div_fftx<double, double>(const Dims &, const Metrics &, const double *,
                                              const double *, const double *, Cx<double> *,
                                              const FFTPlan &, const Cx<double> *,
                                              const Cx<double> *, hipStream_t);
template void launch_div_fftx<float, float>(const Dims &, const Metrics &, const float *, const float *,
                                            const float *, Cx<float> *, const FFTPlan &,
                                            const Cx<float> *, const Cx<float> *, hipStream_t);
template void launch_fftz<double>(const Dims &, Cx<double> *, const FFTPlan &, const Cx<double> *, int,
                                  hipStream_t);
template void launch_fftz<float>(const Dims &, Cx<float> *, const FFTPlan &, const Cx<float> *, int,
                                 hipStream_t);
template void launch_thomas<double>(const Dims &, Cx<double> *, double *, const double *,
                                    const double *, const double *, const double *, const double *,
                                    hipStream_t);
template void launch_thomas<float>(const Dims &, Cx<float> *, double *, const double *, const double *,
                                   const double *, const double *, const double *, hipStream_t);
template void launch_ifftx_proju<double, double, double>(const Dims &, const Metrics &,
                                                         const Cx<double> *, double *, double *,
                                                         const FFTPlan &, const Cx<double> *,
                                                         const Cx<double> *, double, hipStream_t);
template void launch_ifftx_proju<float, float, float>(const Dims &, const Metrics &, const Cx<float> *,
                                                      float *, float *, const FFTPlan &,
                                                      const Cx<float> *, const Cx<float> *, double,
                                                      hipStream_t);

}  // namespace cf
