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
// Spectrum layout in HBM: spec[(j' * mh + kx) * nzp + kz], j' = j-2, complex,
// kz fastest.  The x transforms read and write it in runs of consecutive kz
// (one workgroup holds consecutive z rows of one y plane), and the fused
// z-transform / y-solve kernels read whole contiguous kz rows.
//
// The FFT engine is an LDS Stockham autosort transform (radix 8, 6, 4, 2, 3, 5, 7) over a
// batch of lines per workgroup; each pass stages its butterflies through
// registers so a single LDS buffer suffices.
#include <cstdlib>

#include <algorithm>

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
// radix 8 as two radix-4 transforms of the even and odd samples and one
// radix-2 stage with the twiddles W8^k = exp(sign 2 pi i k / 8)
template <typename T>
__device__ __forceinline__ void dft8(Cx<T> *a, int sign) {
    Cx<T> e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
    dft4(e, sign);
    dft4(o, sign);
    const T c = (T)0.70710678118654752440084436210485;
    const T sg = sign < 0 ? (T)-1 : (T)1;
    const Cx<T> w1 = {c, sg * c}, w3 = {-c, sg * c};
    Cx<T> t[4];
    t[0] = o[0];
    t[1] = cmul(o[1], w1);
    t[2] = Cx<T>{-sg * o[2].y, sg * o[2].x};  // times sign * i
    t[3] = cmul(o[3], w3);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        a[k] = cadd(e[k], t[k]);
        a[k + 4] = csub(e[k], t[k]);
    }
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

// radix 6 as two radix-3 transforms of the even and odd samples and one
// radix-2 stage with the twiddles W6^k = exp(sign 2 pi i k / 6), k = 0, 1, 2
template <typename T>
__device__ __forceinline__ void dft6(Cx<T> *a, int sign) {
    Cx<T> e[3] = {a[0], a[2], a[4]}, o[3] = {a[1], a[3], a[5]};
    dft3(e, sign);
    dft3(o, sign);
    const T h = (T)0.5, s3 = (T)0.86602540378443864676372317075294;
    const T sg = sign < 0 ? (T)-1 : (T)1;
    const Cx<T> t1 = cmul(o[1], Cx<T>{h, sg * s3});
    const Cx<T> t2 = cmul(o[2], Cx<T>{-h, sg * s3});
    a[0] = cadd(e[0], o[0]);
    a[3] = csub(e[0], o[0]);
    a[1] = cadd(e[1], t1);
    a[4] = csub(e[1], t1);
    a[2] = cadd(e[2], t2);
    a[5] = csub(e[2], t2);
}

// cos and sin of 2 pi q / R for the odd radices 5 and 7, q = 1..(R-1)/2
__host__ __device__ constexpr double root_cos(int R, int q) {
    return R == 5 ? (q == 1 ? 0.309016994374947424102293417182819059 : -0.809016994374947424102293417182819059)
                  : (q == 1 ? 0.623489801858733530525004884004239811
                            : (q == 2 ? -0.222520933956314404288902564496794759
                                      : -0.900968867902419126236102319507445051));
}
__host__ __device__ constexpr double root_sin(int R, int q) {
    return R == 5 ? (q == 1 ? 0.951056516295153572116439333379382143 : 0.587785252292473129168705954639072769)
                  : (q == 1 ? 0.78183148246802980870844452667405775
                            : (q == 2 ? 0.974927912181823607018131682993931217
                                      : 0.433883739117558120475768332848358755));
}
// DFT of odd prime length R (5, 7) from the input pairs (k, R-k):
// t_k = a_k + a_{R-k}, u_k = a_k - a_{R-k},
//   X(m)   = a_0 + sum_k t_k cos(2 pi k m/R) + sign i sum_k u_k sin(2 pi k m/R)
//   X(R-m) = a_0 + sum_k t_k cos(2 pi k m/R) - sign i sum_k u_k sin(2 pi k m/R)
// for m = 1..(R-1)/2; cos and sin of 2 pi q/R with q = k m mod R folded into
// the first half (cos even, sin odd about R/2)
template <int R, typename T>
__device__ __forceinline__ void dft_odd(Cx<T> *a, int sign) {
    constexpr int H = (R - 1) / 2;
    Cx<T> t[H], u[H];
#pragma unroll
    for (int k = 1; k <= H; k++) {
        t[k - 1] = cadd(a[k], a[R - k]);
        u[k - 1] = csub(a[k], a[R - k]);
    }
    const Cx<T> a0 = a[0];
    Cx<T> s0 = a0;
#pragma unroll
    for (int k = 0; k < H; k++) s0 = cadd(s0, t[k]);
    a[0] = s0;
#pragma unroll
    for (int m = 1; m <= H; m++) {
        Cx<T> re = a0, im = {(T)0, (T)0};
#pragma unroll
        for (int k = 1; k <= H; k++) {
            const int q = (k * m) % R;
            const T c = (T)root_cos(R, q <= H ? q : R - q);
            const T sn = (T)(q <= H ? root_sin(R, q) : -root_sin(R, R - q));
            re = Cx<T>{re.x + c * t[k - 1].x, re.y + c * t[k - 1].y};
            im = Cx<T>{im.x + sn * u[k - 1].x, im.y + sn * u[k - 1].y};
        }
        // sign i im
        const Cx<T> r = sign < 0 ? Cx<T>{im.y, -im.x} : Cx<T>{-im.y, im.x};
        a[m] = cadd(re, r);
        a[R - m] = csub(re, r);
    }
}

// LDS lines are padded with one element after every 8 (FP64 complex) or 16
// (FP32 complex) so that the strided writes of the early Stockham passes
// (stride R) and the transposed accesses spread over the LDS banks; fpad
// maps a line position to its slot.  (Measured against 1 per 32 and none:
// the z transform ran 0.158 -> 0.150 ms in FP64 and 0.094 -> 0.072 ms in
// FP32 at 384 points.)
template <typename T>
__host__ __device__ __forceinline__ int fpad(int e) {
    return sizeof(T) == 8 ? e + (e >> 3) : e + (e >> 4);
}
// slots of a padded line of n elements
template <typename T>
__host__ __device__ __forceinline__ int fft_line(int n) { return fpad<T>(n - 1) + 1; }

// Twiddle buffer of a length-n transform: the plain table exp(-2 pi i q/n),
// q = 0..n, followed by the per-pass table (n-1 entries); FP64 transforms
// use the second, FP32 ones the first.
template <typename T>
__device__ __forceinline__ int tw_base(int n) { return sizeof(T) == 8 ? n + 1 : 0; }

// a = q*b + r for 0 <= a < 2^20 and 1 <= b <= 2^12 with a precomputed 1/b:
// the float quotient is off by at most one, which the remainder test fixes
__device__ __forceinline__ void small_divmod(int a, int b, float inv_b, int &q, int &r) {
    q = (int)((float)a * inv_b);
    r = a - q * b;
    if (r < 0) { q -= 1; r += b; }
    else if (r >= b) { q += 1; r -= b; }
}

// e = q n + r along e = e0, e0 + step, e0 + 2 step, ...: two divisions at
// construction, then one add, one compare and at most one correction per step
struct StepDiv {
    int q, r, dq, dr, n;
    __device__ __forceinline__ StepDiv(int e0, int step, int n_) : n(n_) {
        q = e0 / n;
        r = e0 - q * n;
        dq = step / n;
        dr = step - dq * n;
    }
    __device__ __forceinline__ void next() {
        q += dq;
        r += dr;
        if (r >= n) { r -= n; q += 1; }
    }
};

// One Stockham pass over `nlines` lines of length n held in LDS (line stride
// ls).  MAXIT bounds the butterflies per thread: callers guarantee
// nlines*n <= 8*WG.  NC > 0 fixes n = NC, Ns = NSC and sign = SIGNC at
// compile time, so that the index arithmetic divides by constants.
template <int R, int WG, typename T, int NC = 0, int NSC = 1, int SIGNC = -1>
__device__ __forceinline__ void fft_pass(Cx<T> *buf, int ls, int nlines, int n_, int Ns_,
                                         const Cx<T> *__restrict__ tw, int sign_) {
    constexpr int MAXIT = (8 + R - 1) / R;
    const int n = NC ? NC : n_, Ns = NC ? NSC : Ns_, sign = NC ? SIGNC : sign_;
    const int nb = n / R, items = nb * nlines;
    const int stw = n / (Ns * R);
    const float inb = 1.0f / (float)nb, ins = 1.0f / (float)Ns;
    auto split = [&](int it, int &line, int &jj, int &q, int &k) {
        if constexpr (NC > 0) {
            line = it / (NC / R);
            jj = it - line * (NC / R);
            q = jj / NSC;
            k = jj - q * NSC;
        } else {
            small_divmod(it, nb, inb, line, jj);
            small_divmod(jj, Ns, ins, q, k);
        }
    };
    Cx<T> v[MAXIT][R];
#pragma unroll
    for (int c = 0; c < MAXIT; c++) {
        const int it = threadIdx.x + c * WG;
        if (it < items) {
            int line, jj, q, k;
            split(it, line, jj, q, k);
            const Cx<T> *src = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; r++) {
                Cx<T> a = src[fpad<T>(jj + r * nb)];
                if (r) {
                    // FP64: pass table, exp(-2 pi i r k / (Ns R)) at Ns-1 + k(R-1) + r-1
                    // (consecutive in k, fewer LDS bank conflicts); FP32: the plain
                    // table exp(-2 pi i q / n) at q = r k n/(Ns R) measured faster
                    Cx<T> w = sizeof(T) == 8 ? tw[Ns - 1 + k * (R - 1) + (r - 1)] : tw[r * k * stw];
                    if (sign > 0) w.y = -w.y;
                    a = cmul(a, w);
                }
                v[c][r] = a;
            }
            if constexpr (R == 2) dft2(v[c]);
            else if constexpr (R == 4) dft4(v[c], sign);
            else if constexpr (R == 8) dft8(v[c], sign);
            else if constexpr (R == 3) dft3(v[c], sign);
            else if constexpr (R == 6) dft6(v[c], sign);
            else dft_odd<R>(v[c], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXIT; c++) {
        const int it = threadIdx.x + c * WG;
        if (it < items) {
            int line, jj, q, k;
            split(it, line, jj, q, k);
            const int base = q * Ns * R + k;
            Cx<T> *dst = buf + line * ls;
#pragma unroll
            for (int r = 0; r < R; r++) dst[fpad<T>(base + r * Ns)] = v[c][r];
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
        if (R == 8) fft_pass<8, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        else if (R == 4) fft_pass<4, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        else if (R == 2) fft_pass<2, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        else if (R == 3) fft_pass<3, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        else if (R == 6) fft_pass<6, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        else if (R == 5) fft_pass<5, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        else fft_pass<7, WG>(buf, ls, nlines, p.n, Ns, tw, sign);
        Ns *= R;
    }
}

// The plan of a length-n transform (the host's make_plan): radix 8 first
// when r8, then 4, one 6 for a remaining 2 x 3, then 2, 3, 5 and 7.
struct CPlan {
    int nr;
    int radix[16];
};
__host__ __device__ constexpr CPlan cplan(int n, bool r8) {
    CPlan p{0, {}};
    int m = n;
    while (r8 && m % 8 == 0 && p.nr < 16) { p.radix[p.nr++] = 8; m /= 8; }
    while (m % 4 == 0 && p.nr < 16) { p.radix[p.nr++] = 4; m /= 4; }
    if (m % 6 == 0 && p.nr < 16) { p.radix[p.nr++] = 6; m /= 6; }  // the last 2 with a 3: one pass
    while (m % 2 == 0 && p.nr < 16) { p.radix[p.nr++] = 2; m /= 2; }
    while (m % 3 == 0 && p.nr < 16) { p.radix[p.nr++] = 3; m /= 3; }
    while (m % 5 == 0 && p.nr < 16) { p.radix[p.nr++] = 5; m /= 5; }
    while (m % 7 == 0 && p.nr < 16) { p.radix[p.nr++] = 7; m /= 7; }
    return p;
}

// lds_fft for a length N and sign known at compile time
template <int N, bool R8, int WG, int SIGN, int PASS = 0, int NS = 1, typename T>
__device__ __forceinline__ void lds_fft_c(Cx<T> *buf, int ls, int nlines, const Cx<T> *tw) {
    constexpr CPlan p = cplan(N, R8);
    if constexpr (PASS < p.nr) {
        constexpr int R = p.radix[PASS];
        fft_pass<R, WG, T, N, NS, SIGN>(buf, ls, nlines, N, NS, tw, SIGN);
        lds_fft_c<N, R8, WG, SIGN, PASS + 1, NS * R>(buf, ls, nlines, tw);
    }
}

// passes PASS .. END-1 of lds_fft_c, for a caller that runs the first pass
// from registers or the last one into global memory itself
template <int N, bool R8, int WG, int SIGN, int PASS, int END, int NS, typename T>
__device__ __forceinline__ void lds_fft_range(Cx<T> *buf, int ls, int nlines, const Cx<T> *tw) {
    constexpr CPlan p = cplan(N, R8);
    if constexpr (PASS < END) {
        constexpr int R = p.radix[PASS];
        fft_pass<R, WG, T, N, NS, SIGN>(buf, ls, nlines, N, NS, tw, SIGN);
        lds_fft_range<N, R8, WG, SIGN, PASS + 1, END, NS * R>(buf, ls, nlines, tw);
    }
}

// The twiddled inputs of butterfly (jj, k) of a Stockham pass (Ns, R) of a
// length-n line, as fft_pass forms them: a[r] times the pass's twiddle r k
template <int R, typename T>
__device__ __forceinline__ void fft_twiddle(Cx<T> (&a)[R], int n, int Ns, int k, const Cx<T> *tw, int sign) {
    const int stw = n / (Ns * R);
#pragma unroll
    for (int r = 1; r < R; r++) {
        Cx<T> w = sizeof(T) == 8 ? tw[Ns - 1 + k * (R - 1) + (r - 1)] : tw[r * k * stw];
        if (sign > 0) w.y = -w.y;
        a[r] = cmul(a[r], w);
    }
}
template <int R, typename T>
__device__ __forceinline__ void fft_dft(Cx<T> (&a)[R], int sign) {
    if constexpr (R == 2) dft2(a);
    else if constexpr (R == 4) dft4(a, sign);
    else if constexpr (R == 8) dft8(a, sign);
    else if constexpr (R == 3) dft3(a, sign);
    else if constexpr (R == 6) dft6(a, sign);
    else dft_odd<R>(a, sign);
}

// Per-row offsets of the x transforms, written to LDS once per workgroup so
// that the element loops add the x index instead of re-deriving (j, k) and
// three-dimensional offsets per element.
struct XRow {
    int a, b, c, d;  // forward: U, V, W and W(k-1) at i = 2; inverse: phi, U at i = 2
    int s;           // spectrum offset of (row, kx = 0): jp mh nzp + kp
    int kr;          // z index kp of the row; -1 for rows beyond the last
    int j;           // y index of the row (2..nyg-1)
    int pad;
    double r1, r2;   // forward: rdyf(j), rdzf(k); inverse: the bulk profile term added to u for dt
};

// ---------------------------------------------------------------------------
// div(u*) fused with the forward real FFT along x.
// compute_pseudo_pressure_rhs (projection.f90:45-53) for i = 2..nxp+1 then
// the x part of the 2-D DFT (projection.f90:84-86), as a length-M complex FFT
// of z[m] = div[2m] + i div[2m+1] (M = nxp/2) and the standard split.
// U(1) and W(k=1) are read at their periodic images U(nx-1), W(nz-1) (the
// values apply_boundary_conditions gives them); V(j=1) is the wall row,
// which only the boundary kernel writes (zero).
// ---------------------------------------------------------------------------
template <typename S, typename T, int WG, int MC>
__global__ void __launch_bounds__(WG)
k_div_fftx(Dims d, Metrics m, const S *__restrict__ U, const S *__restrict__ V,
           const S *__restrict__ W, Cx<T> *__restrict__ spec, FFTPlan pM,
           const Cx<T> *__restrict__ twM, const Cx<T> *__restrict__ twN, int rb,
           const double *__restrict__ ysrc, int64_t cs) {
    extern __shared__ unsigned char smem_raw[];
    U = case_ptr(U, cs); V = case_ptr(V, cs); W = case_ptr(W, cs);
    spec = case_ptr(spec, cs); ysrc = case_ptr(ysrc, cs);
    Cx<T> *lds = reinterpret_cast<Cx<T> *>(smem_raw);
    const int M = MC ? MC : pM.n, N = 2 * M, ls = fft_line<T>(M);
    const int nrows = d.nyp * d.nzp;
    // FP32: row blocks remapped per XCD (neighbouring blocks share the
    // 128-byte lines of their runs of consecutive kz, an FP32 run of 8 rows
    // being half a line, and blocks on one XCD share its L2).  FP64 keeps the
    // dispatch order, in which the block holding row j-1 (read for V) ran 6
    // blocks earlier on the same XCD rather than 48: remapped, the FP64 reads
    // grew from 1.04 to 1.27 times the algorithmic bytes
    const int row0 = (sizeof(T) == 4 ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x) * rb;
    const int nx = d.nx, nxg = d.nxg, nyg = d.nyg, ny = d.ny;
    T *ldr = reinterpret_cast<T *>(lds);
    Cx<T> *tw = lds + rb * ls;  // twiddles of the length-M transform
    XRow *rows = reinterpret_cast<XRow *>(tw + M);
    for (int e = threadIdx.x; e < M; e += WG) tw[e] = twM[tw_base<T>(M) + e];
    const int mh = M + 1;
    if (threadIdx.x < rb) {
        const int r = row0 + threadIdx.x;
        XRow t{};
        t.kr = -1;
        if (r < nrows) {
            const int jr = r / d.nzp, kr = r - jr * d.nzp;
            const int j = 2 + jr, k = 2 + kr;
            t.a = (int)i3(nx, nyg, 2, j, k);
            t.b = (int)i3(nxg, ny, 2, j, k);
            t.c = (int)i3(nxg, nyg, 2, j, k);
            // W(:,:,1) is read at its periodic image W(:,:,nz-1)
            t.d = k == 2 ? t.c + (d.nz - 3) * nxg * nyg : t.c - nxg * nyg;
            t.s = jr * mh * d.nzp + kr;
            t.kr = kr;
            t.j = j;
            t.r1 = m.rdyf[j];
            t.r2 = m.rdzf[k];
        }
        rows[threadIdx.x] = t;
    }
    __syncthreads();
    // Elements in chunks of XCH per thread: every global load of a chunk is
    // issued before the first is used (no branch around them: an element
    // outside the rows reads element 0 and is discarded), so a thread waits
    // on memory once per chunk instead of once per element.
    constexpr int XCH = 4;
    StepDiv el(threadIdx.x, WG, N);  // element (row lr = q, ip = r)
    for (int e0 = threadIdx.x; e0 < rb * N; e0 += XCH * WG) {
        S u1[XCH], u0[XCH], v1[XCH], v0[XCH], w1[XCH], w0[XCH];
        double rx[XCH];
        int lrs[XCH], ips[XCH];
#pragma unroll
        for (int c = 0; c < XCH; c++, el.next()) {
            const int lr = el.q, ip = el.r;
            lrs[c] = lr;
            ips[c] = ip;
            const bool in = e0 + c * WG < rb * N;
            const XRow &t = rows[min(lr, rb - 1)];  // read unconditionally (selects below, no branches)
            const int ta = t.a, tb = t.b, tc = t.c, td = t.d;
            const bool ok = in && t.kr >= 0;
            // the x and z ghost faces are read at their periodic images, so the
            // fields need no ghost fill between the RK update and this kernel
            const int qu = ok ? ta + ip : 0, qum = !ok ? 0 : (ip == 0 ? ta + (nx - 3) : ta + ip - 1);  // U(1) = U(nx-1)
            const int qv = ok ? tb + ip : nxg, qw = ok ? tc + ip : 0, qwm = ok ? td + ip : 0;
            u1[c] = U[qu];
            u0[c] = U[qum];
            v1[c] = V[qv];
            v0[c] = V[qv - nxg];
            w1[c] = W[qw];
            w0[c] = W[qwm];
            rx[c] = m.rdxf[ok ? 2 + ip : 2];
        }
#pragma unroll
        for (int c = 0; c < XCH; c++) {
            if (e0 + c * WG >= rb * N) break;
            const int lr = lrs[c], ip = ips[c];
            const XRow &t = rows[lr];
            T dv = (T)0;
            if (t.kr >= 0) {
                dv = ((T)u1[c] - (T)u0[c]) * (T)rx[c] + ((T)v1[c] - (T)v0[c]) * (T)t.r1 +
                     ((T)w1[c] - (T)w0[c]) * (T)t.r2;
                if (ysrc) {
                    if (t.j == 2) dv += (T)ysrc[t.kr * d.nxp + ip];
                    if (t.j == nyg - 1) dv += (T)ysrc[(d.nzp + t.kr) * d.nxp + ip];
                }
            }
            ldr[2 * (lr * ls + fpad<T>(ip >> 1)) + (ip & 1)] = dv;
        }
    }
    __syncthreads();
    if constexpr (MC > 0) lds_fft_c<MC, FFT_X_RADIX8, WG, -1>(lds, ls, rb, tw);
    else lds_fft<WG>(lds, ls, rb, pM, tw, -1);
    // consecutive threads take consecutive rows (kz) of one kx: runs of rb
    StepDiv ek(threadIdx.x, WG, rb);  // (kx = q, row lr = r)
    for (int e = threadIdx.x; e < rb * mh; e += WG, ek.next()) {
        const int kx = ek.q, lr = ek.r;
        const int sb = rows[lr].s;
        if (rows[lr].kr < 0) continue;
        const Cx<T> *z = lds + lr * ls;
        const Cx<T> a = z[fpad<T>(kx == M ? 0 : kx)];
        const Cx<T> bc = z[fpad<T>(kx == 0 ? 0 : M - kx)];
        const Cx<T> b = {bc.x, -bc.y};  // conj(Z[M-k])
        const Cx<T> E = {(T)0.5 * (a.x + b.x), (T)0.5 * (a.y + b.y)};
        const Cx<T> dd = csub(a, b);
        const Cx<T> O = {(T)0.5 * dd.y, -(T)0.5 * dd.x};  // (a-b)/(2i)
        const Cx<T> X = cadd(E, cmul(twN[kx], O));
        spec[sb + kx * d.nzp] = X;
    }
}

// ---------------------------------------------------------------------------
// C2C FFT along z (in place).  With kz fastest, column c = j' * mh + kx is
// the contiguous run spec[c * nzp .. c * nzp + nzp), and a workgroup's batch
// of consecutive columns is one contiguous block of memory.
// ---------------------------------------------------------------------------
template <typename T, int NC, int SIGNC>
__global__ void __launch_bounds__(512)
k_fftz(Dims d, Cx<T> *__restrict__ spec, FFTPlan pZ, const Cx<T> *__restrict__ twZ, int bcols,
       int sign, int64_t cs) {
    extern __shared__ unsigned char smem_raw[];
    spec = case_ptr(spec, cs);
    Cx<T> *lds = reinterpret_cast<Cx<T> *>(smem_raw);
    const int n = NC ? NC : pZ.n, ls = fft_line<T>(n) + 1;
    const int64_t ncols = (int64_t)d.nyp * d.mh;
    const int64_t c0 = (int64_t)blockIdx.x * bcols;
    const int nb = (int)min((int64_t)bcols, ncols - c0);
    Cx<T> *base = spec + c0 * n;
    Cx<T> *tw = lds + bcols * ls;
    for (int e = threadIdx.x; e < n; e += 512) tw[e] = twZ[tw_base<T>(n) + e];
    for (int e = threadIdx.x; e < nb * n; e += 512) {
        const int b = e / n, kz = e - b * n;
        lds[b * ls + fpad<T>(kz)] = base[e];
    }
    __syncthreads();
    if constexpr (NC > 0) lds_fft_c<NC, true, 512, SIGNC>(lds, ls, nb, tw);
    else lds_fft<512>(lds, ls, nb, pZ, tw, sign);
    for (int e = threadIdx.x; e < nb * n; e += 512) {
        const int b = e / n, kz = e - b * n;
        base[e] = lds[b * ls + fpad<T>(kz)];
    }
}

// ---------------------------------------------------------------------------
// Tridiagonal solve in y, two lanes per (kx,kz) mode: the real and the
// imaginary part of the right-hand side share the matrix, so each lane runs
// the forward sweep and back substitution of projection.f90:90-113 for one
// of them (FP64 arithmetic), recomputing the same pivots.  Rows go in blocks
// of LB, the next block's values loaded while the current block is
// eliminated.  The forward sweep writes nothing per row: at the last row of
// each block it stores c' and d' (FP64 checkpoints, one pair per lane).  The
// back substitution walks the blocks from the top; for each it reads the
// block's right-hand side again, recomputes c' and d' from the previous
// block's checkpoint -- the same operations on the same values, so the same
// c' and d' -- and walks the block backwards writing the solution in place.
// Three passes over the spectrum (read, read, write) and 4/LB of a pass of
// FP64 scratch, no serial phase: the faster form whenever the modes fill the
// GPU (the sweep is latency bound, and two lanes per mode double the waves
// in flight).
// ---------------------------------------------------------------------------
// rows per block: as many as keep three waves per SIMD (FP64 20 rows, 148
// VGPRs; FP32 24 rows, 156 VGPRs), so that the 2 x 193 x 384 lanes of the
// 384x256x384 grid stay resident in one round, with fewer blocks (fewer
// load latencies per lane, less checkpoint traffic) than 16-row blocks
constexpr int TH_LB = 20, TH_LB_F32 = 24;
template <typename T> constexpr int th_lb() { return sizeof(T) == 4 ? TH_LB_F32 : TH_LB; }
size_t thomas_scratch_elems(const Dims &d) {
    return (size_t)4 * ((d.nyp + std::min(TH_LB, TH_LB_F32) - 1) / std::min(TH_LB, TH_LB_F32)) * d.mh * d.nzp;
}

// pivot reciprocal 1 / (D(j) - dl(j-1) c'(j-1)); c'(j) = du(j) times it.
// The hardware reciprocal refined by two Newton steps (within an ulp of the
// rounded quotient; the pivots are never zero, subnormal or infinite): 7
// dependent operations from c'(j-1) to c'(j) where an IEEE division needs 12,
// and this chain is what the row-by-row sweeps wait on.
__device__ __forceinline__ double thomas_pivot(double D, double lo, double cprev) {
    const double x = D - lo * cprev;
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_thomas_lane(Dims d, Cx<T> *__restrict__ spec, double *__restrict__ ckpt, const double *__restrict__ kxx,
              const double *__restrict__ kzz, const double *__restrict__ dd, const double *__restrict__ dl,
              const double *__restrict__ du, int64_t cs) {
    extern __shared__ double sco[];  // dd | dl | du
    constexpr int LB = th_lb<T>();
    spec = case_ptr(spec, cs);
    ckpt = case_ptr(ckpt, cs);
    const int nzp = d.nzp, nyp = d.nyp;
    double *s_dd = sco, *s_dl = sco + nyp, *s_du = sco + 2 * nyp;
    for (int e = threadIdx.x; e < nyp; e += blockDim.x) { s_dd[e] = dd[e]; s_dl[e] = dl[e]; s_du[e] = du[e]; }
    __syncthreads();
    const int nmodes = d.mh * nzp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // 2 q + part
    if (t >= 2 * nmodes) return;
    const int q = t >> 1;
    const int kx = q / nzp, kz = q - kx * nzp;
    const double kx2 = kxx[kx], kz2 = kzz[kz];
    const bool zm = kx == 0 && kz == 0;
    const int64_t js = 2 * (int64_t)nmodes;          // row stride in T
    T *col = reinterpret_cast<T *>(spec) + t;        // this lane's part of mode q, row 0
    const int nb = (nyp + LB - 1) / LB;
    const int64_t cks = 2 * (int64_t)nmodes;         // checkpoint stride per block
    double *ck_c = ckpt, *ck_d = ckpt + (int64_t)nb * cks;
    auto Dj = [&](int j) {
        const double D = (s_dd[j] + kx2) + kz2;
        return (j == 0 && zm) ? 1.5 * D : D;
    };
    T cur[LB], nxt[LB];
    auto loadblk = [&](int b, T *dst) {
#pragma unroll
        for (int u = 0; u < LB; u++) {
            const int j = b * LB + u;
            dst[u] = j < nyp ? col[j * js] : (T)0;
        }
    };
    // forward sweep: checkpoints only
    double cp = 0.0, dp = 0.0;
    loadblk(0, cur);
    for (int b = 0; b < nb; b++) {
        if (b + 1 < nb) loadblk(b + 1, nxt);
#pragma unroll
        for (int u = 0; u < LB; u++) {
            const int j = b * LB + u;
            if (j < nyp) {
                const double lo = j > 0 ? s_dl[j - 1] : 0.0;
                const double r = thomas_pivot(Dj(j), lo, cp);
                dp = ((double)cur[u] - lo * dp) * r;
                cp = (j < nyp - 1) ? s_du[j] * r : 0.0;
            }
        }
        ck_c[(int64_t)b * cks + t] = cp;
        ck_d[(int64_t)b * cks + t] = dp;
#pragma unroll
        for (int u = 0; u < LB; u++) cur[u] = nxt[u];
    }
    // back substitution, block by block from the top
    double x = dp;  // row nyp-1: x = d'
    loadblk(nb - 1, cur);
    double c0n = nb >= 2 ? ck_c[(int64_t)(nb - 2) * cks + t] : 0.0;
    double d0n = nb >= 2 ? ck_d[(int64_t)(nb - 2) * cks + t] : 0.0;
    for (int b = nb - 1; b >= 0; b--) {
        double cprev = c0n, dprev = d0n;  // c', d' at the row before the block
        if (b > 0) {
            loadblk(b - 1, nxt);
            c0n = b >= 2 ? ck_c[(int64_t)(b - 2) * cks + t] : 0.0;
            d0n = b >= 2 ? ck_d[(int64_t)(b - 2) * cks + t] : 0.0;
        }
        double c[LB], dv[LB];
#pragma unroll
        for (int u = 0; u < LB; u++) {
            const int j = b * LB + u;
            c[u] = 0.0;
            dv[u] = 0.0;
            if (j < nyp) {
                const double lo = j > 0 ? s_dl[j - 1] : 0.0;
                const double r = thomas_pivot(Dj(j), lo, cprev);
                dprev = ((double)cur[u] - lo * dprev) * r;
                dv[u] = dprev;
                c[u] = (j < nyp - 1) ? s_du[j] * r : 0.0;
                cprev = c[u];
            }
        }
#pragma unroll
        for (int u = LB - 1; u >= 0; u--) {
            const int j = b * LB + u;
            if (j < nyp - 1) x = dv[u] - c[u] * x;
            if (j < nyp) col[j * js] = (T)x;
        }
#pragma unroll
        for (int u = 0; u < LB; u++) cur[u] = nxt[u];
    }
}

// ---------------------------------------------------------------------------
// z transforms fused with the y solve over kx slabs (projection.f90:84-113:
// the z part of the 2-D DFT, the zgtsv per mode, the inverse).  A slab is
// the nyp rows of nzp consecutive kz at one kx, owned by one workgroup or by
// two (one per half of its rows), and nothing outside the slab is read or
// written.  Rows go in batches of ZR, the next batch's loads in flight while
// the current one is worked on.
//   Forward: a batch's rows are transformed along z with the same Stockham
//   passes as k_fftz (fixed lengths: the first pass in registers from the
//   loads; run-time plans: the rows staged in LDS first), and thread kz
//   advances the elimination of mode (kx, kz) through them, real and
//   imaginary parts sharing the pivots.  The eliminated right-hand side is
//   stored in place of the row, and the pivot ratio of the batch's last row
//   goes to an FP64 checkpoint.
//   Backward: batches from the middle of the slab to the wall.  A batch's
//   stored values are loaded, the pivot ratios recomputed from the
//   checkpoint of the batch nearer the wall (the same operations on the same
//   values, so the same ratios), and the substitution writes the solution
//   into the LDS lines, which are transformed back along z (fixed lengths:
//   the last pass stores to the spectrum).
// The spectrum moves four times (read, write; read, write) instead of seven
// (z transform 2, lane solve 3, inverse z transform 2).  Arithmetic is FP64;
// with FP32 storage the eliminated values are rounded to FP32 like the
// right-hand side and the solution around them.
// ---------------------------------------------------------------------------
// Rows per batch ZR and row groups G: with G = 2 (fixed-plan lengths, 2 n
// threads) thread t loads and stores the rows u = g, g + 2, ... of column
// t mod n (g = t / n), so a batch of 16 rows costs each thread as many
// registers as 8 rows with one thread per column; the sweeps run on the
// first n threads.  G = 1: one thread per column (run-time plans, n <= WG),
// 8 rows.
//
// halves = 1: one workgroup per slab eliminates rows 0..nyp-1 from the
// bottom wall (c', d' in projection.f90's zgtsv order, the operations of
// k_thomas_lane, so in FP64 storage the same bits as the unfused chain).
// halves = 2: the slab is solved from both walls (a twisted factorisation):
// half 0 eliminates rows 0..m-1 as above, half 1 rows nyp-1..m upward from
// the top wall, x_j = e'_j - b'_j x_(j-1) with
//   q = 1 / (D_j - du_j b'_(j+1)),  b'_j = dl_(j-1) q,  e'_j = (r_j - du_j e'_(j+1)) q,
// m = ceil(nyp/2).  The backward kernel joins the halves at rows m-1 and m,
//   x_(m-1) = (d'_(m-1) - c'_(m-1) e'_m) / (1 - c'_(m-1) b'_m),  x_m = e'_m - b'_m x_(m-1),
// which both halves compute from the same values, so both get the same
// bits, and substitutes outward: 2 mh workgroups of half the rows.  The
// forward kernel also keeps each half's last d' (e') at FP64 in the scratch
// after the checkpoints: the backward kernel reads the other half's from
// there, since the other workgroup overwrites that row of the spectrum with
// its solution.

// twiddles, the y operator and the z wavenumbers into LDS: the sweeps then
// read no global memory, so their first row waits for no load in flight
template <typename T, int WG>
__device__ __forceinline__ void zs_stage_consts(Dims d, int n, Cx<T> *tw, const Cx<T> *__restrict__ twZ,
                                                double *s_dd, double *s_dl, double *s_du, double *s_kz,
                                                const double *__restrict__ dd, const double *__restrict__ dl,
                                                const double *__restrict__ du, const double *__restrict__ kzz) {
    for (int e = threadIdx.x; e < n; e += WG) { tw[e] = twZ[tw_base<T>(n) + e]; s_kz[e] = kzz[e]; }
    for (int e = threadIdx.x; e < d.nyp; e += WG) { s_dd[e] = dd[e]; s_dl[e] = dl[e]; s_du[e] = du[e]; }
}

// rows of a half, batches per half, and the row of batch b, line u
struct ZsHalf {
    int h, nyp, m, nh, nbh, zr;
    __device__ ZsHalf(int h_, int nyp_, int zr_, int halves) : h(h_), nyp(nyp_), zr(zr_) {
        m = halves >= 2 ? (nyp + 1) / 2 : nyp;
        nh = h == 0 ? m : nyp - m;
        nbh = (nh + zr - 1) / zr;
    }
    __device__ int nl(int b) const { return max(0, min(zr, nh - b * zr)); }
    __device__ int row(int b, int u) const { return h == 0 ? b * zr + u : nyp - 1 - (b * zr + u); }
};
// checkpoint batches per half (half 0 has the more rows)
__host__ __device__ inline int zs_ck_batches(int nyp, int zr, int halves) {
    return ((halves >= 2 ? (nyp + 1) / 2 : nyp) + zr - 1) / zr;
}

// halves: 1, one workgroup per slab sweeping all rows (with G = 2 on the
// first thread group); 2, one workgroup per half slab; 3 (G = 2 only), both
// halves in one workgroup, thread group g sweeping half g, each batch ZR/2
// rows of either half (lines g ZR/2 ..), so the sweep per batch is half as
// long and the grid stays at one workgroup per slab.
template <typename T, int WG, int NC, int ZR, int G, bool TW>
__global__ void __launch_bounds__(WG)
k_zsolve_fwd(Dims d, Cx<T> *__restrict__ spec, double *__restrict__ ckpt, FFTPlan pZ,
             const Cx<T> *__restrict__ twZ, const double *__restrict__ kxx, const double *__restrict__ kzz,
             const double *__restrict__ dd, const double *__restrict__ dl, const double *__restrict__ du,
             int halves, int64_t cs) {
    static_assert(G == 1 || (NC > 0 && WG == G * NC), "row groups need a fixed length");
    static_assert(!TW || G == 2, "both halves in one workgroup need two thread groups");
    extern __shared__ unsigned char smem_raw[];
    constexpr int RPT = ZR / G;  // rows per thread per batch
    spec = case_ptr(spec, cs);
    ckpt = case_ptr(ckpt, cs);
    const int n = NC ? NC : pZ.n, ls = fft_line<T>(n) + 1, nyp = d.nyp;
    Cx<T> *lds = reinterpret_cast<Cx<T> *>(smem_raw);
    Cx<T> *tw = lds + ZR * ls;
    double *s_dd = reinterpret_cast<double *>(tw + n), *s_dl = s_dd + nyp, *s_du = s_dl + nyp, *s_kz = s_du + nyp;
    zs_stage_consts<T, WG>(d, n, tw, twZ, s_dd, s_dl, s_du, s_kz, dd, dl, du, kzz);
    constexpr bool twin = TW;  // halves = 3
    const int g = G == 1 ? 0 : (int)threadIdx.x / n, kz = (int)threadIdx.x - g * n;
    const int kx = halves == 2 ? (int)blockIdx.x >> 1 : (int)blockIdx.x;
    const int zr = twin ? ZR / 2 : ZR;  // rows of one half per batch
    const ZsHalf hf(halves == 2 ? (int)blockIdx.x & 1 : (twin ? g : 0), nyp, zr, halves);
    const int nbl = twin ? ZsHalf(0, nyp, zr, halves).nbh : hf.nbh;  // batches of the loop
    const int l0 = twin ? g * zr : 0;                                 // the half's first LDS line
    const bool act = kz < n;                    // a column (all threads when G = 2)
    const bool sweep = act && (twin || g == 0);  // runs an elimination of mode (kx, kz)
    const double kx2 = kxx[kx];
    const bool zm = kx == 0 && kz == 0;
    const int64_t rs = (int64_t)d.mh * n;  // row stride in elements
    Cx<T> *col = spec + (int64_t)kx * n + kz;
    const int nbc = zs_ck_batches(nyp, zr, halves);
    double *ck = ckpt + (int64_t)hf.h * nbc * rs + (int64_t)kx * n + kz;
    auto Dj = [&](int j) {
        const double D = (s_dd[j] + kx2) + s_kz[kz];
        return (j == 0 && zm) ? 1.5 * D : D;
    };
    // row of load slot i: twin, the half's row i; otherwise row g + G i
    auto slot = [&](int i) { return twin ? i : g + G * i; };
    // REG: fixed lengths whose plan starts with radix 8 and whose first-pass
    // butterflies of a batch number at most one per thread: a thread loads
    // the eight inputs of its butterfly (line L, jj) from global memory and
    // runs the first Stockham pass in registers, so the rows are not staged
    // in LDS first
    constexpr bool REG = NC > 0 && RPT == 8 && cplan(NC > 0 ? NC : 8, true).radix[0] == 8 && (NC / 8) * ZR <= WG;
    constexpr int NB1 = NC > 0 ? NC / 8 : 1;  // first-pass butterflies per line
    const int bt = (int)threadIdx.x, bL = bt / NB1, bj = bt - bL * NB1;
    const bool bact = bt < NB1 * ZR;
    Cx<T> *slab = spec + (int64_t)kx * n;
    // the half and batch row of this thread's line, fixed along the march
    const ZsHalf bh(twin ? bL / zr : hf.h, nyp, zr, halves);
    const int bu = twin ? bL - bh.h * zr : bL;
    Cx<T> cur[RPT], nxt[RPT];
    auto loadblk = [&](int b, Cx<T> *dst) {
        if constexpr (REG) {
            const int j = (bact && bu < bh.nl(b)) ? bh.row(b, bu) : -1;
#pragma unroll
            for (int r = 0; r < 8; r++) dst[r] = j >= 0 ? slab[j * rs + bj + r * NB1] : Cx<T>{(T)0, (T)0};
        } else {
            const int nl = hf.nl(b);
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const int u = slot(i);
                dst[i] = (act && u < nl) ? col[hf.row(b, u) * rs] : Cx<T>{(T)0, (T)0};
            }
        }
    };
    // half 0: p = c', (dr, di) = d'; half 1: p = b', (dr, di) = e'
    double p = 0.0, dr = 0.0, di = 0.0;
    loadblk(0, cur);
    for (int b = 0; b < nbl; b++) {
        const int nl = hf.nl(b);
        const int nlf = twin ? ZR : nl;  // twin: both halves' lines (a short half's spare lines unused)
        if constexpr (REG) {
            if (b + 1 < nbl) loadblk(b + 1, nxt);
            // the previous batch's sweep has read its lines (and, at b = 0,
            // the twiddles are staged)
            __syncthreads();
            if (bact && bL < nlf) {
                Cx<T> a[8];
#pragma unroll
                for (int r = 0; r < 8; r++) a[r] = cur[r];
                fft_twiddle<8>(a, NC, 1, 0, tw, -1);
                fft_dft<8>(a, -1);
                Cx<T> *dst = lds + bL * ls;
#pragma unroll
                for (int r = 0; r < 8; r++) dst[fpad<T>(bj * 8 + r)] = a[r];
            }
            __syncthreads();
            lds_fft_range<(NC > 0 ? NC : 8), true, WG, -1, 1, cplan(NC > 0 ? NC : 8, true).nr, 8>(lds, ls, nlf, tw);
        } else {
            if (act) {
#pragma unroll
                for (int i = 0; i < RPT; i++)
                    if (slot(i) < nl) lds[(l0 + slot(i)) * ls + fpad<T>(kz)] = cur[i];
            }
            if (b + 1 < nbl) loadblk(b + 1, nxt);
            __syncthreads();  // staged rows (and, at b = 0, the constants) visible
            if constexpr (NC > 0) lds_fft_c<NC, true, WG, -1>(lds, ls, nlf, tw);
            else lds_fft<WG>(lds, ls, nlf, pZ, tw, -1);
        }
        if (sweep && nl > 0) {
#pragma unroll
            for (int u = 0; u < ZR; u++) {
                if (u < nl) {
                    const int j = hf.row(b, u);
                    const Cx<T> v = lds[(l0 + u) * ls + fpad<T>(kz)];
                    // coupling to the row already eliminated, and to the next
                    const double cin = hf.h == 0 ? (j > 0 ? s_dl[j - 1] : 0.0) : (j < nyp - 1 ? s_du[j] : 0.0);
                    const double cout = hf.h == 0 ? s_du[j] : (j > 0 ? s_dl[j - 1] : 0.0);
                    const double r = thomas_pivot(Dj(j), cin, p);
                    dr = ((double)v.x - cin * dr) * r;
                    di = ((double)v.y - cin * di) * r;
                    p = cout * r;
                    col[j * rs] = Cx<T>{(T)dr, (T)di};
                }
            }
            ck[(int64_t)b * rs] = p;
            if (b == hf.nbh - 1) {  // the half's middle row: the value the join needs
                double *jn = ckpt + (int64_t)(halves >= 2 ? 2 : 1) * nbc * rs + (int64_t)kx * n + kz;
                jn[(int64_t)(2 * hf.h) * rs] = dr;
                jn[(int64_t)(2 * hf.h + 1) * rs] = di;
            }
        }
        // G = 1 or twin: a thread reads and next writes only its own LDS
        // column of its own lines, and the transform's last pass ends with a
        // barrier; G = 2 with one half: the other group stages rows of the
        // same columns next (REG: the barrier at the top of the next batch)
        if constexpr (G > 1 && !twin && !REG) __syncthreads();
#pragma unroll
        for (int i = 0; i < RPT; i++) cur[i] = nxt[i];
    }
}

template <typename T, int WG, int NC, int ZR, int G, bool TW>
__global__ void __launch_bounds__(WG)
k_zsolve_bwd(Dims d, Cx<T> *__restrict__ spec, const double *__restrict__ ckpt, FFTPlan pZ,
             const Cx<T> *__restrict__ twZ, const double *__restrict__ kxx, const double *__restrict__ kzz,
             const double *__restrict__ dd, const double *__restrict__ dl, const double *__restrict__ du,
             int halves, int64_t cs) {
    static_assert(G == 1 || (NC > 0 && WG == G * NC), "row groups need a fixed length");
    static_assert(!TW || G == 2, "both halves in one workgroup need two thread groups");
    extern __shared__ unsigned char smem_raw[];
    constexpr int RPT = ZR / G;
    spec = case_ptr(spec, cs);
    ckpt = case_ptr(ckpt, cs);
    const int n = NC ? NC : pZ.n, ls = fft_line<T>(n) + 1, nyp = d.nyp;
    Cx<T> *lds = reinterpret_cast<Cx<T> *>(smem_raw);
    Cx<T> *tw = lds + ZR * ls;
    double *s_dd = reinterpret_cast<double *>(tw + n), *s_dl = s_dd + nyp, *s_du = s_dl + nyp, *s_kz = s_du + nyp;
    zs_stage_consts<T, WG>(d, n, tw, twZ, s_dd, s_dl, s_du, s_kz, dd, dl, du, kzz);
    constexpr bool twin = TW;  // halves = 3
    const int g = G == 1 ? 0 : (int)threadIdx.x / n, kz = (int)threadIdx.x - g * n;
    const int kx = halves == 2 ? (int)blockIdx.x >> 1 : (int)blockIdx.x;
    const int zr = twin ? ZR / 2 : ZR;
    const ZsHalf hf(halves == 2 ? (int)blockIdx.x & 1 : (twin ? g : 0), nyp, zr, halves);
    const int nbl = twin ? ZsHalf(0, nyp, zr, halves).nbh : hf.nbh;
    const int l0 = twin ? g * zr : 0;
    const bool act = kz < n;
    const bool sweep = act && (twin || g == 0);
    const double kx2 = kxx[kx];
    const bool zm = kx == 0 && kz == 0;
    const int64_t rs = (int64_t)d.mh * n;
    Cx<T> *col = spec + (int64_t)kx * n + kz;
    const int nbc = zs_ck_batches(nyp, zr, halves);
    const double *ck0 = ckpt + (int64_t)kx * n + kz;  // half 0's checkpoints
    const double *ck = ck0 + (int64_t)hf.h * nbc * rs;
    auto Dj = [&](int j) {
        const double D = (s_dd[j] + kx2) + s_kz[kz];
        return (j == 0 && zm) ? 1.5 * D : D;
    };
    auto slot = [&](int i) { return twin ? i : g + G * i; };
    // REGL (fixed lengths): the last Stockham pass reads the LDS lines and
    // stores its outputs straight to the spectrum, a run of consecutive kz
    // per output index, instead of writing them to LDS for a store pass
    constexpr bool REGL = NC > 0;
    Cx<T> *slab = spec + (int64_t)kx * n;
    auto line_row = [&](int b, int L) {
        if (!twin) return L < hf.nl(b) ? hf.row(b, L) : -1;
        const ZsHalf hh(L / zr, nyp, zr, halves);
        const int u = L - hh.h * zr;
        return u < hh.nl(b) ? hh.row(b, u) : -1;
    };
    Cx<T> cur[RPT], nxt[RPT];
    auto loadblk = [&](int b, Cx<T> *dst) {
        const int nl = hf.nl(b);
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const int u = slot(i);
            dst[i] = (act && u < nl) ? col[hf.row(b, u) * rs] : Cx<T>{(T)0, (T)0};
        }
    };
    const int nb = nbl;
    loadblk(nb - 1, cur);
    // the solution at the half's middle row: halves = 1, x(nyp-1) = d'(nyp-1);
    // two halves, the join of x(m-1) and x(m) from the last c', d' of half 0
    // and b', e' of half 1
    const int m = hf.m, jmid = hf.h == 0 ? m - 1 : m;
    double xr = 0.0, xi = 0.0, c0n = 0.0;
    if (sweep) {
        const double *jn = ck0 + (int64_t)(halves >= 2 ? 2 : 1) * nbc * rs;
        const double dmr = jn[0], dmi = jn[rs];
        if (halves >= 2) {
            const ZsHalf h1(1, nyp, zr, halves);
            const double cm = ck0[(int64_t)(nbc - 1) * rs];  // half 0 has nbc batches
            const double bm = ck0[((int64_t)nbc + h1.nbh - 1) * rs];
            const double emr = jn[2 * rs], emi = jn[3 * rs];
            const double den = 1.0 - cm * bm;
            const double x1r = (dmr - cm * emr) / den;
            const double x1i = (dmi - cm * emi) / den;
            if (hf.h == 0) { xr = x1r; xi = x1i; }
            else { xr = emr - bm * x1r; xi = emi - bm * x1i; }
        } else {
            xr = dmr;
            xi = dmi;
        }
        c0n = nb >= 2 ? ck[(int64_t)(nb - 2) * rs] : 0.0;
    }
    const double xmr = xr, xmi = xi;
    for (int b = nb - 1; b >= 0; b--) {
        const int nl = hf.nl(b);
        double pprev = c0n;  // c' (b') at the row before the batch, wall side
        // e' / d' of the batch into the LDS lines (G = 1 or twin: each thread
        // its own column of its own lines; G = 2 with one half: both groups,
        // read by the first after the barrier)
        if (act) {
#pragma unroll
            for (int i = 0; i < RPT; i++)
                if (slot(i) < nl) lds[(l0 + slot(i)) * ls + fpad<T>(kz)] = cur[i];
        }
        if (b > 0) {
            loadblk(b - 1, nxt);
            c0n = (sweep && b >= 2) ? ck[(int64_t)(b - 2) * rs] : 0.0;
        }
        if constexpr (G > 1 && !twin) __syncthreads();
        if (sweep && nl > 0) {
            double c[ZR];
#pragma unroll
            for (int u = 0; u < ZR; u++) {
                c[u] = 0.0;
                if (u < nl) {
                    const int j = hf.row(b, u);
                    const double cin = hf.h == 0 ? (j > 0 ? s_dl[j - 1] : 0.0) : (j < nyp - 1 ? s_du[j] : 0.0);
                    const double cout = hf.h == 0 ? s_du[j] : (j > 0 ? s_dl[j - 1] : 0.0);
                    const double r = thomas_pivot(Dj(j), cin, pprev);
                    c[u] = cout * r;
                    pprev = c[u];
                }
            }
#pragma unroll
            for (int u = ZR - 1; u >= 0; u--) {
                if (u < nl) {
                    const int j = hf.row(b, u);
                    const Cx<T> dv = lds[(l0 + u) * ls + fpad<T>(kz)];
                    if (j == jmid) {
                        xr = xmr;
                        xi = xmi;
                    } else {
                        xr = (double)dv.x - c[u] * xr;
                        xi = (double)dv.y - c[u] * xi;
                    }
                    lds[(l0 + u) * ls + fpad<T>(kz)] = Cx<T>{(T)xr, (T)xi};
                }
            }
        }
        __syncthreads();
        const int nlf = twin ? ZR : nl;
        if constexpr (REGL) {
            constexpr int NCL = NC > 0 ? NC : 8;
            constexpr CPlan pl = cplan(NCL, true);
            constexpr int RL = pl.radix[pl.nr - 1], NSL = NCL / RL, MAXL = (8 + RL - 1) / RL;
            lds_fft_range<NCL, true, WG, 1, 0, pl.nr - 1, 1>(lds, ls, nlf, tw);
            // the last pass (Ns = NSL, so q = 0 and k = jj): butterfly jj of
            // a line writes outputs jj + r NSL
#pragma unroll
            for (int c = 0; c < MAXL; c++) {
                const int it = (int)threadIdx.x + c * WG;
                if (it < NSL * nlf) {
                    const int L = it / NSL, jj = it - L * NSL;
                    const Cx<T> *src = lds + L * ls;
                    Cx<T> a[RL];
#pragma unroll
                    for (int r = 0; r < RL; r++) a[r] = src[fpad<T>(jj + r * NSL)];
                    fft_twiddle<RL>(a, NCL, NSL, jj, tw, 1);
                    fft_dft<RL>(a, 1);
                    const int j = line_row(b, L);
                    if (j >= 0) {
#pragma unroll
                        for (int r = 0; r < RL; r++) slab[j * rs + jj + r * NSL] = a[r];
                    }
                }
            }
            __syncthreads();  // the next batch writes the lines this pass read
        } else {
            lds_fft<WG>(lds, ls, nlf, pZ, tw, 1);
            if (act) {
#pragma unroll
                for (int i = 0; i < RPT; i++) {
                    const int u = slot(i);
                    if (u < nl) col[hf.row(b, u) * rs] = lds[(l0 + u) * ls + fpad<T>(kz)];
                }
            }
            // the next batch's staging writes only the rows this thread has
            // just stored from, and the sweep reads them after the barrier above
        }
#pragma unroll
        for (int i = 0; i < RPT; i++) cur[i] = nxt[i];
    }
}

// ---------------------------------------------------------------------------
// Tridiagonal solve in y per (kx,kz) (projection.f90:90-113).
//   D(j) = Dyy(j,j) + kxx(kx) + kzz(kz),  DL, DU from Dyy,  D(2) *= 3/2 for
//   the (0,0) mode.  The matrix is diagonally dominant (kxx, kzz <= 0), so
//   elimination without pivoting is used; LAPACK zgtsv's row interchanges,
//   when they happen, change results at round-off only.
// Partitioned solve, one read and one write of the spectrum: a tile is one
// kx and KZ consecutive kz; thread (kz, s) holds rows [s*MS, s*MS+MS) of
// its system in registers and solves that segment for the right-hand side
// (y) and for unit couplings to the neighbouring segments' end values
// (v: lower, w: upper), so x = y + alpha L(s-1) + beta F(s+1) with F, L the
// first and last unknown of each segment.  The 2x2-block system of the
// segment end values is solved by one thread per system through LDS, then
// every thread assembles and stores its rows.
// The workgroups are persistent: each walks the tiles with stride gridDim.x
// and issues the loads of its next tile before solving the current one, so
// that the memory latency of a tile overlaps the sweeps and the reduced
// solve of the previous (one workgroup fills a CU's registers).  The y
// operator and the wavenumbers sit in LDS, so no other global load queues
// behind the prefetch.
// Arithmetic is FP64 in every mode.
// ---------------------------------------------------------------------------
template <typename T, int MS>
__global__ void __launch_bounds__(512)
k_thomas(Dims d, Cx<T> *__restrict__ spec, const double *__restrict__ kxx, const double *__restrict__ kzz,
         const double *__restrict__ dd, const double *__restrict__ dl, const double *__restrict__ du,
         int KZ, int P, int64_t cs) {
    // [8 fields per (segment, kz), field-major: sh[f * NT + slot]] | dd | dl | du | kxx | kzz
    extern __shared__ double sh[];
    spec = case_ptr(spec, cs);
    const int nzp = d.nzp, nyp = d.nyp, mh = d.mh;
    const int nkzb = (nzp + KZ - 1) / KZ, ntiles = mh * nkzb;
    const int tid = threadIdx.x, kzl = tid % KZ, sg = tid / KZ;
    const int NT = blockDim.x;
    double *s_dd = sh + 8 * NT, *s_dl = s_dd + nyp, *s_du = s_dl + nyp, *s_kx = s_du + nyp, *s_kz = s_kx + mh;
    for (int e = tid; e < nyp; e += NT) { s_dd[e] = dd[e]; s_dl[e] = dl[e]; s_du[e] = du[e]; }
    for (int e = tid; e < mh; e += NT) s_kx[e] = kxx[e];
    for (int e = tid; e < nzp; e += NT) s_kz[e] = kzz[e];
    const int a = sg * MS;
    const int len = sg < P ? min(MS, nyp - a) : 0;
    const int64_t rs = (int64_t)mh * nzp;
    // couplings to the neighbouring segments' end values
    const double lo_s = (sg < P && sg > 0) ? -dl[a - 1] : 0.0;        // times L(s-1)
    const double up_s = (sg < P && a + len < nyp) ? -du[a + len - 1] : 0.0;  // times F(s+1)
    auto load = [&](int tile, double *r, double *im) {
        const int kx = tile / nkzb, kz = (tile - kx * nkzb) * KZ + kzl;
        const bool act = tile < ntiles && sg < P && kz < nzp;
        const Cx<T> *c = spec + (int64_t)kx * nzp + kz;
#pragma unroll
        for (int i = 0; i < MS; i++) {
            r[i] = im[i] = 0.0;
            if (act && i < len) {
                const Cx<T> b = c[(a + i) * rs];
                r[i] = (double)b.x;
                im[i] = (double)b.y;
            }
        }
    };
    double yr[MS], yi[MS];
    int tile = blockIdx.x;
    load(tile, yr, yi);
    __syncthreads();
    for (; tile < ntiles; tile += gridDim.x) {
        double nr[MS], ni[MS];
        load(tile + gridDim.x, nr, ni);  // the next tile, in flight while this one is solved
        const int kx = tile / nkzb, kz = (tile - kx * nkzb) * KZ + kzl;
        const bool active = sg < P && kz < nzp;
        const double kx2 = s_kx[kx], kz2 = kz < nzp ? s_kz[kz] : 0.0;
        const bool zm = kx == 0 && kz == 0;
        double v[MS], w[MS], cp[MS];
        // segment forward sweep for the three right-hand sides (b, e_first, e_last)
        {
            double cprev = 0.0, yrp = 0.0, yip = 0.0, vp = 0.0, wp = 0.0;
#pragma unroll
            for (int i = 0; i < MS; i++) {
                v[i] = w[i] = cp[i] = 0.0;
                if (i < len) {
                    const int j = a + i;
                    double D = (s_dd[j] + kx2) + kz2;
                    if (j == 0 && zm) D = 1.5 * D;
                    const double lo = i > 0 ? s_dl[j - 1] : 0.0;
                    const double r = 1.0 / (D - lo * cprev);
                    cp[i] = (i < len - 1) ? s_du[j] * r : 0.0;
                    yr[i] = (yr[i] - lo * yrp) * r;
                    yi[i] = (yi[i] - lo * yip) * r;
                    v[i] = ((i == 0 ? 1.0 : 0.0) - lo * vp) * r;
                    w[i] = ((i == len - 1 ? 1.0 : 0.0) - lo * wp) * r;
                    cprev = cp[i]; yrp = yr[i]; yip = yi[i]; vp = v[i]; wp = w[i];
                }
            }
#pragma unroll
            for (int i = MS - 2; i >= 0; i--) {
                if (i + 1 < len) {
                    yr[i] -= cp[i] * yr[i + 1];
                    yi[i] -= cp[i] * yi[i + 1];
                    v[i] -= cp[i] * v[i + 1];
                    w[i] -= cp[i] * w[i + 1];
                }
            }
        }
        const double lo_c = active ? lo_s : 0.0, up_c = active ? up_s : 0.0;
#define SH(f, slot) sh[(size_t)(f) * NT + (slot)]
        if (active) {
            double ymr = 0, ymi = 0, vm = 0, wm = 0;
#pragma unroll
            for (int i = 0; i < MS; i++)
                if (i == len - 1) { ymr = yr[i]; ymi = yi[i]; vm = v[i]; wm = w[i]; }
            SH(0, tid) = yr[0]; SH(1, tid) = yi[0]; SH(2, tid) = ymr; SH(3, tid) = ymi;
            SH(4, tid) = lo_c * v[0]; SH(5, tid) = lo_c * vm;  // alpha at the first / last row
            SH(6, tid) = up_c * w[0]; SH(7, tid) = up_c * wm;  // beta at the first / last row
        }
        __syncthreads();
        // end values: F(t) = P(t) + Q(t) F(t+1), L(t) = G(t) + H(t) F(t+1)
        if (sg == 0 && kz < nzp) {
            double Gr = 0, Gi = 0, H = 0;
            for (int t = 0; t < P; t++) {
                const int q = t * KZ + kzl;
                const double y1r = SH(0, q), y1i = SH(1, q), ymr = SH(2, q), ymi = SH(3, q);
                const double a1 = SH(4, q), am = SH(5, q), b1 = SH(6, q), bm = SH(7, q);
                const double rD = 1.0 / (1.0 - a1 * H);
                const double Pr = (y1r + a1 * Gr) * rD, Pi = (y1i + a1 * Gi) * rD, Q = b1 * rD;
                const double Gnr = ymr + am * Gr + am * H * Pr, Gni = ymi + am * Gi + am * H * Pi;
                const double Hn = am * H * Q + bm;
                SH(0, q) = Pr; SH(1, q) = Pi; SH(2, q) = Q; SH(3, q) = Gnr; SH(4, q) = Gni; SH(5, q) = Hn;
                Gr = Gnr; Gi = Gni; H = Hn;
            }
            double Fr = 0, Fi = 0;
            for (int t = P - 1; t >= 0; t--) {
                const int q = t * KZ + kzl;
                const double Pr = SH(0, q), Pi = SH(1, q), Q = SH(2, q), Gnr = SH(3, q), Gni = SH(4, q),
                             Hn = SH(5, q);
                const double Lr = Gnr + Hn * Fr, Li = Gni + Hn * Fi;
                Fr = Pr + Q * Fr;
                Fi = Pi + Q * Fi;
                SH(0, q) = Fr; SH(1, q) = Fi; SH(2, q) = Lr; SH(3, q) = Li;
            }
        }
        __syncthreads();
        if (active) {
            double Lpr = 0, Lpi = 0, Fnr = 0, Fni = 0;
            if (sg > 0) {
                const int q = (sg - 1) * KZ + kzl;
                Lpr = SH(2, q); Lpi = SH(3, q);
            }
            if (sg + 1 < P) {
                const int q = (sg + 1) * KZ + kzl;
                Fnr = SH(0, q); Fni = SH(1, q);
            }
            Cx<T> *col = spec + (int64_t)kx * nzp + kz;
#pragma unroll
            for (int i = 0; i < MS; i++) {
                if (i < len) {
                    const double al = lo_c * v[i], be = up_c * w[i];
                    const double xr = yr[i] + al * Lpr + be * Fnr, xi = yi[i] + al * Lpi + be * Fni;
                    col[(a + i) * rs] = Cx<T>{(T)xr, (T)xi};
                }
            }
        }
#undef SH
        __syncthreads();  // the reduced-system slots are reused by the next tile
#pragma unroll
        for (int i = 0; i < MS; i++) { yr[i] = nr[i]; yi[i] = ni[i]; }
    }
}

// ---------------------------------------------------------------------------
// Inverse real FFT along x (projection.f90:116-120: BACKWARD, divide by
// nxp*nzp, real part), writes phi and projects U on the same row
// (projection.f90:208-211), phi(nx) being the periodic copy of phi(2).
// ---------------------------------------------------------------------------
template <typename S, typename T, typename P, int WG, int MC>
__global__ void __launch_bounds__(WG)
k_ifftx_proju(Dims d, Metrics m, const Cx<T> *__restrict__ spec, P *__restrict__ phi,
              S *__restrict__ U, FFTPlan pM, const Cx<T> *__restrict__ twM,
              const Cx<T> *__restrict__ twN, int rb, double norm, const Scalars *__restrict__ csc, DtParts dtp,
              int gcopy, int64_t cs) {
    extern __shared__ unsigned char smem_raw[];
    spec = case_ptr(spec, cs); phi = case_ptr(phi, cs); U = case_ptr(U, cs); csc = case_ptr(csc, cs);
    dtp.pu = case_ptr(dtp.pu, cs); dtp.pv = case_ptr(dtp.pv, cs); dtp.pw = case_ptr(dtp.pw, cs);
    dtp.ubar = case_ptr(dtp.ubar, cs); dtp.sc = case_ptr(dtp.sc, cs);
    Cx<T> *lds = reinterpret_cast<Cx<T> *>(smem_raw);
    const int M = MC ? MC : pM.n, N = 2 * M, mh = M + 1, ls = fft_line<T>(mh);
    const int nrows = d.nyp * d.nzp;
    // row blocks remapped per XCD: neighbouring blocks share the 128-byte
    // lines of their runs of consecutive kz (an FP32 run of 8 rows is half a
    // line), and blocks on one XCD share its L2
    const int row0 = xcd_remap(blockIdx.x, gridDim.x) * rb;
    Cx<T> *tw = lds + rb * ls;  // twiddles of the length-M transform
    for (int e = threadIdx.x; e < M; e += WG) tw[e] = twM[tw_base<T>(M) + e];
    const float inz = 1.0f / (float)d.nzp;
    // consecutive threads read consecutive rows (kz) of one kx: runs of rb
    {
        // all loads of the thread first (rb * mh <= 8 WG + rb, so 9 cover
        // WG >= 256), then the LDS writes; a plain loop takes any remainder
        constexpr int MAXL = 9;
        T vr[MAXL], vi[MAXL];
        StepDiv ek(threadIdx.x, WG, rb);  // (kx = q, row lr = r)
#pragma unroll
        for (int c = 0; c < MAXL; c++, ek.next()) {
            vr[c] = vi[c] = (T)0;
            const int r = row0 + ek.r;
            if (threadIdx.x + c * WG < rb * mh && r < nrows) {
                int jp, kp;
                small_divmod(r, d.nzp, inz, jp, kp);
                const Cx<T> v = spec[((int64_t)jp * mh + ek.q) * d.nzp + kp];
                vr[c] = v.x;
                vi[c] = v.y;
            }
        }
        StepDiv es(threadIdx.x, WG, rb);
#pragma unroll
        for (int c = 0; c < MAXL; c++, es.next())
            if (threadIdx.x + c * WG < rb * mh) lds[es.r * ls + fpad<T>(es.q)] = Cx<T>{vr[c], vi[c]};
        for (int e = threadIdx.x + MAXL * WG; e < rb * mh; e += WG, ek.next()) {
            const int kx = ek.q, lr = ek.r;
            const int r = row0 + lr;
            T vx = (T)0, vy = (T)0;
            if (r < nrows) {
                int jp, kp;
                small_divmod(r, d.nzp, inz, jp, kp);
                const Cx<T> v = spec[((int64_t)jp * mh + kx) * d.nzp + kp];
                vx = v.x;
                vy = v.y;
            }
            lds[lr * ls + fpad<T>(kx)] = Cx<T>{vx, vy};
        }
    }
    // row table of the output loop
    XRow *rows = reinterpret_cast<XRow *>(tw + M);
    if (threadIdx.x < rb) {
        const int r = row0 + threadIdx.x;
        XRow t{};
        t.kr = -1;
        if (r < nrows) {
            const int jp = r / d.nzp, kp = r - jp * d.nzp;
            t.a = (kp * d.nyp + jp) * d.nxp;
            t.b = (int)i3(d.nx, d.nyg, 2, 2 + jp, 2 + kp);
            t.kr = kp;
            t.j = 2 + jp;
            if (dtp.pu && dtp.ubar) t.r1 = dtp.add_corr ? dtp.ubar[2 + jp] + dtp.sc->corr : dtp.ubar[2 + jp];
        }
        rows[threadIdx.x] = t;
    }
    __syncthreads();
    // packed length-M input of the complex inverse, in place: each thread
    // owns the pair (k, M-k) (k = 0 pairs with M) and writes Z(k), Z(M-k)
    // from the two spectrum values it read.
    //   Z(k) = (X(k) + conj X(M-k)) + i (X(k) - conj X(M-k)) exp(+2 pi i k/N)
    const int hp = M / 2 + 1;
    StepDiv ep(threadIdx.x, WG, hp);  // (row lr = q, k = r)
    for (int e = threadIdx.x; e < rb * hp; e += WG, ep.next()) {
        const int lr = ep.q, k = ep.r, k2 = M - k;
        Cx<T> *row = lds + lr * ls;
        const Cx<T> A = row[fpad<T>(k)], B = row[fpad<T>(k2)];
        auto z = [&](Cx<T> a, Cx<T> bc, int kk) {
            const Cx<T> b = {bc.x, -bc.y};
            const Cx<T> s = cadd(a, b), df = csub(a, b);
            Cx<T> wc = twN[kk];
            wc.y = -wc.y;  // exp(+2 pi i k/N)
            const Cx<T> t = cmul(df, wc);
            return Cx<T>{s.x - t.y, s.y + t.x};  // s + i t
        };
        row[fpad<T>(k)] = z(A, B, k);
        if (k != 0 && k2 != k) row[fpad<T>(k2)] = z(B, A, k2);
    }
    __syncthreads();
    if constexpr (MC > 0) lds_fft_c<MC, FFT_X_RADIX8, WG, 1>(lds, ls, rb, tw);
    else lds_fft<WG>(lds, ls, rb, pM, tw, +1);
    const T *zr = reinterpret_cast<const T *>(lds);
    const double rnorm = 1.0 / norm;
    const double corr = csc ? csc->corr : 0.0;
    // compute_dt partial: element e = tid + c WG of the loop below sits at
    // x index (tid + c WG) mod N, so a thread meets NS = N / gcd(N, WG)
    // distinct x indices, cycling with c (NS = 3 for N = 2^a 3, 1 for N =
    // 2^a): it keeps the smallest |u| per index and divides once per index at
    // the end (correctly rounded division is monotone); NS > 3 divides per value
    int NS = N, g2 = WG;
    while (g2) { const int t2 = NS % g2; NS = g2; g2 = t2; }
    NS = N / NS;
    double umin0 = INFINITY, umin1 = INFINITY, umin2 = INFINITY, lu = 0.0;
    int slot = 0;
    // one element: phi, and U projected with its previous value uo and the
    // spacing rx (loaded here unless `loaded`); sl is the element's dt slot
    auto elem = [&](int lr, int ip, const XRow &t, int sl, S uo, double rx, bool loaded) {
        const double ph = (double)zr[2 * (lr * ls + fpad<T>(ip >> 1)) + (ip & 1)] * rnorm;
        phi[t.a + ip] = (P)ph;
        const int ipn = ip + 1 == N ? 0 : ip + 1;
        const double phn = (double)zr[2 * (lr * ls + fpad<T>(ipn >> 1)) + (ipn & 1)] * rnorm;
        if (U) {
            const int i = 2 + ip;
            const int qu = t.b + ip;
            if (!loaded) {
                uo = U[qu];
                rx = m.rdxc[i];
            }
            const double up = (double)uo - (phn - ph) * rx;
            const S us = (S)(csc ? up + corr : up);
            U[qu] = us;
            if (gcopy) {  // periodic and Dirichlet-wall ghost copies of U
                const int gi = ghost_of_face(i, d.nx), gk = ghost_of_ctr(2 + t.kr, d.nzg);
                const int gj = t.j == 2 ? 1 : (t.j == d.nyg - 1 ? d.nyg : 0);
                if (gi | gj | gk) put_ghost_copies(U, d.nx, d.nyg, i, t.j, 2 + t.kr, us, gi, gj, gk);
            }
            if (dtp.pu) {
                double uf = (double)us;
                if (dtp.ubar) uf = uf + t.r1;
                const double au = fabs(uf);
                if (NS > 3) lu = fmax(lu, m.dxc[i] / au);
                else if (sl == 0) umin0 = fmin(umin0, au);
                else if (sl == 1) umin1 = fmin(umin1, au);
                else umin2 = fmin(umin2, au);
            }
        }
    };
    StepDiv el(threadIdx.x, WG, N);  // (row lr = q, ip = r)
    if constexpr (sizeof(S) == 8) {
        // elements in chunks of XCH per thread, the U values and spacings of
        // a chunk loaded before any of its stores (the unique U faces an
        // element writes are never another element's ghost copies), so a
        // thread waits on memory once per chunk instead of once per element
        // (0.320 -> 0.304 ms at 384x256x384; the FP32 kernel is slower so)
        constexpr int XCH = 4;
        for (int e0 = threadIdx.x; e0 < rb * N; e0 += XCH * WG) {
            S uo[XCH];
            double rx[XCH];
            int lrs[XCH], ips[XCH];
#pragma unroll
            for (int c = 0; c < XCH; c++, el.next()) {
                lrs[c] = el.q;
                ips[c] = el.r;
                const XRow &t = rows[min(el.q, rb - 1)];
                const bool ok = U && e0 + c * WG < rb * N && t.kr >= 0;
                uo[c] = ok ? U[t.b + el.r] : (S)0;
                rx[c] = m.rdxc[ok ? 2 + el.r : 2];
            }
#pragma unroll
            for (int c = 0; c < XCH; c++) {
                if (e0 + c * WG >= rb * N) break;
                const XRow &t = rows[lrs[c]];
                const int sl = slot;
                slot = slot + 1 == NS ? 0 : slot + 1;
                if (t.kr < 0) continue;
                elem(lrs[c], ips[c], t, sl, uo[c], rx[c], true);
            }
        }
    } else {
#pragma unroll 4
        for (int e = threadIdx.x; e < rb * N; e += WG, el.next()) {
            const int lr = el.q, ip = el.r;
            const XRow &t = rows[lr];
            const int sl = slot;
            slot = slot + 1 == NS ? 0 : slot + 1;
            if (t.kr < 0) continue;
            elem(lr, ip, t, sl, (S)0, 0.0, false);
        }
    }
    if (dtp.pu) {  // one partial maximum per workgroup
        if (NS <= 3) {
            const int t = threadIdx.x;
            if (umin0 < INFINITY) lu = fmax(lu, m.dxc[2 + (int)(t % N)] / umin0);
            if (NS > 1 && umin1 < INFINITY) lu = fmax(lu, m.dxc[2 + (int)((t + WG) % N)] / umin1);
            if (NS > 2 && umin2 < INFINITY) lu = fmax(lu, m.dxc[2 + (int)((t + 2 * WG) % N)] / umin2);
        }
        __shared__ double sh[WG / 64];
        for (int o = 32; o > 0; o >>= 1) lu = fmax(lu, __shfl_down(lu, o, 64));
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = lu;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int q = 0; q < WG / 64; q++) r = fmax(r, sh[q]);
            dtp.pu[blockIdx.x] = r;
        }
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

// Transform lengths compiled with their plan fixed (constant index
// arithmetic): half lengths M = nxp/2 of the x transforms and z lengths nzp
// of the grids 64..384; other lengths use the run-time plan.  env
// CF_FFT_STATIC=0 selects the run-time plan for every length.
#define CF_X_LENGTHS(F) F(32) F(48) F(64) F(96) F(128) F(192)
#define CF_Z_LENGTHS(F) F(64) F(96) F(128) F(192) F(256) F(384)
static bool fft_static_sizes() { return env_int("CF_FFT_STATIC", 1) != 0; }

// threads per workgroup of the x transforms (env CF_FFT_XWG: 256 or 512)
static int fft_x_wg() { return env_int("CF_FFT_XWG", 256) == 512 ? 512 : 256; }

// host twin of fft_line for a complex element of esize bytes
static int line_slots(int n, int esize) { return esize == 16 ? fft_line<double>(n) : fft_line<float>(n); }

int fft_rows_per_wg(int M, int elems_per_row, int esize) {
    int rb = 8 * fft_x_wg() / M;
    if (rb > 16) rb = 16;
    // cap on rows per workgroup (more workgroups per CU); 8 rows = runs of 8
    // consecutive kz in the spectrum, 128 bytes in FP64.  Measured best at 8
    // in both directions in FP64 at 384x256x384 (the inverse took 0.322 ms
    // with 10 rows, 0.298 ms with 8) and forward in FP32; the FP32 inverse
    // reads whole 128-byte lines with 16 rows (0.182 -> 0.179 ms)
    const bool inv = elems_per_row > M;
    const int cap = env_int(inv ? "CF_FFT_ROWS_INV" : "CF_FFT_ROWS_FWD", (inv && esize == 8) ? 16 : 8);
    if (cap > 0 && rb > cap) rb = cap;
    const int lim = (65536 / esize - M) / line_slots(elems_per_row, esize);  // + M twiddles
    if (rb > lim) rb = lim;
    if (rb < 1) rb = 1;
    return rb;
}
// columns per workgroup for the z transform: at most 4096 complex elements
// (512 threads) and 64 KiB of LDS with the n+1 line padding, and by default
// at most 1536 elements in FP64 and 3072 in FP32 (measured at length 384: 4
// FP64 columns took 0.128 ms against 0.148 ms with 9; 8 FP32 columns 0.073
// against 0.075 ms with 10).
int fft_cols_per_wg(int n, int esize) {
    int b = 4096 / n;
    if (b > 16) b = 16;
    const int dflt = std::max(1, (esize == 16 ? 1536 : 3072) / n);
    const int cap = env_int("CF_FFT_COLS", dflt);
    if (cap > 0 && b > cap) b = cap;
    const int lim = (65536 / esize - n) / (line_slots(n, esize) + 1);  // + n twiddles
    if (b > lim) b = lim;
    if (b < 1) b = 1;
    return b;
}

// y-solve geometry: MS rows per thread segment, P = ceil(nyp/MS) segments,
// KZ systems per workgroup (a power of two, KZ * P <= 512)
// (8 rows: 16 would keep 80 doubles per thread and spill)
static int thomas_ms() { return env_int("CF_TH_MS", 8) == 4 ? 4 : 8; }
static void thomas_geometry(const Dims &d, int ms, int &KZ, int &P) {
    P = (d.nyp + ms - 1) / ms;
    KZ = 1;
    while (KZ * 2 * P <= 512 && KZ * 2 <= 64) KZ *= 2;
}
bool thomas_supported(const Dims &d) { return (d.nyp + 3) / 4 <= 512; }

template <typename S, typename T>
void launch_div_fftx(const Dims &d, const Metrics &m, const S *U, const S *V, const S *W,
                     Cx<T> *spec, const FFTPlan &pM, const Cx<T> *twM, const Cx<T> *twN,
                     hipStream_t st, const double *ysrc) {
    const int rb = fft_rows_per_wg(pM.n, pM.n, (int)sizeof(Cx<T>));
    const int nrows = d.nyp * d.nzp;
    const size_t shm = ((size_t)rb * fft_line<T>(pM.n) + pM.n) * sizeof(Cx<T>) + (size_t)rb * sizeof(XRow);
    const dim3 grid((nrows + rb - 1) / rb, launch_batch().n);
    const int64_t cs = launch_batch().cs;
    if (fft_x_wg() == 512) {
        hipLaunchKernelGGL((k_div_fftx<S, T, 512, 0>), grid, dim3(512), shm, st, d, m, U, V, W, spec, pM, twM, twN,
                           rb, ysrc, cs);
        return;
    }
#define CF_DIVX(MC)                                                                                        \
    case MC:                                                                                               \
        hipLaunchKernelGGL((k_div_fftx<S, T, 256, MC>), grid, dim3(256), shm, st, d, m, U, V, W, spec, pM, \
                           twM, twN, rb, ysrc, cs);                                                            \
        return;
    switch (fft_static_sizes() ? pM.n : 0) {
        CF_X_LENGTHS(CF_DIVX)
        default:
            hipLaunchKernelGGL((k_div_fftx<S, T, 256, 0>), grid, dim3(256), shm, st, d, m, U, V, W, spec, pM, twM,
                               twN, rb, ysrc, cs);
    }
#undef CF_DIVX
}

template <typename T>
void launch_fftz(const Dims &d, Cx<T> *spec, const FFTPlan &pZ, const Cx<T> *twZ, int sign,
                 hipStream_t st) {
    const int b = fft_cols_per_wg(pZ.n, (int)sizeof(Cx<T>));
    const int64_t ncols = (int64_t)d.nyp * d.mh;
    const size_t shm = ((size_t)b * (fft_line<T>(pZ.n) + 1) + pZ.n) * sizeof(Cx<T>);
    const dim3 grid((unsigned)((ncols + b - 1) / b), launch_batch().n);
    const int64_t cs = launch_batch().cs;
#define CF_FFTZ(NC)                                                                                           \
    case NC:                                                                                                  \
        if (sign < 0)                                                                                         \
            hipLaunchKernelGGL((k_fftz<T, NC, -1>), grid, dim3(512), shm, st, d, spec, pZ, twZ, b, sign, cs);     \
        else                                                                                                  \
            hipLaunchKernelGGL((k_fftz<T, NC, 1>), grid, dim3(512), shm, st, d, spec, pZ, twZ, b, sign, cs);      \
        return;
    switch (fft_static_sizes() ? pZ.n : 0) {
        CF_Z_LENGTHS(CF_FFTZ)
        default:
            hipLaunchKernelGGL((k_fftz<T, 0, 0>), grid, dim3(512), shm, st, d, spec, pZ, twZ, b, sign, cs);
    }
#undef CF_FFTZ
}

// lane form when there are enough modes to fill the GPU one lane each (both
// precisions); the partitioned form below that
bool thomas_uses_scratch(int nmodes) { return env_int("CF_TH_LANE", nmodes >= 65536 ? 1 : 0) != 0; }

template <typename T>
void launch_thomas(const Dims &d, Cx<T> *spec, double *cscr, const double *kxx, const double *kzz,
                   const double *dd, const double *dl, const double *du, hipStream_t st) {
    const int nmodes = d.mh * d.nzp;
    if (cscr && thomas_uses_scratch(nmodes)) {
        hipLaunchKernelGGL((k_thomas_lane<T>), dim3((2 * nmodes + 255) / 256, launch_batch().n), dim3(256),
                           3 * d.nyp * sizeof(double), st, d, spec, cscr, kxx, kzz, dd, dl, du, launch_batch().cs);
        return;
    }
    int KZ, P;
    const int ms = thomas_ms();
    thomas_geometry(d, ms, KZ, P);
    const int nkzb = (d.nzp + KZ - 1) / KZ;
    const int threads = KZ * P;
    const size_t shm = ((size_t)threads * 8 + 3 * (size_t)d.nyp + d.mh + d.nzp) * sizeof(double);
    // persistent workgroups: two per CU (one resident at a time; the second
    // wave of workgroups evens out the tile counts)
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ncu = v;
    }
    const int grid = std::min(d.mh * nkzb, ncu * std::max(1, env_int("CF_TH_WGS_PER_CU", 2)));
    if (ms == 4)
        hipLaunchKernelGGL((k_thomas<T, 4>), dim3(grid, launch_batch().n), dim3(threads), shm, st, d, spec, kxx, kzz, dd, dl,
                           du, KZ, P, launch_batch().cs);
    else
        hipLaunchKernelGGL((k_thomas<T, 8>), dim3(grid, launch_batch().n), dim3(threads), shm, st, d, spec, kxx, kzz, dd, dl,
                           du, KZ, P, launch_batch().cs);
}

// fused z transforms and y solve: one workgroup per kx slab, so only where
// the slabs fill the GPU and a workgroup covers a row (env CF_ZSOLVE: 1 on,
// 0 off)
// (and where a workgroup's LDS, FP64 rows of the batch, twiddles and y
// operator, fits in the 160 KiB of a CU)
static int zsolve_rows(int n) { return (n == 384 || n == 256 || n == 192 || n == 128) ? 16 : 8; }
bool zsolve_fused(const Dims &d) {
    const size_t lds = ((size_t)zsolve_rows(d.nzp) * (fft_line<double>(d.nzp) + 1) + d.nzp) * 16 + 8 * (3 * (size_t)d.nyp + d.nzp);
    if (d.nzp > 512 || d.nyp < 4 || lds > 160 * 1024) return false;
    return env_int("CF_ZSOLVE", d.mh >= 128 ? 1 : 0) != 0;
}
// checkpoints: one c' per batch and mode, batches of at least 8 rows
// FP64 checkpoints (one per batch, half and mode) and the middle values of
// the halves (complex, one per half and mode), for batches of 8 rows or more
static size_t zsolve_scratch_elems(const Dims &d) {
    const size_t a = (size_t)2 * zs_ck_batches(d.nyp, 8, 2) + 4, b = (size_t)zs_ck_batches(d.nyp, 8, 1) + 2;
    return std::max(a, b) * d.mh * d.nzp;
}
bool ysolve_uses_scratch(const Dims &d) { return zsolve_fused(d) || thomas_uses_scratch(d.mh * d.nzp); }
size_t ysolve_scratch_elems(const Dims &d) { return std::max(thomas_scratch_elems(d), zsolve_scratch_elems(d)); }

template <typename T, int WG, int NC, int ZR, int G>
static void launch_zsolve_v(const Dims &d, Cx<T> *spec, double *cscr, const FFTPlan &pZ, const Cx<T> *twZ,
                            const double *kxx, const double *kzz, const double *dd, const double *dl,
                            const double *du, int halves, hipStream_t st) {
    const int n = pZ.n;
    const size_t shm = ((size_t)ZR * (fft_line<T>(n) + 1) + n) * sizeof(Cx<T>) + (3 * (size_t)d.nyp + n) * sizeof(double);
    if (G == 1 && halves == 3) halves = 2;  // both halves in one workgroup needs two thread groups
    const dim3 grid((halves == 2 ? 2 : 1) * d.mh, launch_batch().n);
    const int64_t cs = launch_batch().cs;
    if constexpr (G == 2) {
        if (halves == 3) {
            hipLaunchKernelGGL((k_zsolve_fwd<T, WG, NC, ZR, G, true>), grid, dim3(WG), shm, st, d, spec, cscr, pZ, twZ,
                               kxx, kzz, dd, dl, du, halves, cs);
            hipLaunchKernelGGL((k_zsolve_bwd<T, WG, NC, ZR, G, true>), grid, dim3(WG), shm, st, d, spec, cscr, pZ, twZ,
                               kxx, kzz, dd, dl, du, halves, cs);
            return;
        }
    }
    hipLaunchKernelGGL((k_zsolve_fwd<T, WG, NC, ZR, G, false>), grid, dim3(WG), shm, st, d, spec, cscr, pZ, twZ, kxx,
                       kzz, dd, dl, du, halves, cs);
    hipLaunchKernelGGL((k_zsolve_bwd<T, WG, NC, ZR, G, false>), grid, dim3(WG), shm, st, d, spec, cscr, pZ, twZ, kxx,
                       kzz, dd, dl, du, halves, cs);
}

// Forms measured at 384x256x384 in one run (forward + backward): FP64, both
// halves in one 768-thread workgroup 0.391-0.397 ms, one half sweeping all
// rows 0.429 ms, a workgroup per half on 384 threads 0.469 ms (386 half slabs
// on 256 CUs: 130 CUs take two, so the slowest CU still does 256 rows);
// FP32 0.260, 0.294 and 0.251 ms.
template <typename T>
void launch_zsolve(const Dims &d, Cx<T> *spec, double *cscr, const FFTPlan &pZ, const Cx<T> *twZ,
                   const double *kxx, const double *kzz, const double *dd, const double *dl, const double *du,
                   hipStream_t st) {
    const int n = pZ.n;
    int halves = env_int("CF_ZS_HALVES", sizeof(T) == 4 ? 2 : 3);
    halves = halves == 2 ? 2 : (halves == 3 ? 3 : 1);
    const bool groups = env_int("CF_ZS_G", sizeof(T) == 4 ? 1 : 2) == 2;
#define CF_ZSV(WG, NC, ZR, G) launch_zsolve_v<T, WG, NC, ZR, G>(d, spec, cscr, pZ, twZ, kxx, kzz, dd, dl, du, halves, st)
    switch (fft_static_sizes() ? n : 0) {
        case 384: if (groups) CF_ZSV(768, 384, 16, 2); else CF_ZSV(384, 384, 8, 1); return;
        case 256: if (groups) CF_ZSV(512, 256, 16, 2); else CF_ZSV(256, 256, 8, 1); return;
        case 192: if (groups) CF_ZSV(384, 192, 16, 2); else CF_ZSV(256, 192, 8, 1); return;
        case 128: if (groups) CF_ZSV(256, 128, 16, 2); else CF_ZSV(256, 128, 8, 1); return;
        default:
            if (n <= 256) CF_ZSV(256, 0, 8, 1);
            else CF_ZSV(512, 0, 8, 1);
    }
#undef CF_ZSV
}

int ifftx_blocks(const Dims &d, int esize) {
    const int M = d.nxp / 2;
    const int rb = fft_rows_per_wg(M, M + 1, 2 * esize);
    return (d.nyp * d.nzp + rb - 1) / rb;
}

template <typename S, typename T, typename P>
void launch_ifftx_proju(const Dims &d, const Metrics &m, const Cx<T> *spec, P *phi, S *U,
                        const FFTPlan &pM, const Cx<T> *twM, const Cx<T> *twN, double norm,
                        hipStream_t st, const Scalars *corr, const DtParts *dtpp, int gcopy) {
    const DtParts dtp = dtpp ? *dtpp : DtParts{nullptr, nullptr, nullptr, nullptr, nullptr, 0};
    const int rb = fft_rows_per_wg(pM.n, pM.n + 1, (int)sizeof(Cx<T>));
    const int nrows = d.nyp * d.nzp;
    const size_t shm = ((size_t)rb * fft_line<T>(pM.n + 1) + pM.n) * sizeof(Cx<T>) + (size_t)rb * sizeof(XRow);
    const dim3 grid((nrows + rb - 1) / rb, launch_batch().n);
    const int64_t cs = launch_batch().cs;
    if (fft_x_wg() == 512) {
        hipLaunchKernelGGL((k_ifftx_proju<S, T, P, 512, 0>), grid, dim3(512), shm, st, d, m, spec, phi, U, pM, twM,
                           twN, rb, norm, corr, dtp, gcopy, cs);
        return;
    }
#define CF_IFFTX(MC)                                                                                          \
    case MC:                                                                                                  \
        hipLaunchKernelGGL((k_ifftx_proju<S, T, P, 256, MC>), grid, dim3(256), shm, st, d, m, spec, phi, U, pM, \
                           twM, twN, rb, norm, corr, dtp, gcopy, cs);                                                    \
        return;
    switch (fft_static_sizes() ? pM.n : 0) {
        CF_X_LENGTHS(CF_IFFTX)
        default:
            hipLaunchKernelGGL((k_ifftx_proju<S, T, P, 256, 0>), grid, dim3(256), shm, st, d, m, spec, phi, U, pM,
                               twM, twN, rb, norm, corr, dtp, gcopy, cs);
    }
#undef CF_IFFTX
}

template void launch_div_fftx<double, double>(const Dims &, const Metrics &, const double *,
                                              const double *, const double *, Cx<double> *,
                                              const FFTPlan &, const Cx<double> *,
                                              const Cx<double> *, hipStream_t, const double *);
template void launch_div_fftx<float, float>(const Dims &, const Metrics &, const float *, const float *,
                                            const float *, Cx<float> *, const FFTPlan &,
                                            const Cx<float> *, const Cx<float> *, hipStream_t, const double *);
template void launch_div_fftx<double, float>(const Dims &, const Metrics &, const double *, const double *,
                                            const double *, Cx<float> *, const FFTPlan &, const Cx<float> *,
                                            const Cx<float> *, hipStream_t, const double *);
template void launch_fftz<double>(const Dims &, Cx<double> *, const FFTPlan &, const Cx<double> *, int,
                                  hipStream_t);
template void launch_fftz<float>(const Dims &, Cx<float> *, const FFTPlan &, const Cx<float> *, int,
                                 hipStream_t);
template void launch_thomas<double>(const Dims &, Cx<double> *, double *, const double *, const double *,
                                    const double *, const double *, const double *, hipStream_t);
template void launch_thomas<float>(const Dims &, Cx<float> *, double *, const double *, const double *,
                                   const double *, const double *, const double *, hipStream_t);
template void launch_zsolve<double>(const Dims &, Cx<double> *, double *, const FFTPlan &, const Cx<double> *,
                                    const double *, const double *, const double *, const double *,
                                    const double *, hipStream_t);
template void launch_zsolve<float>(const Dims &, Cx<float> *, double *, const FFTPlan &, const Cx<float> *,
                                   const double *, const double *, const double *, const double *,
                                   const double *, hipStream_t);
template void launch_ifftx_proju<double, double, double>(const Dims &, const Metrics &,
                                                         const Cx<double> *, double *, double *,
                                                         const FFTPlan &, const Cx<double> *,
                                                         const Cx<double> *, double, hipStream_t,
                                                         const Scalars *, const DtParts *, int);
template void launch_ifftx_proju<float, float, float>(const Dims &, const Metrics &, const Cx<float> *,
                                                      float *, float *, const FFTPlan &,
                                                      const Cx<float> *, const Cx<float> *, double,
                                                      hipStream_t, const Scalars *, const DtParts *, int);
template void launch_ifftx_proju<float, float, double>(const Dims &, const Metrics &, const Cx<float> *,
                                                       double *, float *, const FFTPlan &,
                                                       const Cx<float> *, const Cx<float> *, double,
                                                       hipStream_t, const Scalars *, const DtParts *, int);

}  // namespace cf
