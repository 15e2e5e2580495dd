// Learned wall model: a 4 -> 64 -> 64 -> 64 -> 1 tanh MLP mapping
// (u_par, h, nu, dp/dx) to the wall stress tau_w, trained on the GPU from
// samples produced by the EQWM kernel, and used inside the wall model with
// the ODE solver as fallback outside the training hull.
//
// Precision: the two 64x64 layers run on MFMA in BF16 with FP32
// accumulation: gfx950's v_mfma_f32_16x16x32_bf16 (K = 32, two per 64-wide
// reduction) for the forward and backward products, and
// v_mfma_f32_16x16x16_bf16 for the weight gradients, whose reduction runs
// over the 16 samples of a chunk; the 4->64 input layer and the 64->1
// output layer are FP32 FMAs (K = 4 and N = 1 do not fill a matrix tile).
// Master weights and Adam moments are FP32.
//
// Training: one single-wavefront workgroup owns 16 samples at a time, with
// its activations in LDS (the backward pass needs them in both layouts).
// Gradients of the 64x64 layers accumulate in MFMA accumulators across all
// chunks a wavefront processes, then one partial per wavefront is reduced
// in a fixed order (deterministic).  Inference (k_mlp_infer) keeps the
// activations in registers from features to stress.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "cf_internal.hpp"
#include "cf_mlp.hpp"

namespace cf {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short short8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// 16x16x32 BF16 product: lane (r16, q4) holds positions 8 q4 + j (j < 8) of
// the 32-long reduction, for row r16 of A and column r16 of B; the
// accumulator holds rows 4 q4 + i of column r16, as for K = 16.  A and B
// share the position -> reduction-index map, so any one map gives the same
// product: each kernel picks the map that matches where its data lies.
// Operands are assembled as BF16 bit patterns in a short vector and cast as
// a whole (element-wise __bf16 insertion from bit patterns was measured to
// give wrong products).
__device__ inline floatx4 mfma32(const short8 &a, const short8 &b, const floatx4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

__host__ __device__ inline unsigned short f2bf(float f) {
    unsigned u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float bf2f(unsigned short b) {
    unsigned u = (unsigned)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// parameter vector layout (FP32), also the gradient layout
struct PLayout {
    static constexpr int H = MLP_H;
    static constexpr int W1 = 0, B1 = W1 + 4 * H, W2 = B1 + H, B2 = W2 + H * H, W3 = B2 + H, B3 = W3 + H * H,
                         W4 = B3 + H, B4 = W4 + H, N = B4 + 1;
};

constexpr int LDP = MLP_H + 1;  // padded LDS row

struct WaveLds {
    float xs[16][4];
    float a1[16][LDP], a2[16][LDP], a3[16][LDP];
    float dz[16][LDP], dz2[16][LDP];
    float y[16], dy[16];
};

// forward (+ backward) over `chunks` chunks of 16 samples per wavefront
template <bool BWD>
__global__ void __launch_bounds__(64)
k_mlp(const float *__restrict__ X, const float *__restrict__ Y, int n, int chunks, float inv_b,
      const float *__restrict__ P, const unsigned short *__restrict__ W2b,
      const unsigned short *__restrict__ W3b, float *__restrict__ yout, float *__restrict__ gpart,
      float *__restrict__ lpart) {
    using L = PLayout;
    __shared__ WaveLds s;
    const int l = threadIdx.x;
    const int wave = blockIdx.x;
    const int r16 = l % 16, q4 = l / 16;
    floatx4 gW2[4][4], gW3[4][4];
    float gW1[4] = {0, 0, 0, 0}, gb1 = 0, gb2 = 0, gb3 = 0, gW4 = 0, gb4 = 0, lsum = 0;
    if (BWD) {
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) { gW2[a][b] = floatx4{0, 0, 0, 0}; gW3[a][b] = floatx4{0, 0, 0, 0}; }
    }
    const float b1 = P[L::B1 + l], b2 = P[L::B2 + l], b3 = P[L::B3 + l], w4 = P[L::W4 + l], b4 = P[L::B4];
    float w1[4];
    for (int f = 0; f < 4; f++) w1[f] = P[L::W1 + f * MLP_H + l];

    // D = act(A16x64 * W64x64 + bias) for the MFMA layers: A rows in LDS
    // (FP32, rounded to BF16 on load), W BF16 [in][out]; result rows to LDS
    auto layer = [&](float (*A)[LDP], const unsigned short *Wb, float (*D)[LDP], const float *bias) {
        for (int t = 0; t < 4; t++) {
            floatx4 acc = {0, 0, 0, 0};
            for (int kk = 0; kk < 2; kk++) {
                short8 af, bfr;
                for (int i = 0; i < 8; i++) {
                    const int k = 32 * kk + 8 * q4 + i;
                    af[i] = (short)f2bf(A[r16][k]);
                    bfr[i] = (short)Wb[k * MLP_H + 16 * t + r16];
                }
                acc = mfma32(af, bfr, acc);
            }
            for (int i = 0; i < 4; i++) {
                const int row = 4 * q4 + i, col = 16 * t + r16;
                D[row][col] = tanhf(acc[i] + bias[col]);
            }
        }
    };
    // G(16x64) = dZ(16x64) * W^T : rows of dZ in LDS, W BF16 [in][out]
    auto back = [&](float (*Dz)[LDP], const unsigned short *Wb, float (*Out)[LDP], float (*Act)[LDP]) {
        for (int t = 0; t < 4; t++) {  // output tile over the layer input index
            floatx4 acc = {0, 0, 0, 0};
            for (int kk = 0; kk < 2; kk++) {  // reduction over the layer output index
                short8 af, bfr;
                for (int i = 0; i < 8; i++) {
                    const int k = 32 * kk + 8 * q4 + i;
                    af[i] = (short)f2bf(Dz[r16][k]);
                    bfr[i] = (short)Wb[(16 * t + r16) * MLP_H + k];
                }
                acc = mfma32(af, bfr, acc);
            }
            for (int i = 0; i < 4; i++) {
                const int row = 4 * q4 + i, col = 16 * t + r16;
                const float a = Act[row][col];
                Out[row][col] = acc[i] * (1.0f - a * a);
            }
        }
    };
    // dW(64x64) += Act^T(64x16) * Dz(16x64), K = the 16 samples
    auto wgrad = [&](float (*Act)[LDP], float (*Dz)[LDP], floatx4 (*G)[4]) {
        for (int m = 0; m < 4; m++)
            for (int t = 0; t < 4; t++) {
                bf16x4 af, bfr;
                for (int i = 0; i < 4; i++) {
                    const int smp = 4 * q4 + i;
                    af[i] = (short)f2bf(Act[smp][16 * m + r16]);
                    bfr[i] = (short)f2bf(Dz[smp][16 * t + r16]);
                }
                G[m][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bfr, G[m][t], 0, 0, 0);
            }
    };

    // Cross-lane exchanges go through LDS; every phase ends with a barrier
    // (a single-wavefront workgroup, so it only orders the LDS accesses).
    __shared__ float sb2[MLP_H], sb3[MLP_H];
    sb2[l] = b2;
    sb3[l] = b3;
    __syncthreads();
    for (int c = 0; c < chunks; c++) {
        const int base = (wave * chunks + c) * 16;
        if (base >= n) break;
        {   // inputs: lane -> (sample l/4, feature l%4)
            const int smp = l / 4, f = l % 4;
            s.xs[smp][f] = base + smp < n ? X[(size_t)(base + smp) * 4 + f] : 0.0f;
        }
        __syncthreads();
        // layer 1 (FP32): lane l owns hidden unit l
        for (int smp = 0; smp < 16; smp++) {
            float z = b1;
            for (int f = 0; f < 4; f++) z += s.xs[smp][f] * w1[f];
            s.a1[smp][l] = tanhf(z);
        }
        __syncthreads();
        layer(s.a1, W2b, s.a2, sb2);
        __syncthreads();
        layer(s.a2, W3b, s.a3, sb3);
        __syncthreads();
        {   // output layer (FP32): 4 lanes per sample, 16 hidden units each
            const int smp = l / 4, part = l % 4;
            float acc = 0.f;
            for (int k = 16 * part; k < 16 * part + 16; k++) acc += s.a3[smp][k] * P[L::W4 + k];
            acc += __shfl_xor(acc, 1, 64);
            acc += __shfl_xor(acc, 2, 64);
            if (part == 0) {
                const float yv = acc + b4;
                s.y[smp] = yv;
                if (yout && base + smp < n) yout[base + smp] = yv;
                if (BWD) {
                    const bool ok = base + smp < n;
                    const float e = ok ? yv - Y[base + smp] : 0.0f;
                    s.dy[smp] = 2.0f * e * inv_b;  // d/dy of mean squared error
                    lsum += e * e;
                }
            }
        }
        __syncthreads();
        if (!BWD) continue;
        // output layer gradients and dz3
        for (int smp = 0; smp < 16; smp++) {
            const float dyv = s.dy[smp];
            const float a = s.a3[smp][l];
            gW4 += a * dyv;
            s.dz[smp][l] = dyv * w4 * (1.0f - a * a);
        }
        if (l == 0)
            for (int smp = 0; smp < 16; smp++) gb4 += s.dy[smp];
        for (int smp = 0; smp < 16; smp++) gb3 += s.dz[smp][l];
        __syncthreads();
        wgrad(s.a2, s.dz, gW3);
        back(s.dz, W3b, s.dz2, s.a2);  // dz2 = (dz3 W3^T) * (1 - a2^2)
        __syncthreads();
        for (int smp = 0; smp < 16; smp++) gb2 += s.dz2[smp][l];
        wgrad(s.a1, s.dz2, gW2);
        __syncthreads();
        back(s.dz2, W2b, s.dz, s.a1);  // dz1 into dz
        __syncthreads();
        for (int smp = 0; smp < 16; smp++) {
            const float d1 = s.dz[smp][l];
            gb1 += d1;
            for (int f = 0; f < 4; f++) gW1[f] += s.xs[smp][f] * d1;
        }
        __syncthreads();
    }
    if (!BWD) return;
    float *g = gpart + (size_t)wave * L::N;
    for (int f = 0; f < 4; f++) g[L::W1 + f * MLP_H + l] = gW1[f];
    g[L::B1 + l] = gb1;
    g[L::B2 + l] = gb2;
    g[L::B3 + l] = gb3;
    g[L::W4 + l] = gW4;
    if (l == 0) g[L::B4] = gb4;
    // MFMA accumulator layout: lane l holds rows 4*q4+i, column r16 of tile (m,t)
    for (int m = 0; m < 4; m++)
        for (int t = 0; t < 4; t++)
            for (int i = 0; i < 4; i++) {
                const int row = 16 * m + 4 * q4 + i, col = 16 * t + r16;
                g[L::W2 + row * MLP_H + col] = gW2[m][t][i];
                g[L::W3 + row * MLP_H + col] = gW3[m][t][i];
            }
    // per-wave loss: lanes 0,4,..,60 hold sample sums
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_down(lsum, o, 64);
    if (l == 0) lpart[wave] = lsum;
}

// sum per-wave gradient partials in a fixed order, then Adam
__global__ void k_mlp_adam(const float *__restrict__ gpart, int nw, float *__restrict__ P,
                           float *__restrict__ m1, float *__restrict__ m2, float *__restrict__ gout,
                           unsigned short *__restrict__ W2b, unsigned short *__restrict__ W3b, float lr,
                           float b1c, float b2c, int update) {
    using L = PLayout;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L::N) return;
    float g = 0.f;
    for (int w = 0; w < nw; w++) g += gpart[(size_t)w * L::N + p];
    if (gout) gout[p] = g;
    if (!update) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float a = beta1 * m1[p] + (1.f - beta1) * g;
    const float b = beta2 * m2[p] + (1.f - beta2) * g * g;
    m1[p] = a;
    m2[p] = b;
    const float w = P[p] - lr * (a / b1c) / (sqrtf(b / b2c) + eps);
    P[p] = w;
    if (p >= L::W2 && p < L::W2 + MLP_H * MLP_H) W2b[p - L::W2] = f2bf(w);
    if (p >= L::W3 && p < L::W3 + MLP_H * MLP_H) W3b[p - L::W3] = f2bf(w);
}

__global__ void k_sum_loss(const float *__restrict__ lp, int nw, float *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < nw; w++) s += lp[w];
        out[0] = (float)s;
    }
}

// features of raw (u, h, nu, dpdx): standardised [ln u, ln h, ln nu, asinh(dpdx/1e-3)]
__host__ __device__ inline void mlp_features(double u, double h, double nu, double dpdx, const float *nrm,
                                             float *f) {
    const double r[4] = {log(u), log(h), log(nu), asinh(dpdx / 1e-3)};
    for (int q = 0; q < 4; q++) f[q] = (float)((r[q] - nrm[q]) / nrm[4 + q]);
}

// ---------------------------------------------------------------------------
// inference
// ---------------------------------------------------------------------------
// The 64x64 layers run as transposed products H^T = W^T A^T on
// v_mfma_f32_16x16x32_bf16: the weights are the A operand and the 16
// samples of a wavefront are the columns.  The accumulator of output tile t
// holds, in lane (r16, q4), units 16t+4q4+i of sample r16.  The next layer's
// reduction block kk (K = 32) takes, in lane (r16, q4), the accumulators of
// tiles 2kk and 2kk+1: positions 8q4 + j hold units 16(2kk + j/4) + 4q4 +
// j%4, and the weight fragments are loaded with the same map -- so
// activations stay in registers from the features to the output.  The BF16 A-operand fragments
// of both 64x64 layers are loaded into registers once per wavefront and
// reused over its chunks of 16 samples; the FP32 input- and output-layer
// weights and the biases are read from LDS as float4 per 4 units.
//
// tanh here is 1 - 2/(2^(2z log2 e) + 1) on v_exp_f32 and v_rcp_f32:
// absolute error below 5e-7 against tanhf, far under the BF16 rounding
// (relative 2^-9) the activations take before the next layer.

// raw inputs: per-sample arrays, or constants (h, nu) and a device scalar
// (dp/dx) where the arrays are null
struct MlpInputs {
    const double *u, *h, *nu, *dpdx;
    double hc, nuc;
    const double *dpc;
};

__device__ inline float tanh_fast(float z) {
    const float e = __builtin_amdgcn_exp2f(z * 2.8853900817779268f);  // exp(2z)
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

constexpr int MLP_INFER_WG = 256;

// BF16 bits of x, rounded to nearest even (v_cvt_pk_bf16_f32; the same
// rounding as f2bf for finite x)
__device__ inline short bf16_bits(float x) { return __builtin_bit_cast(short, (__bf16)x); }

// raw feature q4 of sample q: ln u, ln h, ln nu, asinh(dp/dx / 1e-3)
__device__ inline double mlp_raw_feature(const MlpInputs &in, int q4, int q) {
    if (q4 == 3) return asinh((in.dpdx ? in.dpdx[q] : (in.dpc ? *in.dpc : 0.0)) / 1e-3);
    const double a = q4 == 0 ? in.u[q] : (q4 == 1 ? (in.h ? in.h[q] : in.hc) : (in.nu ? in.nu[q] : in.nuc));
    return log(a);
}

// tau[q] = exp(y * std + mean) of the network output; where the sample lies
// outside the training box, inside[q] = 0 and (mask_tau) tau[q] = 0.  nfall,
// when given, gains the number of outside samples.
__global__ void __launch_bounds__(MLP_INFER_WG)
k_mlp_infer(int n, MlpInputs in, const float *__restrict__ P, const unsigned short *__restrict__ W2b,
            const unsigned short *__restrict__ W3b, const float *__restrict__ nrm, const float *__restrict__ hull,
            int mask_tau, double *__restrict__ tau, int *__restrict__ inside, int *__restrict__ nfall) {
    using L = PLayout;
    // FP32 parameters of the input and output layers and the biases, shared
    // by the workgroup: [W1 (4 x 64) | b1 | b2 | b3 | w4], unit-fastest
    __shared__ float4 sp[8 * MLP_H / 4];
    for (int e = threadIdx.x; e < 8 * MLP_H; e += MLP_INFER_WG) {
        const int blk = e / MLP_H, u = e % MLP_H;
        const int src = blk < 4 ? L::W1 + blk * MLP_H + u
                                : (blk == 4 ? L::B1 : blk == 5 ? L::B2 : blk == 6 ? L::B3 : L::W4) + u;
        reinterpret_cast<float *>(sp)[e] = P[src];
    }
    __syncthreads();
    const int l = threadIdx.x & 63;
    const int r16 = l & 15, q4 = l >> 4;
    const int wave = blockIdx.x * (MLP_INFER_WG / 64) + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * (MLP_INFER_WG / 64);
    // lane (r16, q4) owns units 16t + 4q4 + i: the float4 at sp[blk*16 + 4t + q4]
    auto par = [&](int blk, int t) { return sp[blk * (MLP_H / 4) + 4 * t + q4]; };
    short8 a2[4][2], a3[4][2];  // [output tile][reduction block of 32]
    for (int t = 0; t < 4; t++)
        for (int kk = 0; kk < 2; kk++)
            for (int j = 0; j < 8; j++) {
                const int k = 16 * (2 * kk + j / 4) + 4 * q4 + j % 4;
                a2[t][kk][j] = (short)W2b[k * MLP_H + 16 * t + r16];
                a3[t][kk][j] = (short)W3b[k * MLP_H + 16 * t + r16];
            }
    const float b4 = P[L::B4];
    // lane q4 standardises feature q4 of sample r16; features that do not
    // vary per sample are evaluated once
    const double mu = nrm[q4], sd = nrm[4 + q4];
    const float lo = hull[q4], hi = hull[4 + q4];
    const double ym = nrm[8], ys = nrm[9];
    const bool per_sample = q4 == 0 || (q4 == 1 && in.h) || (q4 == 2 && in.nu) || (q4 == 3 && in.dpdx);
    const double rconst = per_sample ? 0.0 : mlp_raw_feature(in, q4, 0);
    int c = wave;
    double rnext = (per_sample && c * 16 + r16 < n) ? mlp_raw_feature(in, q4, c * 16 + r16) : rconst;
    for (; c * 16 < n; c += nwaves) {
        const int q = c * 16 + r16;
        const bool valid = q < n;
        const double r = rnext;
        {   // the next chunk's input, in flight while this one computes
            const int qn = q + 16 * nwaves;
            if (per_sample && qn < n) rnext = mlp_raw_feature(in, q4, qn);
        }
        float f = (float)((r - mu) / sd);
        const bool fin = valid && f >= lo && f <= hi;  // false for NaN and infinities
        if (!valid || !isfinite(f)) f = 0.0f;
        const unsigned long long m = __ballot(fin);
        const bool ok = ((m >> r16) & (m >> (r16 + 16)) & (m >> (r16 + 32)) & (m >> (r16 + 48)) & 1ull) != 0;
        float x[4];
        for (int g = 0; g < 4; g++) x[g] = __shfl(f, r16 + 16 * g, 64);
        // layer 1 (FP32)
        bf16x4 h1[4];
        for (int t = 0; t < 4; t++) {
            const float4 bb = par(4, t), w0 = par(0, t), w1 = par(1, t), w2 = par(2, t), w3 = par(3, t);
            const float zb[4] = {bb.x, bb.y, bb.z, bb.w};
            const float z0[4] = {w0.x, w0.y, w0.z, w0.w}, z1[4] = {w1.x, w1.y, w1.z, w1.w};
            const float z2[4] = {w2.x, w2.y, w2.z, w2.w}, z3[4] = {w3.x, w3.y, w3.z, w3.w};
            for (int i = 0; i < 4; i++) {
                float z = zb[i];
                z += x[0] * z0[i];
                z += x[1] * z1[i];
                z += x[2] * z2[i];
                z += x[3] * z3[i];
                h1[t][i] = bf16_bits(tanh_fast(z));
            }
        }
        // the B operands of a layer: tiles 2kk and 2kk+1 side by side
        auto bpair = [](const bf16x4 *h, int kk) {
            short8 b;
            for (int j = 0; j < 8; j++) b[j] = h[2 * kk + j / 4][j % 4];
            return b;
        };
        // layer 2
        bf16x4 h2[4];
        const short8 b10 = bpair(h1, 0), b11 = bpair(h1, 1);
        for (int t = 0; t < 4; t++) {
            floatx4 acc = {0, 0, 0, 0};
            acc = mfma32(a2[t][0], b10, acc);
            acc = mfma32(a2[t][1], b11, acc);
            const float4 bb = par(5, t);
            const float zb[4] = {bb.x, bb.y, bb.z, bb.w};
            for (int i = 0; i < 4; i++) h2[t][i] = bf16_bits(tanh_fast(acc[i] + zb[i]));
        }
        // layer 3 and the output layer (FP32)
        float y = 0.0f;
        const short8 b20 = bpair(h2, 0), b21 = bpair(h2, 1);
        for (int t = 0; t < 4; t++) {
            floatx4 acc = {0, 0, 0, 0};
            acc = mfma32(a3[t][0], b20, acc);
            acc = mfma32(a3[t][1], b21, acc);
            const float4 bb = par(6, t), ww = par(7, t);
            const float zb[4] = {bb.x, bb.y, bb.z, bb.w}, zw[4] = {ww.x, ww.y, ww.z, ww.w};
            for (int i = 0; i < 4; i++) y += tanh_fast(acc[i] + zb[i]) * zw[i];
        }
        y += __shfl_xor(y, 16, 64);
        y += __shfl_xor(y, 32, 64);
        y += b4;
        if (valid && q4 == 0) {
            tau[q] = (!mask_tau || ok) ? exp((double)y * ys + ym) : 0.0;
            if (inside) inside[q] = ok ? 1 : 0;
        }
        if (nfall) {
            const unsigned long long out = __ballot(valid && q4 == 0 && !ok);
            if (l == 0 && out) atomicAdd(nfall, (int)__popcll(out));
        }
    }
}

// waves of the inference grid: 2048 (two per SIMD on 256 CUs) at most, so
// that each wave reuses its registered weights over several chunks
inline int mlp_infer_blocks(int n) {
    const int chunks = (n + 15) / 16;
    const int waves = std::min(chunks, 2048);
    return (waves + MLP_INFER_WG / 64 - 1) / (MLP_INFER_WG / 64);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define MCK(call)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " #call);  \
            return -1;                                                                     \
        }                                                                                  \
    } while (0)

Mlp::~Mlp() {
    for (void *p : owned) hipFree(p);
    for (void *p : {(void *)X, (void *)Yd, (void *)yo, (void *)gpart, (void *)lpart})
        if (p) hipFree(p);
    if (stream) hipStreamDestroy(stream);
}

int Mlp::init(unsigned seed) {
    using L = PLayout;
    MCK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    auto al = [&](void **p, size_t b) -> int {
        MCK(hipMalloc(p, b));
        owned.push_back(*p);
        return 0;
    };
    if (al((void **)&P, L::N * 4) || al((void **)&m1, L::N * 4) || al((void **)&m2, L::N * 4) ||
        al((void **)&W2b, MLP_H * MLP_H * 2) || al((void **)&W3b, MLP_H * MLP_H * 2) ||
        al((void **)&nrm_d, 16 * 4) || al((void **)&hull_d, 8 * 4) || al((void **)&loss_d, 4))
        return -1;
    // Xavier-uniform init from a SplitMix64 stream
    std::vector<float> h(L::N, 0.f);
    unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
    auto rnd = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int fin, int fout) {
        const double a = std::sqrt(6.0 / (fin + fout));
        for (int q = 0; q < fin * fout; q++) h[off + q] = (float)((2 * rnd() - 1) * a);
    };
    fill(L::W1, 4, MLP_H);
    fill(L::W2, MLP_H, MLP_H);
    fill(L::W3, MLP_H, MLP_H);
    fill(L::W4, MLP_H, 1);
    if (set_params(h.data())) return -1;
    MCK(hipMemset(m1, 0, L::N * 4));
    MCK(hipMemset(m2, 0, L::N * 4));
    float nrm[16] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 1};
    MCK(hipMemcpy(nrm_d, nrm, sizeof nrm, hipMemcpyHostToDevice));
    float hull[8] = {-1e30f, -1e30f, -1e30f, -1e30f, 1e30f, 1e30f, 1e30f, 1e30f};
    MCK(hipMemcpy(hull_d, hull, sizeof hull, hipMemcpyHostToDevice));
    adam_t = 0;
    return 0;
}

int Mlp::set_params(const float *h) {
    using L = PLayout;
    MCK(hipMemcpy(P, h, L::N * 4, hipMemcpyHostToDevice));
    std::vector<unsigned short> w2(MLP_H * MLP_H), w3(MLP_H * MLP_H);
    for (int q = 0; q < MLP_H * MLP_H; q++) { w2[q] = f2bf(h[L::W2 + q]); w3[q] = f2bf(h[L::W3 + q]); }
    MCK(hipMemcpy(W2b, w2.data(), w2.size() * 2, hipMemcpyHostToDevice));
    MCK(hipMemcpy(W3b, w3.data(), w3.size() * 2, hipMemcpyHostToDevice));
    return 0;
}

int Mlp::get_params(float *h) {
    MCK(hipStreamSynchronize(stream));
    MCK(hipMemcpy(h, P, PLayout::N * 4, hipMemcpyDeviceToHost));
    return 0;
}

int Mlp::ensure(size_t ndata, int nwaves) {
    if (ndata > cap_data) {
        if (X) hipFree(X);
        if (Yd) hipFree(Yd);
        if (yo) hipFree(yo);
        MCK(hipMalloc(&X, ndata * 4 * 4));
        MCK(hipMalloc(&Yd, ndata * 4));
        MCK(hipMalloc(&yo, ndata * 4));
        cap_data = ndata;
    }
    if ((size_t)nwaves > cap_waves) {
        if (gpart) hipFree(gpart);
        if (lpart) hipFree(lpart);
        MCK(hipMalloc(&gpart, (size_t)nwaves * PLayout::N * 4));
        MCK(hipMalloc(&lpart, (size_t)nwaves * 4));
        cap_waves = nwaves;
    }
    return 0;
}

// standardisation from the data, training hull = feature box of the data
int Mlp::fit_normalisation(const double *feat_raw, const double *tau, size_t n) {
    double mean[5] = {0, 0, 0, 0, 0}, var[5] = {0, 0, 0, 0, 0};
    std::vector<double> r(n * 5);
    for (size_t q = 0; q < n; q++) {
        r[q * 5 + 0] = std::log(feat_raw[q * 4 + 0]);
        r[q * 5 + 1] = std::log(feat_raw[q * 4 + 1]);
        r[q * 5 + 2] = std::log(feat_raw[q * 4 + 2]);
        r[q * 5 + 3] = std::asinh(feat_raw[q * 4 + 3] / 1e-3);
        r[q * 5 + 4] = std::log(tau[q]);
        for (int c = 0; c < 5; c++) mean[c] += r[q * 5 + c];
    }
    for (int c = 0; c < 5; c++) mean[c] /= (double)n;
    for (size_t q = 0; q < n; q++)
        for (int c = 0; c < 5; c++) var[c] += (r[q * 5 + c] - mean[c]) * (r[q * 5 + c] - mean[c]);
    float nrm[16] = {0};
    for (int c = 0; c < 4; c++) {
        nrm[c] = (float)mean[c];
        const double sd = std::sqrt(var[c] / (double)n);
        nrm[4 + c] = (float)(sd > 1e-12 ? sd : 1.0);
    }
    nrm[8] = (float)mean[4];
    nrm[9] = (float)std::sqrt(var[4] / (double)n);
    float hull[8];
    for (int c = 0; c < 4; c++) { hull[c] = 1e30f; hull[4 + c] = -1e30f; }
    for (size_t q = 0; q < n; q++)
        for (int c = 0; c < 4; c++) {
            const float f = (float)((r[q * 5 + c] - nrm[c]) / nrm[4 + c]);
            hull[c] = std::min(hull[c], f);
            hull[4 + c] = std::max(hull[4 + c], f);
        }
    std::memcpy(nrm_h, nrm, sizeof nrm_h);
    std::memcpy(hull_h, hull, sizeof hull_h);
    MCK(hipMemcpy(nrm_d, nrm, sizeof nrm, hipMemcpyHostToDevice));
    MCK(hipMemcpy(hull_d, hull, sizeof hull, hipMemcpyHostToDevice));
    return 0;
}

int Mlp::upload(const double *feat_raw, const double *tau, size_t n) {
    if (ensure(n, 1)) return -1;
    std::vector<float> x(n * 4), y(n);
    for (size_t q = 0; q < n; q++) {
        float f[4];
        mlp_features(feat_raw[q * 4], feat_raw[q * 4 + 1], feat_raw[q * 4 + 2], feat_raw[q * 4 + 3], nrm_h, f);
        for (int c = 0; c < 4; c++) x[q * 4 + c] = f[c];
        if (tau) y[q] = (float)((std::log(tau[q]) - nrm_h[8]) / nrm_h[9]);
    }
    MCK(hipMemcpy(X, x.data(), x.size() * 4, hipMemcpyHostToDevice));
    if (tau) MCK(hipMemcpy(Yd, y.data(), y.size() * 4, hipMemcpyHostToDevice));
    ndata = n;
    return 0;
}

// one optimisation step (or gradient only) on samples [off, off+bs)
int Mlp::step(size_t off, int bs, float lr, bool update, float *grad_out, float *loss_out) {
    // 16 samples per wavefront: a 4096-sample batch spreads over 256 CUs
    const int chunks = 1;
    const int nw = (bs + 16 * chunks - 1) / (16 * chunks);
    if (ensure(ndata, nw)) return -1;
    hipLaunchKernelGGL(k_mlp<true>, dim3(nw), dim3(64), 0, stream, X + off * 4, Yd + off, bs, chunks,
                       1.0f / (float)bs, P, W2b, W3b, (float *)nullptr, gpart, lpart);
    float *gdev = nullptr;
    if (grad_out) {
        if (!gbuf) { MCK(hipMalloc(&gbuf, PLayout::N * 4)); owned.push_back(gbuf); }
        gdev = gbuf;
    }
    if (update) adam_t++;
    const float b1c = 1.f - std::pow(0.9f, (float)std::max(adam_t, 1));
    const float b2c = 1.f - std::pow(0.999f, (float)std::max(adam_t, 1));
    hipLaunchKernelGGL(k_mlp_adam, dim3((PLayout::N + 255) / 256), dim3(256), 0, stream, gpart, nw, P, m1, m2,
                       gdev, W2b, W3b, lr, b1c, b2c, update ? 1 : 0);
    if (loss_out) {
        hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(64), 0, stream, lpart, nw, loss_d);
        MCK(hipMemcpyAsync(loss_out, loss_d, 4, hipMemcpyDeviceToHost, stream));
    }
    if (grad_out) MCK(hipMemcpyAsync(grad_out, gdev, PLayout::N * 4, hipMemcpyDeviceToHost, stream));
    MCK(hipGetLastError());
    if (loss_out || grad_out) MCK(hipStreamSynchronize(stream));
    return 0;
}

// wall-model hook: network stress at every wall point inside the training
// box, inside flags for the ODE fallback, and the count of fallbacks (the
// caller zeroes it at the start of the wall-model evaluation)
void Mlp::wall_predict(int n, const double *u_les, double h, double nu, const double *dpdx_dev,
                       double *tau_pred, int *use_pred, Scalars *sc, hipStream_t st) {
    if (n <= 0) return;
    const MlpInputs in{u_les, nullptr, nullptr, nullptr, h, nu, dpdx_dev};
    hipLaunchKernelGGL(k_mlp_infer, dim3(mlp_infer_blocks(n)), dim3(MLP_INFER_WG), 0, st, n, in, P, W2b, W3b, nrm_d,
                       hull_d, 1, tau_pred, use_pred, &sc->n_fallback_mlp);
}

}  // namespace cf

// ===========================================================================
// extern "C"
// ===========================================================================
using cf::set_error;

struct cf_mlp {
    cf::Mlp m;
};
namespace cf {
Mlp *mlp_of(cf_mlp *p) { return &p->m; }
}

extern "C" {

int cf_mlp_create(unsigned seed, cf_mlp **out) {
    if (!out) { set_error("cf_mlp_create: null output"); return -1; }
    cf_mlp *p = new cf_mlp;
    if (p->m.init(seed)) { delete p; return -1; }
    *out = p;
    return 0;
}

int cf_mlp_destroy(cf_mlp *p) {
    delete p;
    return 0;
}

int cf_mlp_nparams(void) { return cf::PLayout::N; }

int cf_mlp_get_params(cf_mlp *p, float *w) {
    if (!p || !w) { set_error("cf_mlp_get_params: null argument"); return -1; }
    return p->m.get_params(w);
}

int cf_mlp_set_params(cf_mlp *p, const float *w) {
    if (!p || !w) { set_error("cf_mlp_set_params: null argument"); return -1; }
    return p->m.set_params(w);
}

int cf_mlp_get_norm(cf_mlp *p, float *nrm, float *hull) {
    if (!p) { set_error("cf_mlp_get_norm: null model"); return -1; }
    if (nrm) std::memcpy(nrm, p->m.nrm_h, sizeof p->m.nrm_h);
    if (hull) std::memcpy(hull, p->m.hull_h, sizeof p->m.hull_h);
    return 0;
}

int cf_mlp_set_norm(cf_mlp *p, const float *nrm, const float *hull) {
    if (!p || !nrm || !hull) { set_error("cf_mlp_set_norm: null argument"); return -1; }
    std::memcpy(p->m.nrm_h, nrm, sizeof p->m.nrm_h);
    std::memcpy(p->m.hull_h, hull, sizeof p->m.hull_h);
    if (hipMemcpy(p->m.nrm_d, nrm, sizeof p->m.nrm_h, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->m.hull_d, hull, sizeof p->m.hull_h, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("cf_mlp_set_norm: copy failed");
        return -1;
    }
    return 0;
}

int cf_mlp_train(cf_mlp *p, long long n, const double *feat, const double *tau, int epochs, int batch,
                 float lr, float lr_final, float *loss_history) {
    if (!p || !feat || !tau || n <= 0 || epochs <= 0 || batch < 16) {
        set_error("cf_mlp_train: bad arguments");
        return -1;
    }
    if (p->m.fit_normalisation(feat, tau, (size_t)n) || p->m.upload(feat, tau, (size_t)n)) return -1;
    const long long nb = n / batch;
    if (nb <= 0) { set_error("cf_mlp_train: fewer samples than one batch"); return -1; }
    const long long total = nb * epochs;
    long long it = 0;
    for (int e = 0; e < epochs; e++) {
        double acc = 0.0;
        for (long long b = 0; b < nb; b++, it++) {
            const float fr = (float)it / (float)std::max(total - 1, 1LL);
            const float lrv = lr * std::pow(lr_final / lr, fr);  // exponential decay
            float l = 0.f;
            const bool last = b == nb - 1;
            if (p->m.step((size_t)(b * batch), batch, lrv, true, nullptr, last ? &l : nullptr)) return -1;
            if (last) acc = l / (double)batch;
        }
        if (loss_history) loss_history[e] = (float)acc;
    }
    if (hipStreamSynchronize(p->m.stream) != hipSuccess) { set_error("cf_mlp_train: sync"); return -1; }
    return 0;
}

int cf_mlp_gradients(cf_mlp *p, int n, const double *feat, const double *tau, float *grad, float *loss) {
    if (!p || !feat || !tau || !grad || n < 1) { set_error("cf_mlp_gradients: bad arguments"); return -1; }
    if (p->m.upload(feat, tau, (size_t)n)) return -1;
    return p->m.step(0, n, 0.f, false, grad, loss);
}

int cf_mlp_predict(cf_mlp *p, long long n, const double *feat, double *tau, int *inside) {
    if (!p || !feat || !tau || n <= 0) { set_error("cf_mlp_predict: bad arguments"); return -1; }
    cf::Mlp &m = p->m;
    double *d = nullptr;
    int *di = nullptr;
    const size_t nn = (size_t)n;
    hipError_t e = hipMalloc(&d, nn * 5 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&di, nn * sizeof(int));
    std::vector<double> cols(nn * 4);
    for (size_t q = 0; q < nn; q++)
        for (int c = 0; c < 4; c++) cols[c * nn + q] = feat[q * 4 + c];
    if (e == hipSuccess) e = hipMemcpy(d, cols.data(), nn * 4 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const cf::MlpInputs in{d, d + nn, d + 2 * nn, d + 3 * nn, 0.0, 0.0, nullptr};
        hipLaunchKernelGGL(cf::k_mlp_infer, dim3(cf::mlp_infer_blocks((int)n)), dim3(cf::MLP_INFER_WG), 0, m.stream,
                           (int)n, in, m.P, m.W2b, m.W3b, m.nrm_d, m.hull_d, 0, d + 4 * nn, di, (int *)nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m.stream);
    if (e == hipSuccess) e = hipMemcpy(tau, d + 4 * nn, nn * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && inside) e = hipMemcpy(inside, di, nn * sizeof(int), hipMemcpyDeviceToHost);
    hipFree(d);
    hipFree(di);
    if (e != hipSuccess) { set_error(std::string("cf_mlp_predict: ") + hipGetErrorString(e)); return -1; }
    return 0;
}

}  // extern "C"
