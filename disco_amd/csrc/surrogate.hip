// surrogate.hip -- the learned far image shell of the periodic sum.
//
// The field of a loop summed over the images with max(|i|, |j|, |k|) in
// (1, img] (the "far shell") is smooth in the atom position.  It is split into
//   * the far-shell sum of the loop's dipole term -- one loopon of the loop's
//     area at its centre (the Kroupa term, Code/kroupa.hpp:52-73), computed
//     exactly by the loopon kernels with an image shell;
//   * a residual, predicted by a small MLP:
//       features (14, padded to 16) -> 128 -> 128 -> 128 -> 3, hard swish,
//     scaled per loop by |b| (R / L)^4 (the residual of a flat loop about its
//     dipole scales with the fourth power of its size over the distance).
// Layer 1 runs in FP32 on the VALU (the atom offset as a BF16 input would be
// quantised to ~L/256).  Layers 2-4 are BF16 GEMMs on MFMA
// (v_mfma_f32_32x32x16_bf16; the split-K weight gradients on
// v_mfma_f32_32x32x8_bf16) with FP32 accumulation; FP32 master weights are
// updated by Adam; all of it by this file's kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "ctx_internal.h"
#include "disco_hip.h"

using disco::fail;

namespace {

typedef uint16_t bf16;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kIn = 16;    // 14 features + 2 zero pads
constexpr int kH = 128;
constexpr int kOut = 32;   // 3 outputs, padded to one 32-wide MFMA tile

// Flat FP32 parameter layout
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn;
constexpr int oW2 = oB1 + kH, oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH, oB3 = oW3 + kH * kH;
constexpr int oW4 = oB3 + kH, oB4 = oW4 + kOut * kH;
constexpr int kParams = oB4 + kOut;   // 39,328

// float -> bfloat16, round to nearest even (NaN kept quiet)
__host__ __device__ inline bf16 f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
// the same rounding (nearest even) by the hardware conversion, on the device
__device__ inline bf16 f2bf_d(float f)
{
    return __builtin_bit_cast(bf16, (__bf16)f);
}
__host__ __device__ inline float bf2f(bf16 h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
// The activation: hard swish, z clamp(z/6 + 1/2, 0, 1) -- an FMA with its clamp
// (one v_med3) and a multiply, no transcendental (the logistic of SiLU costs an
// exp and a reciprocal, which bounded the fused apply).  Derivative: 0 below
// -3, 1 above 3, z/3 + 1/2 between.
__device__ inline float hswish_gate(float z)
{
    return __builtin_amdgcn_fmed3f(fmaf(z, 1.0f / 6.0f, 0.5f), 0.0f, 1.0f);
}
__device__ inline float act_(float z) { return z * hswish_gate(z); }
__device__ inline float act_grad(float z)
{
    const float g = fmaf(z, 1.0f / 6.0f, 0.5f);
    return g <= 0.0f ? 0.0f : (g >= 1.0f ? 1.0f : fmaf(z, 1.0f / 3.0f, 0.5f));
}

// Layer-1 pre-activation of neuron j: the loop part first, in feature order
// 3..15, then the three offset features -- the order the fused apply uses with
// the loop part precomputed per loop, so both paths round identically.
__device__ inline float layer1_loop_part(const float* __restrict__ w, float b, const float* f)
{
    float z = b;
#pragma unroll
    for (int k = 3; k < kIn; k++) z = fmaf(f[k], w[k], z);
    return z;
}
__device__ inline float layer1_offsets(float c, float w0, float w1, float w2, float d0, float d1,
                                       float d2)
{
    return fmaf(d2, w2, fmaf(d1, w1, fmaf(d0, w0, c)));
}

// Features of one (atom, loop) pair; shared by the host feature entry point
// and the device kernels, so training and evaluation see the same inputs.
__host__ __device__ inline void pair_features(double px, double py, double pz,
                                              const disco_surrogate_loop& lp, const double* L,
                                              float* f)
{
    // offsets times the reciprocal box lengths: the fused apply takes each
    // reciprocal once per lane and rounds these features identically
    f[0] = (float)((px - lp.c[0]) * (1.0 / L[0]));
    f[1] = (float)((py - lp.c[1]) * (1.0 / L[1]));
    f[2] = (float)((pz - lp.c[2]) * (1.0 / L[2]));
    f[3] = (float)lp.n[0]; f[4] = (float)lp.n[1]; f[5] = (float)lp.n[2];
    const double bm = sqrt(lp.b[0] * lp.b[0] + lp.b[1] * lp.b[1] + lp.b[2] * lp.b[2]);
    const double ib = bm > 0 ? 1.0 / bm : 0.0;
    f[6] = (float)(lp.b[0] * ib); f[7] = (float)(lp.b[1] * ib); f[8] = (float)(lp.b[2] * ib);
    f[9] = (float)(lp.radius / 60.0);
    for (int k = 0; k < 4; k++) f[10 + k] = lp.shape == k ? 1.0f : 0.0f;
    f[14] = 0.0f;
    f[15] = 0.0f;
}

__host__ __device__ inline double pair_scale(const disco_surrogate_loop& lp, const double* L)
{
    const double bm = sqrt(lp.b[0] * lp.b[0] + lp.b[1] * lp.b[1] + lp.b[2] * lp.b[2]);
    const double Lm = (L[0] + L[1] + L[2]) / 3.0;
    const double r = lp.radius / Lm;
    return bm * r * r * r * r;
}

// ------------------------------------------------------------------ kernels
// Layer 1 in FP32: Z = X W1^T + b1, A = bf16(act(Z)); one thread per output.
__global__ void layer1_kernel(const float* __restrict__ X, const float* __restrict__ P,
                              float* __restrict__ Z, bf16* __restrict__ A, int M)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)M * kH) return;
    const int r = (int)(t / kH), j = (int)(t % kH);
    const float* w = P + oW1 + j * kIn;
    const float* x = X + (long long)r * kIn;
    const float z = layer1_offsets(layer1_loop_part(w, P[oB1 + j], x), w[0], w[1], w[2], x[0],
                                   x[1], x[2]);
    if (Z) Z[t] = z;
    A[t] = f2bf_d(act_(z));
}

// C (M x N) = A (M x K) . B (K x N) on MFMA, BF16 operands, FP32 accumulate.
// A(i, k) = A[i*sam + k*sak], B(k, j) = B[k*sbk + j*sbn].  One wave64 per
// 32 x 32 tile, K16 steps with the k of each operand slot as in the fused
// apply (so the forward and the apply round the same sums alike); grid (ceil(N/32), ceil(M/32)); out-of-range loads read zero
// and out-of-range outputs are not stored.
// EPI 0: C = acc (+ bias)                (fp32 out)
// EPI 1: Z = acc + bias, A = bf16(act)   (fp32 Z may be null, bf16 Ab)
template <int EPI>
__global__ __launch_bounds__(64) void gemm_bf16_kernel(
    const bf16* __restrict__ A, long long sam, long long sak, const bf16* __restrict__ B,
    long long sbk, long long sbn, int M, int N, int K, const float* __restrict__ bias,
    float* __restrict__ C, long long ldc, bf16* __restrict__ Ab, long long ldab)
{
    const int lane = threadIdx.x;
    const long long i0 = (long long)blockIdx.y * 32, j0 = (long long)blockIdx.x * 32;
    const long long ia = i0 + (lane & 31), jb = j0 + (lane & 31);
    const int kq = 4 * (lane >> 5);
    v16f acc;
    for (int t = 0; t < 16; t++) acc[t] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        v8s a, b;
        for (int t = 0; t < 8; t++) {
            const int k = k0 + kq + (t & 3) + 8 * (t >> 2);   // the fused apply's k order
            a[t] = (ia < M && k < K) ? (short)A[ia * sam + k * sak] : (short)0;
            b[t] = (jb < N && k < K) ? (short)B[k * sbk + jb * sbn] : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a),
                                                      __builtin_bit_cast(v8bf, b), acc, 0, 0, 0);
    }
    const long long jc = j0 + (lane & 31);
    if (jc >= N) return;
    const float bj = bias ? bias[jc] : 0.0f;
    for (int t = 0; t < 16; t++) {
        const long long ic = i0 + 8 * (t >> 2) + 4 * (lane >> 5) + (t & 3);
        if (ic >= M) continue;
        const float z = acc[t] + bj;
        if (EPI == 0) {
            C[ic * ldc + jc] = z;
        } else {
            if (C) C[ic * ldc + jc] = z;
            Ab[ic * ldab + jc] = f2bf_d(act_(z));
        }
    }
}

// dY = 2 (Y - T) / (3M) on the 3 real outputs (0 on the pads) as BF16; loss
// partial sums of (Y - T)^2 atomically added to *loss.
__global__ void loss_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                            bf16* __restrict__ dY, float* __restrict__ loss, int M)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.0f;
    if (t < (long long)M * kOut) {
        const int r = (int)(t / kOut), j = (int)(t % kOut);
        float g = 0.0f;
        if (j < 3) {
            const float d = Y[t] - T[r * 3 + j];
            e2 = d * d;
            g = 2.0f * d / (3.0f * (float)M);
        }
        dY[t] = f2bf(g);
    }
    // wave reduction, then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) e2 += __shfl_down(e2, o, 64);
    if ((threadIdx.x & 63) == 0 && e2 != 0.0f) atomicAdd(loss, e2 / (3.0f * (float)M));
}

// dZ = dA * act'(Z), as BF16 (and FP32 when dZf is given)
__global__ void act_bwd_kernel(const float* __restrict__ dA, const float* __restrict__ Z,
                               bf16* __restrict__ dZb, float* __restrict__ dZf, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float g = dA[t] * act_grad(Z[t]);
    if (dZb) dZb[t] = f2bf(g);
    if (dZf) dZf[t] = g;
}

// db[j] = sum_r D[r, j] (D given as BF16 or FP32), one thread per column
// Reductions over the batch rows in two deterministic stages: every block sums
// a slice of kSlice rows into its own partial, a second kernel adds the
// partials in block order.
constexpr int kSlice = 256;

// column sums of an M x N matrix (N <= 256, dividing 256): partial[b][j]
template <bool BF>
__global__ __launch_bounds__(256) void colsum_partial_kernel(const bf16* __restrict__ Db,
                                                             const float* __restrict__ Df,
                                                             float* __restrict__ part, int M,
                                                             int N)
{
    __shared__ float red[256];
    const int R = 256 / N, j = threadIdx.x % N, q = threadIdx.x / N;
    const long long r0 = (long long)blockIdx.x * kSlice;
    float sum = 0.0f;
    for (int r = q; r < kSlice; r += R) {
        const long long row = r0 + r;
        if (row < M) sum += BF ? bf2f(Db[row * N + j]) : Df[row * N + j];
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    if (q == 0) {
        float t = 0.0f;
        for (int k = 0; k < R; k++) t += red[k * N + j];
        part[(long long)blockIdx.x * N + j] = t;
    }
}

// layer-1 weight gradient dW1[j][k] = sum_r dZ[r][j] X[r][k] (128 x 16) over a
// slice: each thread 8 outputs (j = t / 2, k = 8 (t % 2) .. + 8)
__global__ __launch_bounds__(256) void layer1_wgrad_partial_kernel(const float* __restrict__ dZ,
                                                                   const float* __restrict__ X,
                                                                   float* __restrict__ part,
                                                                   int M)
{
    const int j = threadIdx.x >> 1, k0 = 8 * (threadIdx.x & 1);
    const long long r0 = (long long)blockIdx.x * kSlice;
    const long long r1 = r0 + kSlice < M ? r0 + kSlice : M;
    float acc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) acc[k] = 0.0f;
    for (long long r = r0; r < r1; r++) {
        const float d = dZ[r * kH + j];
        const float4 xa = *reinterpret_cast<const float4*>(X + r * kIn + k0);
        const float4 xb = *reinterpret_cast<const float4*>(X + r * kIn + k0 + 4);
        acc[0] += d * xa.x; acc[1] += d * xa.y; acc[2] += d * xa.z; acc[3] += d * xa.w;
        acc[4] += d * xb.x; acc[5] += d * xb.y; acc[6] += d * xb.z; acc[7] += d * xb.w;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) part[(long long)blockIdx.x * (kH * kIn) + j * kIn + k0 + k] = acc[k];
}

// out[t] = sum over b < nb of part[b][t], t < len (block order)
__global__ void sum_partials_kernel(const float* __restrict__ part, int nb, int len,
                                    float* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len) return;
    float s = 0.0f;
    for (int b = 0; b < nb; b++) s += part[(long long)b * len + t];
    out[t] = s;
}

// Split-K MFMA GEMM for the weight gradients (K = the batch): block z sums
// k in [z ks, min(K, (z + 1) ks)) of C = A . B into part[z] (M x N, row-major).
__global__ __launch_bounds__(64) void gemm_bf16_splitk_kernel(
    const bf16* __restrict__ A, long long sam, long long sak, const bf16* __restrict__ B,
    long long sbk, long long sbn, int M, int N, int K, int ks, float* __restrict__ part)
{
    const int lane = threadIdx.x;
    const long long i0 = (long long)blockIdx.y * 32, j0 = (long long)blockIdx.x * 32;
    const long long ia = i0 + (lane & 31), jb = j0 + (lane & 31);
    const int kq = 4 * (lane >> 5);
    const int kb = blockIdx.z * ks, ke = kb + ks < K ? kb + ks : K;
    v16f acc;
    for (int t = 0; t < 16; t++) acc[t] = 0.0f;
    for (int k0 = kb; k0 < ke; k0 += 8) {
        v4s a, b;
        for (int t = 0; t < 4; t++) {
            const int k = k0 + kq + t;
            a[t] = (ia < M && k < ke) ? (short)A[ia * sam + k * sak] : (short)0;
            b[t] = (jb < N && k < ke) ? (short)B[k * sbk + jb * sbn] : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    const long long jc = j0 + (lane & 31);
    if (jc >= N) return;
    float* P = part + (long long)blockIdx.z * M * N;
    for (int t = 0; t < 16; t++) {
        const long long ic = i0 + 8 * (t >> 2) + 4 * (lane >> 5) + (t & 3);
        if (ic < M) P[ic * N + jc] = acc[t];
    }
}

// Adam on the FP32 masters (the W4/b4 pads stay zero), then BF16 copies of
// W2, W3, W4 for the MFMA layers.
__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ G,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2, bf16* __restrict__ Wb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kParams) return;
    const bool pad = (t >= oW4 && t < oB4 && (t - oW4) / kH >= 3) || (t >= oB4 + 3);
    float p = P[t];
    if (!pad) {
        const float g = G[t];
        const float mt = b1 * m[t] + (1.0f - b1) * g;
        const float vt = b2 * v[t] + (1.0f - b2) * g * g;
        m[t] = mt;
        v[t] = vt;
        p -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
        P[t] = p;
    }
    if (t >= oW2 && t < oB2) Wb[t - oW2] = f2bf(p);
    else if (t >= oW3 && t < oB3) Wb[kH * kH + t - oW3] = f2bf(p);
    else if (t >= oW4 && t < oB4) Wb[2 * kH * kH + t - oW4] = f2bf(p);
}

__global__ void cast_weights_kernel(const float* __restrict__ P, bf16* __restrict__ Wb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= oB4) return;
    if (t >= oW2 && t < oB2) Wb[t - oW2] = f2bf(P[t]);
    else if (t >= oW3 && t < oB3) Wb[kH * kH + t - oW3] = f2bf(P[t]);
    else if (t >= oW4 && t < oB4) Wb[2 * kH * kH + t - oW4] = f2bf(P[t]);
}

// rows of a batch gathered from the dataset by index
__global__ void gather_kernel(const float* __restrict__ F, const float* __restrict__ T,
                              const int* __restrict__ idx, float* __restrict__ X,
                              float* __restrict__ Tb, int M)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M) return;
    const long long s = idx[t];
    for (int k = 0; k < kIn; k++) X[(long long)t * kIn + k] = F[s * kIn + k];
    for (int k = 0; k < 3; k++) Tb[t * 3 + k] = T[s * 3 + k];
}

// features of the pairs (atom a0 + r / nl, loop r % nl), r in [0, M)
__global__ void pair_features_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                     const double* __restrict__ z, long long a0,
                                     const disco_surrogate_loop* __restrict__ loops, int nl,
                                     double Lx, double Ly, double Lz, float* __restrict__ X, int M)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    const long long a = a0 + r / nl;
    const double L[3] = {Lx, Ly, Lz};
    pair_features(x[a], y[a], z[a], loops[r % nl], L, X + (long long)r * kIn);
}

// u[a] += gain * sum_l scale_l * Y[(a - a0) * nl + l, 0..2]
__global__ void far_out_kernel(const float* __restrict__ Y, const disco_surrogate_loop* __restrict__ loops,
                               int nl, double Lx, double Ly, double Lz, double gain, long long a0,
                               int na, double* __restrict__ ux, double* __restrict__ uy,
                               double* __restrict__ uz)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na) return;
    const double L[3] = {Lx, Ly, Lz};
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int l = 0; l < nl; l++) {
        const double s = gain * pair_scale(loops[l], L);
        const float* y = Y + ((long long)t * nl + l) * kOut;
        sx += s * y[0]; sy += s * y[1]; sz += s * y[2];
    }
    ux[a0 + t] += sx; uy[a0 + t] += sy; uz[a0 + t] += sz;
}

// ------------------------------------------------------------------ fused apply
// Per loop: the layer-1 pre-activation with the loop's own features (n, b, R,
// shape: constant for the loop) folded in, and the output scale gain |b|(R/L)^4.
__global__ void loop_consts_kernel(const float* __restrict__ P,
                                   const disco_surrogate_loop* __restrict__ loops, int nl,
                                   double Lx, double Ly, double Lz, double gain,
                                   float* __restrict__ cl, float* __restrict__ sl)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nl * kH) return;
    const int l = t / kH, j = t % kH;
    const double L[3] = {Lx, Ly, Lz};
    float f[kIn];
    pair_features(loops[l].c[0], loops[l].c[1], loops[l].c[2], loops[l], L, f);
    cl[t] = layer1_loop_part(P + oW1 + j * kIn, P[oB1 + j], f);
    if (j == 0) sl[l] = (float)(gain * pair_scale(loops[l], L));
}

// The fused apply runs its layers on v_mfma_f32_32x32x16_bf16: 8 bf16 of k per
// lane, k' = 8 (lane / 32) + i.  The B operand (activations) of one K16 step is
// two K8 groups of the previous accumulator, neurons 16t + 4h + i and
// 16t + 8 + 4h + i (h = lane / 32, i < 4); the weights are stored in LDS with
// bits 2 and 3 of the column swapped, so the matching A operand is the 16
// contiguous bytes at 16t + 8h.  Rows are padded to 136 entries (272 B = 68
// banks): the 8 lanes of one 128-bit read phase start 4 banks apart.
constexpr int kLd = kH + 8;
__device__ __forceinline__ int k16_col(int c) { return (c & ~12) | ((c & 4) << 1) | ((c & 8) >> 1); }

__device__ __forceinline__ v16f mfma_k16(const bf16* __restrict__ wrow, const v8s& b, v16f acc)
{
    const v8s a = *reinterpret_cast<const v8s*>(wrow);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a),
                                                   __builtin_bit_cast(v8bf, b), acc, 0, 0, 0);
}

// One hidden layer of the fused apply: hb (the B operands of its 8 k-steps)
// is replaced by the layer's output in the same layout.
__device__ __forceinline__ void hidden_layer(const bf16* __restrict__ W,
                                             const float* __restrict__ bias, int col, int h,
                                             v8s (&hb)[8])
{
    v8s hn[8];
#pragma unroll
    for (int np = 0; np < 4; np += 2) {
        v16f acc0, acc1;
        for (int t = 0; t < 16; t++) { acc0[t] = 0.0f; acc1[t] = 0.0f; }
        // raised wave priority while this wave issues its MFMA chain, so that
        // another wave's activations fill the SIMD meanwhile (61.1 against
        // 62.7 ms for the 4M-atom x 200-loop apply)
        __builtin_amdgcn_s_setprio(1);
#pragma unroll
        for (int s = 0; s < 8; s++) {
            acc0 = mfma_k16(W + (32 * np + col) * kLd + 16 * s + 8 * h, hb[s], acc0);
            acc1 = mfma_k16(W + (32 * np + 32 + col) * kLd + 16 * s + 8 * h, hb[s], acc1);
        }
        __builtin_amdgcn_s_setprio(0);
        // register t holds neuron 8(t/4) + 4h + t%4 of the tile: K16 group
        // 2np + t/8 (+2 for the second tile), half (t/4)&1, element t%4
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const int n0 = 32 * np + 8 * (t >> 2) + 4 * h + (t & 3);
            const int g = 2 * np + (t >> 3), e = 4 * ((t >> 2) & 1) + (t & 3);
            hn[g][e] = (short)f2bf_d(act_(acc0[t] + bias[n0]));
            hn[g + 2][e] = (short)f2bf_d(act_(acc1[t] + bias[n0 + 32]));
        }
    }
#pragma unroll
    for (int s = 0; s < 8; s++) hb[s] = hn[s];
}

// u[a] += sum over loops of scale_l * MLP(features(a, l)), the whole network in
// one kernel.  A wave holds 32 atoms (one per MFMA column) and walks the loops;
// each layer is computed transposed, out^T = W . h^T, with the weights as the
// MFMA A operand (from LDS) and the activations as B: the accumulator of one
// 32-neuron tile (lane: its column; register t: neuron 8(t/4) + 4(lane/32) +
// t%4) is exactly the B operand of the next layer's two K16 steps, so
// activations never leave registers.  Layer 1 (three offset features; the
// loop's features are in cl) runs on the VALU in FP32.
//
// Occupancy: the weights take 81.5 KB of LDS, so two workgroups fit a CU; with
// 8 waves per workgroup and at most 128 registers per lane (4 waves per SIMD,
// the launch bound) that is 16 waves per CU, twice the 4-wave workgroup's
// count -- the MFMA chains of one wave and the activations of another overlap.
constexpr int kApplyWaves = 8;                  // waves per workgroup, 32 atoms each
constexpr int kApplyThreads = 64 * kApplyWaves;
constexpr int kApplyAtoms = 32 * kApplyWaves;
__global__ __launch_bounds__(kApplyThreads, 4) void apply_fused_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    long long a0, int na, const disco_surrogate_loop* __restrict__ loops, int nl, double Lx,
    double Ly, double Lz, const float* __restrict__ P, const bf16* __restrict__ Wb,
    const float* __restrict__ cl, const float* __restrict__ sl, double* __restrict__ ux,
    double* __restrict__ uy, double* __restrict__ uz)
{
    __shared__ bf16 sW2[kH * kLd], sW3[kH * kLd], sW4[kOut * kLd];
    __shared__ float4 sW1[kH];
    __shared__ float sb2[kH], sb3[kH], sb4[kOut];
    for (int i = threadIdx.x; i < kH * kH; i += kApplyThreads) {
        const int r = i / kH, c = i % kH;
        const int p = r * kLd + k16_col(c);
        sW2[p] = Wb[i];
        sW3[p] = Wb[kH * kH + i];
        if (r < kOut) sW4[p] = Wb[2 * kH * kH + i];
    }
    for (int i = threadIdx.x; i < kH; i += kApplyThreads) {
        const float* w = P + oW1 + i * kIn;
        sW1[i] = make_float4(w[0], w[1], w[2], 0.0f);
        sb2[i] = P[oB2 + i];
        sb3[i] = P[oB3 + i];
        if (i < kOut) sb4[i] = P[oB4 + i];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
    const int ia = blockIdx.x * kApplyAtoms + (threadIdx.x >> 6) * 32 + col;   // atom of this column
    const bool valid = ia < na;
    const long long a = a0 + (valid ? ia : na - 1);
    const double px = x[a], py = y[a], pz = z[a];
    float ox = 0.0f, oy = 0.0f, oz = 0.0f;
    const double iLx = 1.0 / Lx, iLy = 1.0 / Ly, iLz = 1.0 / Lz;
    for (int l = 0; l < nl; l++) {
        // the weights stay in LDS: reloaded per loop, not hoisted into registers
        asm volatile("" ::: "memory");
        const disco_surrogate_loop& lp = loops[l];
        const float d0 = (float)((px - lp.c[0]) * iLx);
        const float d1 = (float)((py - lp.c[1]) * iLy);
        const float d2 = (float)((pz - lp.c[2]) * iLz);
        const float* c = cl + (long long)l * kH;
        // layer 1: this lane's 64 neurons 16s + 4h + i and 16s + 8 + 4h + i, as
        // the B operands of layer 2
        // (the loop constants are read one float4 ahead of their use)
        v8s hb[8];
        const float4* c4 = reinterpret_cast<const float4*>(c) + h;
        float4 cnext = c4[0];
#pragma unroll
        for (int s = 0; s < 8; s++) {
#pragma unroll
            for (int half = 0; half < 2; half++) {
                const int n = 16 * s + 8 * half + 4 * h;
                const float4 cc = cnext;
                if (2 * s + half < 15) cnext = c4[2 * (2 * s + half + 1)];
                const float cv[4] = {cc.x, cc.y, cc.z, cc.w};
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float4 w = sW1[n + i];
                    hb[s][4 * half + i] =
                        (short)f2bf_d(act_(layer1_offsets(cv[i], w.x, w.y, w.z, d0, d1, d2)));
                }
            }
        }
        // layers 2 and 3: 128 x 128, 32-neuron tiles two at a time, 8 k-steps
        hidden_layer(sW2, sb2, col, h, hb);
        hidden_layer(sW3, sb3, col, h, hb);
        // layer 4: 32 outputs (3 used), 8 k-steps
        v16f o4;
        for (int t = 0; t < 16; t++) o4[t] = 0.0f;
#pragma unroll
        for (int s = 0; s < 8; s++) o4 = mfma_k16(sW4 + col * kLd + 16 * s + 8 * h, hb[s], o4);
        const float sc = sl[l];
        ox = fmaf(sc, o4[0] + sb4[0], ox);
        oy = fmaf(sc, o4[1] + sb4[1], oy);
        oz = fmaf(sc, o4[2] + sb4[2], oz);
    }
    if (h == 0 && valid) {             // outputs 0..2 sit in lanes 0..31, registers 0..2
        ux[a] += ox; uy[a] += oy; uz[a] += oz;
    }
}

__global__ void zero_scalar_kernel(float* __restrict__ p) { if (threadIdx.x == 0) *p = 0.0f; }

inline unsigned blocks(long long n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------ state
struct disco_surrogate {
    disco_ctx* ctx = nullptr;
    float* P = nullptr;      // FP32 masters
    float* G = nullptr;      // gradients
    float* m = nullptr;
    float* v = nullptr;
    bf16* Wb = nullptr;      // BF16 W2 | W3 | W4
    float* loss = nullptr;
    long long step = 0;
    double gain = 1.0;       // output = gain * network
    // activation buffers for up to cap rows
    int cap = 0;
    float *X = nullptr, *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *Y = nullptr, *T = nullptr;
    float *dA = nullptr, *dZf = nullptr;
    float* part = nullptr;   // partial sums of the batch reductions
    bf16 *A1 = nullptr, *A2 = nullptr, *A3 = nullptr, *dY = nullptr, *dZ = nullptr;
};

namespace {

void free_acts(disco_surrogate* s)
{
    for (void* p : {(void*)s->X, (void*)s->Z1, (void*)s->Z2, (void*)s->Z3, (void*)s->Y, (void*)s->T,
                    (void*)s->dA, (void*)s->dZf, (void*)s->A1, (void*)s->A2, (void*)s->A3,
                    (void*)s->dY, (void*)s->dZ, (void*)s->part})
        if (p) (void)hipFree(p);
    s->X = s->Z1 = s->Z2 = s->Z3 = s->Y = s->T = s->dA = s->dZf = s->part = nullptr;
    s->A1 = s->A2 = s->A3 = s->dY = s->dZ = nullptr;
    s->cap = 0;
}

int ensure_acts(disco_surrogate* s, int rows)
{
    disco_ctx* ctx = s->ctx;
    if (rows <= s->cap) return DISCO_OK;
    free_acts(s);
    const size_t r = (size_t)rows;
    DISCO_HIP_TRY(ctx, hipMalloc(&s->X, r * kIn * 4));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->Z1, r * kH * 4));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->Z2, r * kH * 4));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->Z3, r * kH * 4));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->Y, r * kOut * 4));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->T, r * 3 * 4));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->dA, r * kH * 4));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->dZf, r * kH * 4));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->A1, r * kH * 2));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->A2, r * kH * 2));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->A3, r * kH * 2));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->dY, r * kOut * 2));
    DISCO_HIP_TRY(ctx, hipMalloc(&s->dZ, r * kH * 2));
    const size_t slices = (r + kSlice - 1) / kSlice;
    DISCO_HIP_TRY(ctx, hipMalloc(&s->part, slices * kH * kH * 4));
    s->cap = rows;
    return DISCO_OK;
}

template <int EPI>
void gemm(hipStream_t st, const bf16* A, long long sam, long long sak, const bf16* B,
          long long sbk, long long sbn, int M, int N, int K, const float* bias, float* C,
          long long ldc, bf16* Ab, long long ldab)
{
    const dim3 grid(blocks(N, 32), blocks(M, 32));
    hipLaunchKernelGGL((gemm_bf16_kernel<EPI>), grid, dim3(64), 0, st, A, sam, sak, B, sbk, sbn,
                       M, N, K, bias, C, ldc, Ab, ldab);
}

// Weight gradient C (M x N, ldc = N) = A . B with K = the batch, split over
// slices of kSlice rows and summed in slice order.
void gemm_wgrad(disco_surrogate* s, hipStream_t st, const bf16* A, long long sam, long long sak,
                const bf16* B, long long sbk, long long sbn, int M, int N, int K, float* C)
{
    const int S = (K + kSlice - 1) / kSlice;
    const dim3 grid(blocks(N, 32), blocks(M, 32), (unsigned)S);
    hipLaunchKernelGGL(gemm_bf16_splitk_kernel, grid, dim3(64), 0, st, A, sam, sak, B, sbk, sbn,
                       M, N, K, kSlice, s->part);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(blocks((long long)M * N, 256)), dim3(256), 0, st,
                       (const float*)s->part, S, M * N, C);
}

// Column sums of an M x N batch matrix (bf16 or float) into db[N].
void colsum(disco_surrogate* s, hipStream_t st, const bf16* Db, const float* Df, float* db,
            int M, int N)
{
    const int S = (M + kSlice - 1) / kSlice;
    if (Db)
        hipLaunchKernelGGL(colsum_partial_kernel<true>, dim3(S), dim3(256), 0, st, Db, Df,
                           s->part, M, N);
    else
        hipLaunchKernelGGL(colsum_partial_kernel<false>, dim3(S), dim3(256), 0, st, Db, Df,
                           s->part, M, N);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, st, (const float*)s->part, S,
                       N, db);
}

// Forward of M rows already in s->X; keeps Z1..Z3 when train.
void forward(disco_surrogate* s, int M, bool train, hipStream_t st)
{
    const bf16* W2 = s->Wb;
    const bf16* W3 = s->Wb + kH * kH;
    const bf16* W4 = s->Wb + 2 * kH * kH;
    hipLaunchKernelGGL(layer1_kernel, dim3(blocks((long long)M * kH, 256)), dim3(256), 0, st,
                       s->X, s->P, train ? s->Z1 : nullptr, s->A1, M);
    // Y = A W^T: A(i,k) = A[i*K + k], B(k,j) = W[j*K + k]
    gemm<1>(st, s->A1, kH, 1, W2, 1, kH, M, kH, kH, s->P + oB2, train ? s->Z2 : nullptr, kH,
            s->A2, kH);
    gemm<1>(st, s->A2, kH, 1, W3, 1, kH, M, kH, kH, s->P + oB3, train ? s->Z3 : nullptr, kH,
            s->A3, kH);
    gemm<0>(st, s->A3, kH, 1, W4, 1, kH, M, kOut, kH, s->P + oB4, s->Y, kOut, nullptr, 0);
}

// Gradient of the mean squared error of M rows (targets in s->T) into s->G.
void backward(disco_surrogate* s, int M, hipStream_t st)
{
    const bf16* W2 = s->Wb;
    const bf16* W3 = s->Wb + kH * kH;
    const bf16* W4 = s->Wb + 2 * kH * kH;
    hipLaunchKernelGGL(zero_scalar_kernel, dim3(1), dim3(64), 0, st, s->loss);
    hipLaunchKernelGGL(loss_kernel, dim3(blocks((long long)M * kOut, 256)), dim3(256), 0, st,
                       s->Y, s->T, s->dY, s->loss, M);
    // layer 4: dW4 = dY^T A3 (32 x 128), db4, dA3 = dY W4 (M x 128)
    gemm_wgrad(s, st, s->dY, 1, kOut, s->A3, kH, 1, kOut, kH, M, s->G + oW4);
    colsum(s, st, s->dY, nullptr, s->G + oB4, M, kOut);
    gemm<0>(st, s->dY, kOut, 1, W4, kH, 1, M, kH, kOut, nullptr, s->dA, kH, nullptr, 0);
    // layer 3
    hipLaunchKernelGGL(act_bwd_kernel, dim3(blocks((long long)M * kH, 256)), dim3(256), 0, st,
                       s->dA, s->Z3, s->dZ, (float*)nullptr, (long long)M * kH);
    gemm_wgrad(s, st, s->dZ, 1, kH, s->A2, kH, 1, kH, kH, M, s->G + oW3);
    colsum(s, st, s->dZ, nullptr, s->G + oB3, M, kH);
    gemm<0>(st, s->dZ, kH, 1, W3, kH, 1, M, kH, kH, nullptr, s->dA, kH, nullptr, 0);
    // layer 2
    hipLaunchKernelGGL(act_bwd_kernel, dim3(blocks((long long)M * kH, 256)), dim3(256), 0, st,
                       s->dA, s->Z2, s->dZ, (float*)nullptr, (long long)M * kH);
    gemm_wgrad(s, st, s->dZ, 1, kH, s->A1, kH, 1, kH, kH, M, s->G + oW2);
    colsum(s, st, s->dZ, nullptr, s->G + oB2, M, kH);
    gemm<0>(st, s->dZ, kH, 1, W2, kH, 1, M, kH, kH, nullptr, s->dA, kH, nullptr, 0);
    // layer 1 (FP32)
    hipLaunchKernelGGL(act_bwd_kernel, dim3(blocks((long long)M * kH, 256)), dim3(256), 0, st,
                       s->dA, s->Z1, (bf16*)nullptr, s->dZf, (long long)M * kH);
    {
        const int S = (M + kSlice - 1) / kSlice;
        hipLaunchKernelGGL(layer1_wgrad_partial_kernel, dim3(S), dim3(256), 0, st,
                           (const float*)s->dZf, (const float*)s->X, s->part, M);
        hipLaunchKernelGGL(sum_partials_kernel, dim3(blocks(kH * kIn, 256)), dim3(256), 0, st,
                           (const float*)s->part, S, kH * kIn, s->G + oW1);
    }
    colsum(s, st, nullptr, s->dZf, s->G + oB1, M, kH);
}

int upload_rows(disco_surrogate* s, const float* feats, const float* targets, int n)
{
    disco_ctx* ctx = s->ctx;
    std::vector<float> X((size_t)n * kIn, 0.0f);
    for (int r = 0; r < n; r++)
        for (int k = 0; k < DISCO_SURR_IN; k++) X[(size_t)r * kIn + k] = feats[(size_t)r * DISCO_SURR_IN + k];
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(s->X, X.data(), X.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    if (targets)
        DISCO_HIP_TRY(ctx, hipMemcpyAsync(s->T, targets, (size_t)n * 3 * 4, hipMemcpyHostToDevice, ctx->stream));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DISCO_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

int disco_hip_surrogate_num_params(void) { return kParams; }

int disco_hip_surrogate_features(const double* p, int n, const disco_surrogate_loop* loop,
                                 const double* L, float* out)
{
    if (n < 0 || (n > 0 && (!p || !loop || !L || !out))) return DISCO_ERR_ARG;
    for (int r = 0; r < n; r++) {
        float f[kIn];
        pair_features(p[3 * r], p[3 * r + 1], p[3 * r + 2], *loop, L, f);
        std::memcpy(out + (size_t)r * DISCO_SURR_IN, f, DISCO_SURR_IN * sizeof(float));
    }
    return DISCO_OK;
}

double disco_hip_surrogate_scale(const disco_surrogate_loop* loop, const double* L)
{
    return (loop && L) ? pair_scale(*loop, L) : 0.0;
}

int disco_hip_surrogate_create(disco_ctx* ctx, unsigned seed, disco_surrogate** out)
{
    if (!ctx || !out) return DISCO_ERR_ARG;
    *out = nullptr;
    disco_surrogate* s = new (std::nothrow) disco_surrogate();
    if (!s) return fail(ctx, DISCO_ERR_OOM, "host allocation");
    s->ctx = ctx;
    auto cleanup = [&](int rc) { disco_hip_surrogate_free(s); return rc; };
    if (hipSetDevice(ctx->device) != hipSuccess) return cleanup(fail(ctx, DISCO_ERR_HIP, "hipSetDevice"));
    for (float** p : {&s->P, &s->G, &s->m, &s->v})
        if (hipMalloc(p, kParams * 4) != hipSuccess) return cleanup(fail(ctx, DISCO_ERR_OOM, "parameters"));
    if (hipMalloc(&s->Wb, 3 * kH * kH * 2) != hipSuccess || hipMalloc(&s->loss, 4) != hipSuccess)
        return cleanup(fail(ctx, DISCO_ERR_OOM, "parameters"));
    // He-uniform initial weights, zero biases, zero pads
    std::vector<float> P(kParams, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int rows, int cols, int live_rows) {
        const float a = std::sqrt(6.0f / (float)cols);
        std::uniform_real_distribution<float> U(-a, a);
        for (int r = 0; r < rows; r++)
            for (int c = 0; c < cols; c++) P[off + r * cols + c] = r < live_rows ? U(rng) : 0.0f;
    };
    fill(oW1, kH, kIn, kH);
    for (int r = 0; r < kH; r++) P[oW1 + r * kIn + 14] = P[oW1 + r * kIn + 15] = 0.0f;
    fill(oW2, kH, kH, kH);
    fill(oW3, kH, kH, kH);
    fill(oW4, kOut, kH, 3);
    const int rc = disco_hip_surrogate_set_params(s, P.data());
    if (rc != DISCO_OK) return cleanup(rc);
    if (hipMemset(s->m, 0, kParams * 4) != hipSuccess || hipMemset(s->v, 0, kParams * 4) != hipSuccess)
        return cleanup(fail(ctx, DISCO_ERR_HIP, "optimizer state"));
    *out = s;
    return DISCO_OK;
}

int disco_hip_surrogate_free(disco_surrogate* s)
{
    if (!s) return DISCO_OK;
    free_acts(s);
    for (void* p : {(void*)s->P, (void*)s->G, (void*)s->m, (void*)s->v, (void*)s->Wb, (void*)s->loss})
        if (p) (void)hipFree(p);
    delete s;
    return DISCO_OK;
}

int disco_hip_surrogate_get_params(disco_surrogate* s, float* out)
{
    if (!s || !out) return DISCO_ERR_ARG;
    DISCO_HIP_TRY(s->ctx, hipMemcpy(out, s->P, kParams * 4, hipMemcpyDeviceToHost));
    return DISCO_OK;
}

int disco_hip_surrogate_set_params(disco_surrogate* s, const float* in)
{
    if (!s || !in) return DISCO_ERR_ARG;
    disco_ctx* ctx = s->ctx;
    DISCO_HIP_TRY(ctx, hipMemcpy(s->P, in, kParams * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(cast_weights_kernel, dim3(blocks(oB4, 256)), dim3(256), 0, ctx->stream,
                       s->P, s->Wb);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DISCO_OK;
}

int disco_hip_surrogate_forward(disco_surrogate* s, const float* feats, int n, float* out)
{
    if (!s || n < 0 || (n > 0 && (!feats || !out))) return DISCO_ERR_ARG;
    if (n == 0) return DISCO_OK;
    disco_ctx* ctx = s->ctx;
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = ensure_acts(s, n);
    if (rc == DISCO_OK) rc = upload_rows(s, feats, nullptr, n);
    if (rc != DISCO_OK) return rc;
    forward(s, n, false, ctx->stream);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    std::vector<float> Y((size_t)n * kOut);
    DISCO_HIP_TRY(ctx, hipMemcpyAsync(Y.data(), s->Y, Y.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < n; r++)
        for (int k = 0; k < 3; k++) out[(size_t)r * 3 + k] = (float)(s->gain * Y[(size_t)r * kOut + k]);
    return DISCO_OK;
}

int disco_hip_surrogate_set_gain(disco_surrogate* s, double gain)
{
    if (!s || !std::isfinite(gain)) return DISCO_ERR_ARG;
    s->gain = gain;
    return DISCO_OK;
}

double disco_hip_surrogate_get_gain(const disco_surrogate* s) { return s ? s->gain : 0.0; }

int disco_hip_surrogate_loss_grad(disco_surrogate* s, const float* feats, const float* targets,
                                  int n, float* loss, float* grads)
{
    if (!s || n <= 0 || !feats || !targets) return DISCO_ERR_ARG;
    disco_ctx* ctx = s->ctx;
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    int rc = ensure_acts(s, n);
    if (rc == DISCO_OK) rc = upload_rows(s, feats, targets, n);
    if (rc != DISCO_OK) return rc;
    forward(s, n, true, ctx->stream);
    backward(s, n, ctx->stream);
    DISCO_HIP_TRY(ctx, hipGetLastError());
    if (loss) DISCO_HIP_TRY(ctx, hipMemcpyAsync(loss, s->loss, 4, hipMemcpyDeviceToHost, ctx->stream));
    if (grads) DISCO_HIP_TRY(ctx, hipMemcpyAsync(grads, s->G, kParams * 4, hipMemcpyDeviceToHost, ctx->stream));
    DISCO_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return DISCO_OK;
}

int disco_hip_surrogate_train(disco_surrogate* s, const float* feats, const float* targets, int n,
                              int epochs, int batch, float lr, unsigned seed, float* final_loss)
{
    if (!s || n <= 0 || !feats || !targets || epochs < 0 || batch <= 0 || !(lr > 0.f))
        return DISCO_ERR_ARG;
    disco_ctx* ctx = s->ctx;
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    batch = std::min(batch, n);
    int rc = ensure_acts(s, batch);
    if (rc != DISCO_OK) return rc;
    // dataset on the device
    float *dF = nullptr, *dT = nullptr;
    int* dI = nullptr;
    std::vector<float> F((size_t)n * kIn, 0.0f);
    for (int r = 0; r < n; r++)
        for (int k = 0; k < DISCO_SURR_IN; k++) F[(size_t)r * kIn + k] = feats[(size_t)r * DISCO_SURR_IN + k];
    auto done = [&](int code) {
        if (dF) (void)hipFree(dF);
        if (dT) (void)hipFree(dT);
        if (dI) (void)hipFree(dI);
        return code;
    };
    if (hipMalloc(&dF, F.size() * 4) != hipSuccess || hipMalloc(&dT, (size_t)n * 12) != hipSuccess ||
        hipMalloc(&dI, (size_t)n * 4) != hipSuccess)
        return done(fail(ctx, DISCO_ERR_OOM, "training set"));
    if (hipMemcpy(dF, F.data(), F.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dT, targets, (size_t)n * 12, hipMemcpyHostToDevice) != hipSuccess)
        return done(fail(ctx, DISCO_ERR_HIP, "training set upload"));
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    const long long steps_total = (long long)epochs * (n / batch);
    long long k = 0;
    float last = 0.0f;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; e++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        if (hipMemcpy(dI, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess)
            return done(fail(ctx, DISCO_ERR_HIP, "permutation upload"));
        for (int b0 = 0; b0 + batch <= n; b0 += batch, k++) {
            hipLaunchKernelGGL(gather_kernel, dim3(blocks(batch, 256)), dim3(256), 0, st, dF, dT,
                               dI + b0, s->X, s->T, batch);
            forward(s, batch, true, st);
            backward(s, batch, st);
            // cosine decay from lr to lr / 100
            const double frac = steps_total > 1 ? (double)k / (double)(steps_total - 1) : 0.0;
            const float lrk = (float)(lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(3.141592653589793 * frac))));
            s->step++;
            const float c1 = 1.0f - std::pow(b1, (float)s->step);
            const float c2 = 1.0f - std::pow(b2, (float)s->step);
            hipLaunchKernelGGL(adam_kernel, dim3(blocks(kParams, 256)), dim3(256), 0, st, s->P,
                               s->G, s->m, s->v, lrk, b1, b2, eps, c1, c2, s->Wb);
        }
        if (hipGetLastError() != hipSuccess) return done(fail(ctx, DISCO_ERR_HIP, "training kernels"));
    }
    if (hipMemcpyAsync(&last, s->loss, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return done(fail(ctx, DISCO_ERR_HIP, "training sync"));
    if (final_loss) *final_loss = last;
    return done(DISCO_OK);
}

int disco_hip_surrogate_apply(disco_surrogate* s, int natoms, const double* x, const double* y,
                              const double* z, double* ux, double* uy, double* uz,
                              const disco_surrogate_loop* loops, int nloops, const double* L,
                              int chunk_rows, void* stream)
{
    if (!s || natoms < 0 || nloops < 0 || !L) return DISCO_ERR_ARG;
    disco_ctx* ctx = s->ctx;
    if (natoms == 0 || nloops == 0) return DISCO_OK;
    if (!x || !y || !z || !ux || !uy || !uz || !loops) return fail(ctx, DISCO_ERR_ARG, "null array");
    DISCO_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    // atoms per launch: all of them unless the caller bounds a launch's length
    const int apc = chunk_rows <= 0 ? natoms : std::max(1, chunk_rows / nloops);
    // loop table, then per loop the layer-1 part of its own features and its scale
    char* buf = nullptr;
    const size_t lb = (size_t)nloops * sizeof(disco_surrogate_loop);
    const size_t cb = (size_t)nloops * kH * sizeof(float);
    DISCO_HIP_TRY(ctx, hipMalloc(&buf, lb + cb + (size_t)nloops * sizeof(float) + 256));
    disco_surrogate_loop* dl = reinterpret_cast<disco_surrogate_loop*>(buf);
    float* cl = reinterpret_cast<float*>(buf + ((lb + 255) / 256) * 256);
    float* sl = cl + (size_t)nloops * kH;
    if (hipMemcpy(dl, loops, lb, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(buf);
        return fail(ctx, DISCO_ERR_HIP, "loop upload");
    }
    hipLaunchKernelGGL(loop_consts_kernel, dim3(blocks((long long)nloops * kH, 256)), dim3(256),
                       0, st, s->P, dl, nloops, L[0], L[1], L[2], s->gain, cl, sl);
    for (long long a0 = 0; a0 < natoms; a0 += apc) {
        const int na = (int)std::min<long long>(apc, natoms - a0);
        hipLaunchKernelGGL(apply_fused_kernel, dim3(blocks(na, kApplyAtoms)), dim3(kApplyThreads),
                           0, st, x, y, z,
                           a0, na, dl, nloops, L[0], L[1], L[2], s->P, s->Wb, cl, sl, ux, uy,
                           uz);
    }
    const hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(st);
    (void)hipFree(buf);
    if (e != hipSuccess) return fail(ctx, DISCO_ERR_HIP, std::string("apply: ") + hipGetErrorString(e));
    return DISCO_OK;
}

}  // extern "C"
