// Learned distance bound: a Fourier-feature MLP on BF16 MFMA (gfx950).
//
//   x -> xhat = (x - c) / h              (scene box -> [-1, 1]^3)
//   f = [sin(B xhat), cos(B xhat)]       (B: 32 x 3, N(0, sigma^2) radians per
//                                         normalised unit, fixed)
//   z1 = W1 f + b1, a1 = silu(z1)        (64 -> 128)
//   z2..z4 likewise                      (128 -> 128, three layers)
//   o = w5 . a4 + b5                     (128 -> 1)
// The network regresses o ~ log(min(d / D, 1) + 1e-3) where d is a query's
// exact closest-silhouette distance and D the scene box diagonal; the seed
// radius is r0 = ((exp(o) - 1e-3) * (1 + rel) + abs) * D.  The exact
// traversal verifies every answer (query.hip), so the network only changes
// how fast a query runs, never its result.
//
// Kernels:
//   k_mlp_infer   out^T = W x^T per layer on v_mfma_f32_16x16x32_bf16 (gfx950's
//                 K = 32 BF16 MFMA): the weights are the A operand, staged in
//                 LDS once per block (112 KB) and shared by its 16 waves; the
//                 activations are the B operand and never leave registers: a
//                 layer's accumulator tiles are turned into the next layer's
//                 B fragments in place, because the hidden features are
//                 numbered so that the output layout of two 16-row tiles is
//                 the B layout of one 32-deep K block (see infer_index);
//                 each wave takes P tiles of 16 points, so every weight
//                 fragment read from LDS feeds P MFMAs (one wave = 16 P points)
//   (the training kernels use v_mfma_f32_16x16x16_bf16, one wave = 16 points)
//   k_mlp_train   forward + backward for a 64-point tile: weight gradients
//                 dW = delta^T a_prev and input deltas delta W both on MFMA,
//                 added into an FP32 gradient with one atomic per weight
//   k_mlp_adam    Adam on FP32 master weights; writes the BF16 copies in the
//                 two MFMA fragment orders (forward and transposed)
// Weight fragment order: for W (out x in), the 16x16 tile (nt, ks) occupies
// 512 contiguous bytes, lane l = 16 q + r holding W[16 nt + r][16 ks + 4 q + i]
// for i = 0..3, so each wave's operand load is one contiguous 512-B chunk.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

namespace snch {
namespace mlpk {

constexpr int NFREQ = 32, NIN = 64, H = 128;
constexpr int OFF_W[4] = {0, 8320, 24832, 41344};
constexpr int OFF_B[4] = {8192, 24704, 41216, 57728};
constexpr int KIN[4] = {64, 128, 128, 128};
constexpr int OFF_W5 = 57856, OFF_B5 = 57984, NPARAM = 57985;
constexpr int SW_OFF[4] = {0, 8192, 24576, 40960}; // offsets in the bf16 fragment arrays
constexpr int NSW = 57344;
constexpr float LOG_EPS = 1e-3f;
#ifndef MLP_INFER_WAVES
#define MLP_INFER_WAVES 16 // waves per inference block (they share one LDS copy of the weights)
#endif
#ifndef MLP_INFER_TILES
#define MLP_INFER_TILES 2 // 16-point tiles per wave at a time (reuse of each weight fragment)
#endif

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ unsigned short f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u); // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(unsigned short h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ v4f mma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
// D (16 x 16) += A (16 x 32) B (32 x 16); lane l = 16 g + r holds A[r][8 g + j],
// B[8 g + j][r] (j = 0..7) and D[4 g + i][r] (i = 0..3)
__device__ __forceinline__ v4f mma32(v8bf a, v8bf b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// fragment index of W_l[n][k] (forward B operand) and of the transposed use
__host__ __device__ __forceinline__ int fwd_index(int l, int n, int k)
{
    const int K = KIN[l];
    return SW_OFF[l] + ((((n >> 4) * (K >> 4) + (k >> 4)) * 64 + ((k & 15) >> 2) * 16 + (n & 15)) << 2) + (k & 3);
}
__host__ __device__ __forceinline__ int bwd_index(int l, int n, int k)
{
    return SW_OFF[l] + ((((k >> 4) * (H >> 4) + (n >> 4)) * 64 + ((n & 15) >> 2) * 16 + (k & 15)) << 2) + (n & 3);
}
// fragment index of W_l[n][k] for k_mlp_infer: the A operand of tile (mt, kb)
// (16 output rows x 32 inputs) is 64 lanes x 8 BF16, lane 16 g + r holding
// row 16 mt + r.  Layer 0 reads the features in order (position 8 g + j of
// block kb is input 32 kb + 8 g + j).  Layers 1-3 read the previous layer's
// outputs as its accumulators leave them: lane 16 g + r of output tiles 2 kb
// and 2 kb + 1 holds rows 4 g + i of each, so position 8 g + j of block kb is
// input 32 kb + 4 g + j (j < 4) or 32 kb + 16 + 4 g + j - 4 (j >= 4).
__host__ __device__ __forceinline__ int infer_index(int l, int n, int k)
{
    const int KB = KIN[l] >> 5, mt = n >> 4, r = n & 15, kb = k >> 5, kl = k & 31;
    int g, j;
    if (l == 0) {
        g = kl >> 3;
        j = kl & 7;
    } else {
        g = (kl & 15) >> 2;
        j = (kl & 3) + (kl >= 16 ? 4 : 0);
    }
    return SW_OFF[l] + (((mt * KB + kb) * 64 + 16 * g + r) << 3) + j;
}

struct Frame {
    float cx, cy, cz, ihx, ihy, ihz, invD;
    float tau; // loss asymmetry: weight 2 tau on under-prediction, 2 (1 - tau) on over-prediction
};

constexpr int ACT_STRIDE = 136; // padded bf16 row (128 + 8) against LDS bank conflicts

// features of one point row into an LDS tile row (lane handles 8 frequencies)
__device__ __forceinline__ void features(const float* __restrict__ fourier, const Frame& fr, float x, float y, float z,
                                         int q, unsigned short* row)
{
    const float xh = (x - fr.cx) * fr.ihx, yh = (y - fr.cy) * fr.ihy, zh = (z - fr.cz) * fr.ihz;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int fi = q * 8 + j;
        const float arg = fourier[3 * fi] * xh + fourier[3 * fi + 1] * yh + fourier[3 * fi + 2] * zh;
        float s, c;
        sincosf(arg, &s, &c);
        row[fi] = f2bf(s);
        row[NFREQ + fi] = f2bf(c);
    }
}

// ------------------------------------------------------------------ infer
// One hidden layer L for P tiles of 16 points: acc = W_L bop (+ bias, SiLU);
// the results become the next layer's B operand in place (L < 3) or are dotted
// with the output weights (L = 3: o = partial w5 . a4 of this lane's rows).
// A function per layer keeps every register array statically indexed.
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
// two floats as one word of two round-to-nearest-even BF16 halves (low first):
// one v_cvt_pk_bf16_f32 on gfx950
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v2f{lo, hi}, v2bf));
}
// SiLU for the inference path: the hardware exp and reciprocal (about 1 ulp,
// far inside the BF16 rounding its result gets)
__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

template <int L, int P>
__device__ __forceinline__ void infer_layer(const v4u* sW, const float* sB, v4u (&bop)[P][4], float (&o)[P],
                                            int lane, int g)
{
    constexpr int KB = KIN[L] >> 5;
    v4f acc[P][8];
#pragma unroll
    for (int pt = 0; pt < P; ++pt)
#pragma unroll
        for (int mt = 0; mt < 8; ++mt)
            acc[pt][mt] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int mt = 0; mt < 8; ++mt) {
            const v8bf a = __builtin_bit_cast(v8bf, sW[(SW_OFF[L] >> 3) + (mt * KB + kb) * 64 + lane]);
#pragma unroll
            for (int pt = 0; pt < P; ++pt)
                acc[pt][mt] = mma32(a, __builtin_bit_cast(v8bf, bop[pt][kb]), acc[pt][mt]);
        }
    // bias + SiLU; lane (g, r) holds outputs 16 mt + 4 g + i of point r
#pragma unroll
    for (int pt = 0; pt < P; ++pt) {
        float s = 0.0f;
#pragma unroll
        for (int mt = 0; mt < 8; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 16 * mt + 4 * g + i;
                acc[pt][mt][i] = silu_fast(acc[pt][mt][i] + sB[L * H + row]);
                if (L == 3)
                    s += sB[4 * H + row] * acc[pt][mt][i];
            }
        o[pt] = s;
        if (L < 3) { // tiles 2 kb and 2 kb + 1 become K block kb of the next layer
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) {
                const v4f a0 = acc[pt][2 * kb], a1 = acc[pt][2 * kb + 1];
                bop[pt][kb] = v4u{pack_bf16(a0[0], a0[1]), pack_bf16(a0[2], a0[3]), pack_bf16(a1[0], a1[1]),
                                  pack_bf16(a1[2], a1[3])};
            }
        }
    }
}

// mode 0: predicted distance; mode 1: seed radius.  NW waves per block share
// the LDS weights; each wave takes P tiles of 16 points at a time.
template <int NW, int P>
__global__ __launch_bounds__(NW * 64) void k_mlp_infer(const v4u* __restrict__ wisw, const float* __restrict__ theta,
                                                       const float* __restrict__ fourier, Frame fr,
                                                       const float* __restrict__ pts, uint32_t n, int mode, float rel,
                                                       float absm, float D, float* __restrict__ out)
{
    __shared__ v4u sW[NSW / 8];
    __shared__ float sB[4 * H + H + 1];
    __shared__ float sFour[3 * NFREQ];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, r = lane & 15, g = lane >> 4;
    for (int i = tid; i < NSW / 8; i += NW * 64)
        sW[i] = wisw[i];
    for (int i = tid; i < 4 * H; i += NW * 64)
        sB[i] = theta[OFF_B[i / H] + (i % H)];
    for (int i = tid; i < H; i += NW * 64)
        sB[4 * H + i] = theta[OFF_W5 + i];
    if (tid == 0)
        sB[5 * H] = theta[OFF_B5];
    for (int i = tid; i < 3 * NFREQ; i += NW * 64)
        sFour[i] = fourier[i];
    __syncthreads();
    const uint32_t ntiles = (n + 15) / 16;
    for (uint32_t t0 = ((uint32_t)blockIdx.x * NW + w) * P; t0 < ntiles; t0 += gridDim.x * NW * P) {
        // a compiler-level memory barrier: the weight fragments are the same in
        // every iteration, and without it the compiler hoists all 112 of their
        // LDS reads out of the loop (448 registers, spilled to scratch)
        asm volatile("" ::: "memory");
        // layer-0 B operand: features of point r of each tile, frequencies 8 g .. 8 g + 7
        v4u bop[P][4];
#pragma unroll
        for (int pt = 0; pt < P; ++pt) {
            const uint32_t p = (t0 + pt) * 16 + r;
            const bool valid = p < n;
            const float x = valid ? pts[3 * p] : 0.0f, y = valid ? pts[3 * p + 1] : 0.0f,
                        z = valid ? pts[3 * p + 2] : 0.0f;
            const float xh = (x - fr.cx) * fr.ihx, yh = (y - fr.cy) * fr.ihy, zh = (z - fr.cz) * fr.ihz;
            float sn[8], cs[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int fi = 8 * g + j;
                const float arg = sFour[3 * fi] * xh + sFour[3 * fi + 1] * yh + sFour[3 * fi + 2] * zh;
                sn[j] = __sinf(arg); // hardware sine and cosine (|arg| stays below ~40 rad):
                cs[j] = __cosf(arg); // errors far below the BF16 rounding of the features
            }
            bop[pt][0] = v4u{pack_bf16(sn[0], sn[1]), pack_bf16(sn[2], sn[3]), pack_bf16(sn[4], sn[5]),
                             pack_bf16(sn[6], sn[7])};
            bop[pt][1] = v4u{pack_bf16(cs[0], cs[1]), pack_bf16(cs[2], cs[3]), pack_bf16(cs[4], cs[5]),
                             pack_bf16(cs[6], cs[7])};
        }
        float o[P];
        infer_layer<0, P>(sW, sB, bop, o, lane, g);
        infer_layer<1, P>(sW, sB, bop, o, lane, g);
        infer_layer<2, P>(sW, sB, bop, o, lane, g);
        infer_layer<3, P>(sW, sB, bop, o, lane, g);
#pragma unroll
        for (int pt = 0; pt < P; ++pt) {
            float s = o[pt];
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            const uint32_t p = (t0 + pt) * 16 + r;
            if (g == 0 && p < n) {
                float d = __expf(s + sB[5 * H]) - LOG_EPS;
                d = d > 0.0f ? d : 0.0f;
                out[p] = mode == 0 ? d * D : (d * (1.0f + rel) + absm) * D;
            }
        }
    }
}

// ------------------------------------------------------------------ train
constexpr int TP = 64; // points per block

__global__ __launch_bounds__(256) void k_mlp_train(const unsigned short* __restrict__ wsw,
                                                   const unsigned short* __restrict__ wtsw,
                                                   const float* __restrict__ theta,
                                                   const float* __restrict__ fourier, Frame fr,
                                                   const float* __restrict__ pts, const float* __restrict__ dist,
                                                   uint32_t n, float inv_batch, float* __restrict__ grad,
                                                   float* __restrict__ loss_acc)
{
    __shared__ unsigned short sF[TP * ACT_STRIDE];     // features (64 used per row)
    __shared__ unsigned short sZ[4][TP * ACT_STRIDE];  // pre-activations z1..z4
    __shared__ unsigned short sD[2][TP * ACT_STRIDE];  // deltas, ping-pong
    __shared__ float sG[TP];
    __shared__ float sRed[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, r = lane & 15, q = lane >> 4;
    const uint32_t base = blockIdx.x * TP;
    {
        const uint32_t p = base + w * 16 + r;
        const bool valid = p < n;
        const float x = valid ? pts[3 * p] : 0.0f, y = valid ? pts[3 * p + 1] : 0.0f, z = valid ? pts[3 * p + 2] : 0.0f;
        features(fourier, fr, x, y, z, q, sF + (w * 16 + r) * ACT_STRIDE);
    }
    __syncthreads();
    // forward, wave w on rows 16w..16w+15
    for (int l = 0; l < 4; ++l) {
        const int K = KIN[l];
        v4f acc[8];
#pragma unroll
        for (int nt = 0; nt < 8; ++nt)
            acc[nt] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < K / 16; ++ks) {
            v4s a;
            if (l == 0) {
                a = *reinterpret_cast<const v4s*>(sF + (w * 16 + r) * ACT_STRIDE + ks * 16 + 4 * q);
            } else {
                const unsigned short* zr = sZ[l - 1] + (w * 16 + r) * ACT_STRIDE + ks * 16 + 4 * q;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a[i] = (short)f2bf(silu(bf2f(zr[i])));
            }
#pragma unroll
            for (int nt = 0; nt < 8; ++nt) {
                const v4s b = reinterpret_cast<const v4s*>(wsw)[(SW_OFF[l] >> 2) + (nt * (K >> 4) + ks) * 64 + lane];
                acc[nt] = mma(a, b, acc[nt]);
            }
        }
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            const int col = nt * 16 + r;
            const float bias = theta[OFF_B[l] + col];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                sZ[l][(w * 16 + 4 * q + i) * ACT_STRIDE + col] = f2bf(acc[nt][i] + bias);
        }
        __syncthreads();
    }
    // output, loss and dL/do
    {
        const int row = w * 16 + r;
        float s = 0.0f;
        for (int k = 32 * q; k < 32 * q + 32; ++k)
            s += theta[OFF_W5 + k] * silu(bf2f(sZ[3][row * ACT_STRIDE + k]));
        s += __shfl_xor(s, 16, 64);
        s += __shfl_xor(s, 32, 64);
        const uint32_t p = base + row;
        float err = 0.0f;
        if (p < n) {
            const float o = s + theta[OFF_B5];
            float dn = dist[p] * fr.invD;
            dn = dn < 1.0f ? dn : 1.0f; // +inf (no silhouette) saturates
            err = o - __logf(dn + LOG_EPS);
        }
        // expectile loss w err^2: w = 2 tau when the prediction is below the
        // target, 2 (1 - tau) above it (tau = 0.5: the plain squared error;
        // tau > 0.5 pushes the prediction up, towards an upper bound)
        const float wgt = 2.0f * (err < 0.0f ? fr.tau : 1.0f - fr.tau);
        if (q == 0)
            sG[row] = 2.0f * wgt * err * inv_batch;
        float e2 = q == 0 ? wgt * err * err : 0.0f;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            e2 += __shfl_xor(e2, o, 64);
        if (lane == 0)
            sRed[w] = e2;
    }
    __syncthreads();
    if (tid == 0)
        atomicAdd(loss_acc, sRed[0] + sRed[1] + sRed[2] + sRed[3]);
    // output layer gradients and delta4
    if (tid < H) {
        float s = 0.0f;
        for (int p = 0; p < TP; ++p)
            s += sG[p] * silu(bf2f(sZ[3][p * ACT_STRIDE + tid]));
        atomicAdd(grad + OFF_W5 + tid, s);
    } else if (tid == H) {
        float s = 0.0f;
        for (int p = 0; p < TP; ++p)
            s += sG[p];
        atomicAdd(grad + OFF_B5, s);
    }
    for (int e = tid; e < TP * H; e += 256) {
        const int p = e / H, k = e % H;
        const float z = bf2f(sZ[3][p * ACT_STRIDE + k]);
        sD[0][p * ACT_STRIDE + k] = f2bf(sG[p] * theta[OFF_W5 + k] * dsilu(z));
    }
    __syncthreads();
    int cur = 0;
    for (int l = 3; l >= 0; --l) {
        const int K = KIN[l];
        const unsigned short* dl = sD[cur];
        // dW_l = delta_l^T a_prev: wave w owns output-row tiles 2w, 2w+1
        for (int mt = 2 * w; mt < 2 * w + 2; ++mt) {
            for (int nt = 0; nt < K / 16; ++nt) {
                v4f acc = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < TP / 16; ++ks) {
                    v4s a, b;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int p = ks * 16 + 4 * q + i;
                        a[i] = (short)dl[p * ACT_STRIDE + mt * 16 + r];
                        const int kc = nt * 16 + r;
                        b[i] = l == 0 ? (short)sF[p * ACT_STRIDE + kc]
                                      : (short)f2bf(silu(bf2f(sZ[l - 1][p * ACT_STRIDE + kc])));
                    }
                    acc = mma(a, b, acc);
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    atomicAdd(grad + OFF_W[l] + (mt * 16 + 4 * q + i) * K + nt * 16 + r, acc[i]);
            }
        }
        if (tid < H) {
            float s = 0.0f;
            for (int p = 0; p < TP; ++p)
                s += bf2f(dl[p * ACT_STRIDE + tid]);
            atomicAdd(grad + OFF_B[l] + tid, s);
        }
        if (l > 0) {
            // delta_{l-1} = (delta_l W_l) * silu'(z_{l-1}): wave w on point rows 16w..
            unsigned short* dn = sD[cur ^ 1];
            v4f acc[8];
#pragma unroll
            for (int nt = 0; nt < 8; ++nt)
                acc[nt] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
            for (int ks = 0; ks < H / 16; ++ks) {
                const v4s a = *reinterpret_cast<const v4s*>(dl + (w * 16 + r) * ACT_STRIDE + ks * 16 + 4 * q);
#pragma unroll
                for (int nt = 0; nt < 8; ++nt) {
                    const v4s b = reinterpret_cast<const v4s*>(wtsw)[(SW_OFF[l] >> 2) + (nt * (H >> 4) + ks) * 64 + lane];
                    acc[nt] = mma(a, b, acc[nt]);
                }
            }
#pragma unroll
            for (int nt = 0; nt < 8; ++nt) {
                const int col = nt * 16 + r;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int row = w * 16 + 4 * q + i;
                    const float z = bf2f(sZ[l - 1][row * ACT_STRIDE + col]);
                    dn[row * ACT_STRIDE + col] = f2bf(acc[nt][i] * dsilu(z));
                }
            }
        }
        __syncthreads();
        cur ^= 1;
    }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; refreshes BF16 fragments
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ theta, float* __restrict__ grad,
                                                  float* __restrict__ m1, float* __restrict__ m2,
                                                  unsigned short* __restrict__ wsw, unsigned short* __restrict__ wtsw,
                                                  unsigned short* __restrict__ wisw, float lr, float b1, float b2,
                                                  float eps, float c1, float c2, int update)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= NPARAM)
        return;
    float th = theta[i];
    if (update) {
        const float g = grad[i];
        grad[i] = 0.0f;
        const float a = b1 * m1[i] + (1.0f - b1) * g;
        const float b = b2 * m2[i] + (1.0f - b2) * g * g;
        m1[i] = a;
        m2[i] = b;
        th -= lr * (a * c1) / (sqrtf(b * c2) + eps);
        theta[i] = th;
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const int K = KIN[l];
        if (i >= OFF_W[l] && i < OFF_W[l] + H * K) {
            const int n = (i - OFF_W[l]) / K, k = (i - OFF_W[l]) % K;
            const unsigned short h = f2bf(th);
            wsw[fwd_index(l, n, k)] = h;
            wtsw[bwd_index(l, n, k)] = h;
            wisw[infer_index(l, n, k)] = h;
        }
    }
}

__global__ void k_fill0(float* p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        p[i] = 0.0f;
}

} // namespace mlpk
} // namespace snch

using namespace snch;
using namespace snch::mlpk;

struct snch_mlp {
    snch_context* ctx = nullptr;
    Frame fr{};
    float D = 1.0f;
    DeviceBuffer theta, grad, m1, m2, wsw, wtsw, wisw, fourier, loss;
    std::vector<float> h_fourier;
    int step = 0;
};

namespace {

uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double u01(uint64_t& s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

int upload_params(snch_mlp* m, const float* params, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(m->theta.p, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, st);
    if (e != hipSuccess)
        return hip_fail(e, "snch_mlp: upload");
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, st, m->theta.as<float>(),
                       m->grad.as<float>(), m->m1.as<float>(), m->m2.as<float>(), m->wsw.as<unsigned short>(),
                       m->wtsw.as<unsigned short>(), m->wisw.as<unsigned short>(), 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
                       0);
    e = hipStreamSynchronize(st);
    return e == hipSuccess ? SNCH_OK : hip_fail(e, "snch_mlp: refresh");
}

} // namespace

extern "C" {

uint32_t snch_mlp_num_params(void) { return NPARAM; }

int snch_mlp_create(snch_context* ctx, const float* lo, const float* hi, float sigma, uint64_t seed, snch_mlp** out)
{
    if (!ctx || !lo || !hi || !out) {
        set_error("snch_mlp_create: null argument");
        return SNCH_ERR_INVALID;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess)
        return hip_fail(e, "hipSetDevice");
    auto* m = new snch_mlp();
    m->ctx = ctx;
    double d2 = 0.0;
    float c[3], ih[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = 0.5f * (lo[k] + hi[k]);
        const float h = 0.5f * (hi[k] - lo[k]);
        ih[k] = h > 0.0f ? 1.0f / h : 1.0f;
        d2 += (double)(hi[k] - lo[k]) * (hi[k] - lo[k]);
    }
    m->D = (float)std::sqrt(d2);
    m->fr = Frame{c[0], c[1], c[2], ih[0], ih[1], ih[2], 1.0f / m->D, 0.5f};
    // Fourier matrix B ~ N(0, sigma^2) (Box-Muller), and He-uniform weights
    uint64_t s = seed * 0x2545F4914F6CDD1Dull + 1;
    m->h_fourier.resize(3 * NFREQ);
    for (auto& v : m->h_fourier) {
        const double g = std::sqrt(-2.0 * std::log(u01(s))) * std::cos(2.0 * M_PI * u01(s));
        v = (float)(sigma * g);
    }
    std::vector<float> p(NPARAM, 0.0f);
    for (int l = 0; l < 4; ++l) {
        const double a = std::sqrt(6.0 / KIN[l]);
        for (int i = 0; i < H * KIN[l]; ++i)
            p[OFF_W[l] + i] = (float)((2.0 * u01(s) - 1.0) * a);
    }
    for (int i = 0; i < H; ++i)
        p[OFF_W5 + i] = (float)((2.0 * u01(s) - 1.0) * std::sqrt(3.0 / H));
    p[OFF_B5] = -2.0f;
    if ((e = m->theta.ensure(sizeof(float) * NPARAM)) != hipSuccess || (e = m->grad.ensure(sizeof(float) * NPARAM)) ||
        (e = m->m1.ensure(sizeof(float) * NPARAM)) || (e = m->m2.ensure(sizeof(float) * NPARAM)) ||
        (e = m->wsw.ensure(sizeof(unsigned short) * NSW)) || (e = m->wtsw.ensure(sizeof(unsigned short) * NSW)) ||
        (e = m->wisw.ensure(sizeof(unsigned short) * NSW)) ||
        (e = m->fourier.ensure(sizeof(float) * 3 * NFREQ)) || (e = m->loss.ensure(sizeof(float) * 4)) ||
        (e = hipMemcpy(m->fourier.p, m->h_fourier.data(), sizeof(float) * 3 * NFREQ, hipMemcpyHostToDevice))) {
        delete m;
        return hip_fail(e, "snch_mlp_create");
    }
    for (DeviceBuffer* b : {&m->grad, &m->m1, &m->m2, &m->loss})
        hipLaunchKernelGGL(k_fill0, dim3((b->bytes / 4 + 255) / 256), dim3(256), 0, 0, b->as<float>(),
                           (int)(b->bytes / 4));
    int rc = upload_params(m, p.data(), 0);
    if (rc != SNCH_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return SNCH_OK;
}

int snch_mlp_destroy(snch_mlp* m)
{
    delete m;
    return SNCH_OK;
}

int snch_mlp_params(snch_mlp* m, float* params_out, const float* params_in, float* fourier_out, float* frame_out)
{
    if (!m) {
        set_error("snch_mlp_params: null model");
        return SNCH_ERR_INVALID;
    }
    if (params_in) {
        int rc = upload_params(m, params_in, 0);
        if (rc != SNCH_OK)
            return rc;
    }
    if (params_out) {
        hipError_t e = hipMemcpy(params_out, m->theta.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            return hip_fail(e, "snch_mlp_params");
    }
    if (fourier_out)
        std::memcpy(fourier_out, m->h_fourier.data(), sizeof(float) * 3 * NFREQ);
    if (frame_out) {
        const float f[8] = {m->fr.cx, m->fr.cy, m->fr.cz, m->fr.ihx, m->fr.ihy, m->fr.ihz, m->fr.invD, m->D};
        std::memcpy(frame_out, f, sizeof(f));
    }
    return SNCH_OK;
}

int snch_mlp_set_loss(snch_mlp* m, float tau)
{
    if (!m || !(tau > 0.0f && tau < 1.0f)) {
        set_error("snch_mlp_set_loss: tau must lie in (0, 1)");
        return SNCH_ERR_INVALID;
    }
    m->fr.tau = tau;
    return SNCH_OK;
}

int snch_mlp_gradient(snch_mlp* m, const float* pts, const float* dist, uint32_t n, float* grad_out, float* loss_out,
                      void* stream)
{
    if (!m || !pts || !dist || n == 0) {
        set_error("snch_mlp_gradient: null argument or empty batch");
        return SNCH_ERR_INVALID;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_fill0, dim3((NPARAM + 255) / 256), dim3(256), 0, st, m->grad.as<float>(), NPARAM);
    hipLaunchKernelGGL(k_fill0, dim3(1), dim3(256), 0, st, m->loss.as<float>(), 4);
    hipLaunchKernelGGL(k_mlp_train, dim3((n + TP - 1) / TP), dim3(256), 0, st, m->wsw.as<unsigned short>(),
                       m->wtsw.as<unsigned short>(), m->theta.as<float>(), m->fourier.as<float>(), m->fr, pts, dist,
                       n, 1.0f / (float)n, m->grad.as<float>(), m->loss.as<float>());
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && grad_out)
        e = hipMemcpyAsync(grad_out, m->grad.p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, st);
    float l = 0.0f;
    if (e == hipSuccess)
        e = hipMemcpyAsync(&l, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, "snch_mlp_gradient");
    hipLaunchKernelGGL(k_fill0, dim3((NPARAM + 255) / 256), dim3(256), 0, st, m->grad.as<float>(), NPARAM);
    if (loss_out)
        *loss_out = l / (float)n;
    return SNCH_OK;
}

int snch_mlp_train(snch_mlp* m, const float* pts, const float* dist, uint32_t n, uint32_t batch, uint32_t epochs,
                   float lr, float* epoch_loss, void* stream)
{
    if (!m || !pts || !dist || n == 0 || batch == 0 || batch % TP != 0) {
        set_error("snch_mlp_train: null argument, empty data or batch not a multiple of 64");
        return SNCH_ERR_INVALID;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        hipLaunchKernelGGL(k_fill0, dim3(1), dim3(256), 0, st, m->loss.as<float>(), 4);
        for (uint32_t off = 0; off < n; off += batch) {
            const uint32_t nb = n - off < batch ? n - off : batch;
            hipLaunchKernelGGL(k_mlp_train, dim3((nb + TP - 1) / TP), dim3(256), 0, st, m->wsw.as<unsigned short>(),
                               m->wtsw.as<unsigned short>(), m->theta.as<float>(), m->fourier.as<float>(), m->fr,
                               pts + 3 * (size_t)off, dist + off, nb, 1.0f / (float)nb, m->grad.as<float>(),
                               m->loss.as<float>());
            ++m->step;
            const float c1 = 1.0f / (1.0f - std::pow(b1, (float)m->step));
            const float c2 = 1.0f / (1.0f - std::pow(b2, (float)m->step));
            hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, st, m->theta.as<float>(),
                               m->grad.as<float>(), m->m1.as<float>(), m->m2.as<float>(),
                               m->wsw.as<unsigned short>(), m->wtsw.as<unsigned short>(), m->wisw.as<unsigned short>(),
                               lr, b1, b2, eps, c1, c2, 1);
        }
        float l = 0.0f;
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(&l, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            return hip_fail(e, "snch_mlp_train");
        if (epoch_loss)
            epoch_loss[ep] = l / (float)n;
    }
    return SNCH_OK;
}

int snch_mlp_predict(snch_mlp* m, const float* pts, uint32_t n, int mode, float rel, float absm, float* out,
                     void* stream)
{
    if (!m || (n && (!pts || !out))) {
        set_error("snch_mlp_predict: null argument");
        return SNCH_ERR_INVALID;
    }
    if (n == 0)
        return SNCH_OK;
    constexpr int NW = MLP_INFER_WAVES, P = MLP_INFER_TILES;
    const uint32_t groups = (n + 16 * NW * P - 1) / (16 * NW * P);
    const uint32_t grid = groups < 2048 ? groups : 2048;
    hipLaunchKernelGGL((k_mlp_infer<NW, P>), dim3(grid), dim3(NW * 64), 0, static_cast<hipStream_t>(stream),
                       m->wisw.as<v4u>(), m->theta.as<float>(), m->fourier.as<float>(), m->fr, pts, n, mode, rel, absm,
                       m->D, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? SNCH_OK : hip_fail(e, "snch_mlp_predict");
}

} // extern "C"
