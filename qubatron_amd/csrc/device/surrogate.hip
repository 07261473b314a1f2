/*
 * surrogate.hip - the learned shadow-visibility surrogate (no reference
 * counterpart: the reference traces every shadow ray,
 * src/qubatron/shaders/octree_fsh.c:426-434).
 *
 *   training   feature kernel -> one fused kernel per minibatch: forward
 *              (bias + ReLU + bf16 between layers), loss and output
 *              gradient, backward through the masked transposed weights and
 *              the weight and bias gradients, 64 samples per wave on the
 *              matrix cores -> sum of the per-block partial gradients ->
 *              Adam on fp32 master weights -> bf16 repack
 *   inference  one fused kernel: weights in LDS, 16 queries per wave, the
 *              four layers chained in registers (mlp.hpp)
 *   render     the decision pass of QBT_RENDER_SURROGATE: predicted bits
 *              outside the band, the rest appended to per-light lists that
 *              the exact shadow kernel traces
 *   collect    training samples from exact renders (QBT_RENDER_COLLECT)
 *
 * Activations are feature-major ([feature][sample], bf16) so that every
 * layer is C = W H with the batch as the N dimension.
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <math.h>

#include <rocprim/rocprim.hpp>

#include "mlp.hpp"
#include "pixel.hpp"
#include "qbt.h"
#include "qbt_internal.hpp"
#include "trace.hpp"

#pragma clang fp contract(off)

using namespace qbt;

#define HIPCHK(x) QBT_HIPCHK(x)

static_assert(kMlpParams == QBT_MLP_PARAMS, "parameter count");
static_assert(kMlpIn == QBT_MLP_INPUTS, "input count");

/* training sample: hit position, normal, light position, label (1 = lit) */
struct RawSample {
    float isp[3], nrm[3], light[3], label;
};
static_assert(sizeof(RawSample) == 40, "sample record");

/* fp32 gradient layout: packed weight matrices (as kP*), then the biases */
constexpr int kGB1 = kPackN, kGB2 = kGB1 + 64, kGB3 = kGB2 + 64, kGB4 = kGB3 + 64, kGradN = kGB4 + 1;
/* biases in LDS for the fused kernels */
constexpr int kLB1 = 0, kLB2 = 64, kLB3 = 128, kLB4 = 192, kLBN = 196;

struct MlpState {
    DevBuf   params, m, v, grads, packed, ctr; /* ctr: collected, lit (u64) */
    DevBuf   samples, perm, keys, vals, sort_tmp;
    DevBuf   X, gpart, lossbuf, io_x, io_p, io_y, io_a, io_c, flat; /* gpart: k_mlp_train's partial gradients */
    size_t   cap = 0;
    uint32_t stride = 1, phase = 0;
    int      t     = 0; /* Adam steps taken */
    float    band_lo = 0.1f, band_hi = 0.9f;
    bool     ready   = false; /* parameters initialised */
    uint64_t epoch   = 0;
    size_t   pos     = 0, perm_n = 0; /* next batch in the permutation of perm_n samples */
};

namespace {

inline unsigned nblk(size_t n, unsigned bs = 256) { return (unsigned) ((n + bs - 1) / bs); }

__global__ void k_fill_f32(float* p, size_t n, float v)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, int n)
{
    int i = threadIdx.x;
    if (i < n) p[i] = 0ull;
}

/* splitmix64 finaliser of a counter */
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* He-uniform weights (bound sqrt(6 / fan_in)), zero biases */
__global__ void k_mlp_init(float* P, uint64_t seed)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    float a = 0.0f;
    if (i < kB1) a = sqrtf(6.0f / 42.0f);
    else if ((i >= kW2 && i < kB2) || (i >= kW3 && i < kB3) || (i >= kW4 && i < kB4)) a = sqrtf(6.0f / 64.0f);
    float u = (float) (mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t) i) >> 40) * (1.0f / 16777216.0f);
    P[i]    = a > 0.0f ? (2.0f * u - 1.0f) * a : 0.0f;
}

/* fp32 parameters -> padded bf16 weight matrices */
__global__ void k_mlp_pack(const float* __restrict__ P, uint16_t* __restrict__ W)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kPackN) return;
    float v = 0.0f;
    if (i < kP2) {
        int o = i / 48, c = i % 48;
        v     = c < kMlpIn ? P[kW1 + o * kMlpIn + c] : 0.0f;
    } else if (i < kP3) {
        v = P[kW2 + (i - kP2)];
    } else if (i < kP4) {
        v = P[kW3 + (i - kP3)];
    } else {
        int r = (i - kP4) / 64, c = (i - kP4) % 64;
        v     = r == 0 ? P[kW4 + c] : 0.0f;
    }
    W[i] = f2bf(v);
}

/* ------------------------------------------------------------------------ */
/* fused training forward and backward                                      */
/* ------------------------------------------------------------------------ */

constexpr int kChunk = 64; /* samples per chunk: four fragment groups of 16, one per wave */

/* one chunk's activations and gradients, [feature][kChunk] bf16 */
struct TrainLds {
    uint16_t x[kMlpInPad * kChunk];
    uint16_t h1[64 * kChunk], h2[64 * kChunk], h3[64 * kChunk];
    uint16_t d1[64 * kChunk], d2[64 * kChunk], d3[64 * kChunk], d4[16 * kChunk];
    float    bias[kLBN];
};

/* A fragment of W (row-major, `in` columns): rows 16 mt.., columns 16 kb.. */
__device__ __forceinline__ bf16x4 wfrag(const uint16_t* __restrict__ W, int in, int mt, int kb, int r, int g)
{
    return *reinterpret_cast<const bf16x4*>(W + (16 * mt + r) * in + 16 * kb + 4 * g);
}

/* A fragment of W^T: element (i, k) = W[16 kb + k][16 mt + i] */
__device__ __forceinline__ bf16x4 wtfrag(const uint16_t* __restrict__ W, int in, int mt, int kb, int r, int g)
{
    bf16x4 v;
    for (int e = 0; e < 4; e++) v[e] = (short) W[(16 * kb + 4 * g + e) * in + 16 * mt + r];
    return v;
}

/* D fragment (features 16 t + 4 g + e of sample r) of group j into a
 * [feature][kChunk] LDS array */
__device__ __forceinline__ void put_frag(uint16_t* __restrict__ a, int t, int j, bf16x4 v, int r, int g)
{
    for (int e = 0; e < 4; e++) a[(16 * t + 4 * g + e) * kChunk + 16 * j + r] = (uint16_t) v[e];
}

/* one 16 x 16 tile of D = A B^T over the chunk's 64 samples (K): A rows
 * 16 mt.. of a, B rows 16 nt.. of b, both [row][kChunk] in LDS */
__device__ __forceinline__ f32x4 chunk_tile(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b, int mt,
                                            int nt, int r, int g)
{
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int j = 0; j < kChunk / 16; j++) {
        const bf16x4 av = *reinterpret_cast<const bf16x4*>(a + (16 * mt + r) * kChunk + 16 * j + 4 * g);
        const bf16x4 bv = *reinterpret_cast<const bf16x4*>(b + (16 * nt + r) * kChunk + 16 * j + 4 * g);
        acc             = mfma16(av, bv, acc);
    }
    return acc;
}

/* the gradient tile (rows 16 mt + 4 g + e, column 16 nt + r of a matrix
 * with `in` columns at offset off) into the block's row of partial sums */
__device__ __forceinline__ void add_tile(float* __restrict__ gp, bool first, int off, int in, int mt, int nt, f32x4 v,
                                         int r, int g)
{
    for (int e = 0; e < 4; e++) {
        float* q = gp + off + (16 * mt + 4 * g + e) * in + 16 * nt + r;
        *q       = first ? v[e] : *q + v[e];
    }
}

__device__ __forceinline__ bf16x4 relu_pack(f32x4 acc, const float* __restrict__ b)
{
    return pack4(acc[0] + b[0] > 0.0f ? acc[0] + b[0] : 0.0f, acc[1] + b[1] > 0.0f ? acc[1] + b[1] : 0.0f,
                 acc[2] + b[2] > 0.0f ? acc[2] + b[2] : 0.0f, acc[3] + b[3] > 0.0f ? acc[3] + b[3] : 0.0f);
}

/* dZ = (W^T dZ_next) where the layer's ReLU output is positive, bf16 */
__device__ __forceinline__ bf16x4 mask_pack(f32x4 acc, bf16x4 h)
{
    return pack4(h[0] > 0 ? acc[0] : 0.0f, h[1] > 0 ? acc[1] : 0.0f, h[2] > 0 ? acc[2] : 0.0f,
                 h[3] > 0 ? acc[3] : 0.0f);
}

/* the weights as matrix-core A fragments, forward (W) and backward (W^T),
 * held in registers for a whole launch */
struct TrainFrags {
    bf16x4 w1[4][3], w2[4][4], w3[4][4], w4[4]; /* W:   rows 16 mt.., columns 16 kb.. */
    bf16x4 t2[4][4], t3[4][4], t4[4];           /* W^T: rows 16 mt.., columns 16 kb.. */
};

__device__ __forceinline__ void load_train_frags(const uint16_t* __restrict__ W, int r, int g, TrainFrags& F)
{
    for (int mt = 0; mt < 4; mt++) {
        for (int kb = 0; kb < 3; kb++) F.w1[mt][kb] = wfrag(W + kP1, 48, mt, kb, r, g);
        for (int kb = 0; kb < 4; kb++) {
            F.w2[mt][kb] = wfrag(W + kP2, 64, mt, kb, r, g);
            F.w3[mt][kb] = wfrag(W + kP3, 64, mt, kb, r, g);
            F.t2[mt][kb] = wtfrag(W + kP2, 64, mt, kb, r, g);
            F.t3[mt][kb] = wtfrag(W + kP3, 64, mt, kb, r, g);
        }
        F.w4[mt] = wfrag(W + kP4, 64, 0, mt, r, g);  /* the output row block, K block mt */
        F.t4[mt] = wtfrag(W + kP4, 64, mt, 0, r, g); /* W4^T rows 16 mt.., its 16 columns */
    }
}

/* Forward and backward of the B columns of X (n valid; labels Y), four
 * waves per block. Block b takes the 64-sample chunks b, b + gridDim.x, ...;
 * per chunk wave w runs fragment group w (16 samples) through the four
 * layers as matrix-core fragments (bias, ReLU and bf16 rounding between
 * layers, the rounding points of the oracle), the loss and output gradient
 * (bf16((p - y) / n)), and the backward pass through the transposed weights
 * masked by the ReLU outputs, writing every activation and gradient to LDS;
 * then the four waves share the chunk's 48 weight-gradient tiles (16 x 16,
 * the 64 samples as K) and the bias row sums, added to the block's own row
 * of gp (kGradN floats, no atomics; k_grad_rows sums the rows). The summed
 * BCE goes to *loss. */
__global__ void __launch_bounds__(256) k_mlp_train(const uint16_t* __restrict__ W, const float* __restrict__ P,
                                                   const uint16_t* __restrict__ X, const float* __restrict__ Y, int B,
                                                   int n, float* __restrict__ gp, double* __restrict__ loss)
{
    __shared__ TrainLds L;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, r = lane & 15, g = lane >> 4;
    for (int i = tid; i < kLBN; i += blockDim.x) {
        float b = 0.0f;
        if (i < 64) b = P[kB1 + i];
        else if (i < 128) b = P[kB2 + i - 64];
        else if (i < 192) b = P[kB3 + i - 128];
        else if (i == kLB4) b = P[kB4];
        L.bias[i] = b;
    }
    TrainFrags F;
    load_train_frags(W, r, g, F);
    float* const  row    = gp + (size_t) blockIdx.x * kGradN;
    const int     chunks = (B + kChunk - 1) / kChunk;
    double        lsum   = 0.0;
    const bf16x4  zero4  = {0, 0, 0, 0};
    for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
        const bool first = c == (int) blockIdx.x;
        const int  base  = c * kChunk;
        for (int i = tid; i < kMlpInPad * kChunk; i += blockDim.x) {
            const int f = i / kChunk, s = base + i % kChunk;
            L.x[i]      = s < B ? X[(size_t) f * B + s] : (uint16_t) 0;
        }
        __syncthreads();
        const int j = wv; /* this wave's fragment group */
        if (base + 16 * j >= B) { /* past the batch: zeros, which add nothing below */
            for (int t = 0; t < 4; t++) {
                put_frag(L.h1, t, j, zero4, r, g), put_frag(L.h2, t, j, zero4, r, g);
                put_frag(L.h3, t, j, zero4, r, g), put_frag(L.d1, t, j, zero4, r, g);
                put_frag(L.d2, t, j, zero4, r, g), put_frag(L.d3, t, j, zero4, r, g);
            }
            put_frag(L.d4, 0, j, zero4, r, g);
        } else {
            /* forward */
            bf16x4 xin[3], h1[4], h2[4], h3[4];
            for (int kb = 0; kb < 3; kb++)
                for (int e = 0; e < 4; e++) xin[kb][e] = (short) L.x[(16 * kb + 4 * g + e) * kChunk + 16 * j + r];
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                for (int kb = 0; kb < 3; kb++) acc = mfma16(F.w1[mt][kb], xin[kb], acc);
                h1[mt] = relu_pack(acc, L.bias + kLB1 + 16 * mt + 4 * g);
                put_frag(L.h1, mt, j, h1[mt], r, g);
            }
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                for (int kb = 0; kb < 4; kb++) acc = mfma16(F.w2[mt][kb], h1[kb], acc);
                h2[mt] = relu_pack(acc, L.bias + kLB2 + 16 * mt + 4 * g);
                put_frag(L.h2, mt, j, h2[mt], r, g);
            }
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                for (int kb = 0; kb < 4; kb++) acc = mfma16(F.w3[mt][kb], h2[kb], acc);
                h3[mt] = relu_pack(acc, L.bias + kLB3 + 16 * mt + 4 * g);
                put_frag(L.h3, mt, j, h3[mt], r, g);
            }
            f32x4 lg = {0.f, 0.f, 0.f, 0.f};
            for (int kb = 0; kb < 4; kb++) lg = mfma16(F.w4[kb], h3[kb], lg);
            /* loss and output gradient: row 0 of the output tile, lanes 0..15 */
            const int s  = base + 16 * j + r;
            float     dz = 0.0f;
            if (g == 0 && s < n) {
                const float p  = 1.0f / (1.0f + expf(-(lg[0] + L.bias[kLB4])));
                const float y  = Y[s];
                const float pc = fminf(fmaxf(p, 1e-7f), 1.0f - 1e-7f);
                lsum += y > 0.5f ? -log((double) pc) : -log(1.0 - (double) pc);
                dz = (p - y) / (float) n;
            }
            bf16x4 d4 = zero4;
            d4[0]     = g == 0 ? (short) f2bf(dz) : (short) 0; /* row 0 of dZ4; rows 1..15 zero */
            put_frag(L.d4, 0, j, d4, r, g);
            /* backward */
            bf16x4 d3[4], d2[4];
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc       = mfma16(F.t4[mt], d4, acc);
                d3[mt]    = mask_pack(acc, h3[mt]);
                put_frag(L.d3, mt, j, d3[mt], r, g);
            }
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                for (int kb = 0; kb < 4; kb++) acc = mfma16(F.t3[mt][kb], d3[kb], acc);
                d2[mt] = mask_pack(acc, h2[mt]);
                put_frag(L.d2, mt, j, d2[mt], r, g);
            }
            for (int mt = 0; mt < 4; mt++) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                for (int kb = 0; kb < 4; kb++) acc = mfma16(F.t2[mt][kb], d2[kb], acc);
                put_frag(L.d1, mt, j, mask_pack(acc, h1[mt]), r, g);
            }
        }
        __syncthreads();
        /* weight gradients of the chunk, dW_l = dZ_l H_{l-1}^T: 48 tiles, wave w takes w, w + 4, ... */
        for (int t = wv; t < 48; t += 4) {
            if (t < 12) {
                const int mt = t / 3, nt = t % 3;
                add_tile(row, first, kP1, 48, mt, nt, chunk_tile(L.d1, L.x, mt, nt, r, g), r, g);
            } else if (t < 28) {
                const int mt = (t - 12) / 4, nt = (t - 12) % 4;
                add_tile(row, first, kP2, 64, mt, nt, chunk_tile(L.d2, L.h1, mt, nt, r, g), r, g);
            } else if (t < 44) {
                const int mt = (t - 28) / 4, nt = (t - 28) % 4;
                add_tile(row, first, kP3, 64, mt, nt, chunk_tile(L.d3, L.h2, mt, nt, r, g), r, g);
            } else {
                const int nt = t - 44;
                add_tile(row, first, kP4, 64, 0, nt, chunk_tile(L.d4, L.h3, 0, nt, r, g), r, g);
            }
        }
        /* bias gradients: row sums, wave w takes layer w + 1 (lane = feature) */
        if (wv < 3 || lane == 0) {
            const uint16_t* d   = wv == 0 ? L.d1 : wv == 1 ? L.d2 : wv == 2 ? L.d3 : L.d4;
            const int       off = wv == 0 ? kGB1 : wv == 1 ? kGB2 : wv == 2 ? kGB3 : kGB4;
            float           acc = 0.0f;
            for (int k = 0; k < kChunk; k++) acc += bf2f(d[lane * kChunk + k]);
            row[off + lane] = first ? acc : row[off + lane] + acc;
        }
        __syncthreads(); /* the next chunk overwrites the LDS arrays */
    }
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (lane == 0 && lsum != 0.0) atomicAdd(loss, lsum);
}

/* G[i] = sum over the rows of the partial gradients */
__global__ void k_grad_rows(const float* __restrict__ gp, int rows, float* __restrict__ G)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGradN) return;
    float acc = 0.0f;
    for (int b = 0; b < rows; b++) acc += gp[(size_t) b * kGradN + i];
    G[i] = acc;
}

/* the 48 (padded) features of a query, bf16, feature-major X[48][B] */
__device__ __forceinline__ void store_features(uint16_t* __restrict__ X, int B, int s, const float u[3],
                                               const float nrm[3], const float lu[3])
{
    for (int f = 0; f < kMlpInPad; f++) X[(size_t) f * B + s] = f2bf(mlp_feature(f, u, nrm, lu));
}

/* minibatch from the sample store: X, labels Y */
__global__ void k_batch_features(const RawSample* __restrict__ S, const uint32_t* __restrict__ perm, size_t off,
                                 int B, float bx, float by, float bz, float bw, uint16_t* __restrict__ X,
                                 float* __restrict__ Y)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    const RawSample q = S[perm[off + (size_t) s]];
    float           u[3], lu[3];
    mlp_scale(q.isp, bx, by, bz, bw, u);
    mlp_scale(q.light, bx, by, bz, bw, lu);
    store_features(X, B, s, u, q.nrm, lu);
    Y[s] = q.label;
}

/* supplied feature rows (n x 42 fp32) -> X; columns n..B-1 zero */
__global__ void k_copy_features(const float* __restrict__ x, int B, int n, uint16_t* __restrict__ X,
                                const float* __restrict__ yin, float* __restrict__ Y)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    for (int f = 0; f < kMlpInPad; f++)
        X[(size_t) f * B + s] = f2bf((s < n && f < kMlpIn) ? x[(size_t) s * kMlpIn + f] : 0.0f);
    if (Y) Y[s] = (s < n && yin) ? yin[s] : 0.0f;
}

/* flat parameter index -> gradient-layout index */
__device__ __forceinline__ int grad_index(int i)
{
    if (i < kB1) return kP1 + (i / kMlpIn) * 48 + (i % kMlpIn);
    if (i < kW2) return kGB1 + (i - kB1);
    if (i < kB2) return kP2 + (i - kW2);
    if (i < kW3) return kGB2 + (i - kB2);
    if (i < kB3) return kP3 + (i - kW3);
    if (i < kW4) return kGB3 + (i - kB3);
    if (i < kB4) return kP4 + (i - kW4);
    return kGB4;
}

/* Adam (Kingma & Ba 2015), bias-corrected moments, fp32 master weights */
__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    float g  = G[grad_index(i)];
    m[i]     = b1 * m[i] + (1.0f - b1) * g;
    v[i]     = b2 * v[i] + (1.0f - b2) * g * g;
    float mh = m[i] / c1;
    float vh = v[i] / c2;
    P[i]     = P[i] - lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_grads_to_flat(const float* __restrict__ G, float* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpParams) out[i] = G[grad_index(i)];
}

/* random sort keys for the sample shuffle */
__global__ void k_perm_keys(uint32_t* keys, uint32_t* vals, size_t n, uint64_t seed)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = (uint32_t) mix64(seed * 0x9E3779B97F4A7C15ull + i);
    vals[i] = (uint32_t) i;
}

/* ------------------------------------------------------------------------ */
/* fused inference                                                          */
/* ------------------------------------------------------------------------ */

struct InferArgs {
    const uint16_t* W; /* packed bf16 weights */
    const float*    P; /* fp32 parameters (the biases are read) */
    float           bx, by, bz, bw;
};

/* the biases in LDS (every lane of a row group reads the same four) */
__device__ __forceinline__ void load_biases_lds(const InferArgs& ia, float* Bs)
{
    for (int i = threadIdx.x; i < kLBN; i += blockDim.x) {
        float b = 0.0f;
        if (i < 64) b = ia.P[kB1 + i];
        else if (i < 128) b = ia.P[kB2 + i - 64];
        else if (i < 192) b = ia.P[kB3 + i - 128];
        else if (i == kLB4) b = ia.P[kB4];
        Bs[i] = b;
    }
    __syncthreads();
}

/* All weights as matrix-core A fragments, held in registers for the whole
 * kernel (48 fragments, 96 VGPRs): lane l holds W[16 mt + l % 16][16 kb +
 * 4 (l / 16) + e]. */
struct MlpFrags {
    bf16x4 w1[4][3], w2[4][4], w3[4][4], w4[4];
};

__device__ __forceinline__ void load_frags(const uint16_t* __restrict__ W, int lane, MlpFrags& F)
{
    const int r = lane & 15, g = lane >> 4;
    for (int mt = 0; mt < 4; mt++) {
        for (int kb = 0; kb < 3; kb++)
            F.w1[mt][kb] = *reinterpret_cast<const bf16x4*>(W + kP1 + (16 * mt + r) * 48 + 16 * kb + 4 * g);
        for (int kb = 0; kb < 4; kb++) {
            F.w2[mt][kb] = *reinterpret_cast<const bf16x4*>(W + kP2 + (16 * mt + r) * 64 + 16 * kb + 4 * g);
            F.w3[mt][kb] = *reinterpret_cast<const bf16x4*>(W + kP3 + (16 * mt + r) * 64 + 16 * kb + 4 * g);
        }
    }
    for (int kb = 0; kb < 4; kb++) F.w4[kb] = *reinterpret_cast<const bf16x4*>(W + kP4 + r * 64 + 16 * kb + 4 * g);
}

/* B fragments of the input of 16 queries: lane l holds features 16 kb +
 * 4 (l / 16) + e of query l % 16 */
__device__ __forceinline__ void query_frag(const float u[3], const float nrm[3], const float lu[3], int g,
                                           bf16x4 xin[3])
{
    for (int kb = 0; kb < 3; kb++) {
        float f[4];
        for (int p = 0; p < 2; p++) {
            const int fi = 16 * kb + 4 * g + 2 * p; /* even */
            if (fi < 36) {
                const int k = fi / 6, a = (fi % 6) >> 1;
                band_sincos(k, pick3(u, a), &f[2 * p], &f[2 * p + 1]);
            } else {
                f[2 * p]     = mlp_feature(fi, u, nrm, lu);
                f[2 * p + 1] = mlp_feature(fi + 1, u, nrm, lu);
            }
        }
        xin[kb] = pack4(f[0], f[1], f[2], f[3]);
    }
}

__device__ __forceinline__ float relu(float v) { return v > 0.0f ? v : 0.0f; }

/* logit of the 16 queries of the fragments (valid in lanes 0..15) */
__device__ __forceinline__ float mlp_logit(const MlpFrags& F, const float* __restrict__ Bs, const bf16x4 xin[3],
                                           int lane)
{
    const int g = lane >> 4;
    bf16x4    h[4], h2[4];
    for (int mt = 0; mt < 4; mt++) { /* layer 1: 64 x 48 */
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < 3; kb++) acc = mfma16(F.w1[mt][kb], xin[kb], acc);
        const float4 b = *reinterpret_cast<const float4*>(Bs + kLB1 + 16 * mt + 4 * g);
        h[mt] = pack4(relu(acc[0] + b.x), relu(acc[1] + b.y), relu(acc[2] + b.z),
                      relu(acc[3] + b.w));
    }
    for (int mt = 0; mt < 4; mt++) { /* layer 2 */
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < 4; kb++) acc = mfma16(F.w2[mt][kb], h[kb], acc);
        const float4 b = *reinterpret_cast<const float4*>(Bs + kLB2 + 16 * mt + 4 * g);
        h2[mt] = pack4(relu(acc[0] + b.x), relu(acc[1] + b.y), relu(acc[2] + b.z),
                       relu(acc[3] + b.w));
    }
    for (int mt = 0; mt < 4; mt++) { /* layer 3 */
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < 4; kb++) acc = mfma16(F.w3[mt][kb], h2[kb], acc);
        const float4 b = *reinterpret_cast<const float4*>(Bs + kLB3 + 16 * mt + 4 * g);
        h[mt] = pack4(relu(acc[0] + b.x), relu(acc[1] + b.y), relu(acc[2] + b.z),
                      relu(acc[3] + b.w));
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f}; /* layer 4: row 0 of 16 x 64 */
    for (int kb = 0; kb < 4; kb++) acc = mfma16(F.w4[kb], h[kb], acc);
    return acc[0] + Bs[kLB4]; /* D[0][query]: lanes 0..15, element 0 */
}

/* Logit of query 16 j + (lane % 16) of the wave's 64 (fragment group j):
 * the group's inputs are handed to the fragment lanes by shuffles, and the
 * result is returned in every lane (it is the one for this lane in the lanes
 * of group j). */
__device__ __forceinline__ float mlp_group(const MlpFrags& F, const float* __restrict__ Bs, const float u[3],
                                           const float nrm[3], const float lu[3], int lane, int j)
{
    const int r = lane & 15, g = lane >> 4;
    const int src = 16 * j + r;
    float     uj[3], nj[3], lj[3];
    for (int a = 0; a < 3; a++) {
        uj[a] = __shfl(u[a], src, 64);
        nj[a] = __shfl(nrm[a], src, 64);
        lj[a] = __shfl(lu[a], src, 64);
    }
    bf16x4 xin[3];
    query_frag(uj, nj, lj, g, xin);
    const float lg = mlp_logit(F, Bs, xin, lane);
    return __shfl(lg, r, 64); /* query 16 j + r's logit sits in lane r */
}

/* Logits of the 64 queries of a wave, one per lane: four fragment groups of 16. */
__device__ __forceinline__ float mlp_logit64(const MlpFrags& F, const float* __restrict__ Bs, const float u[3],
                                             const float nrm[3], const float lu[3], int lane)
{
    const int g   = lane >> 4;
    float     res = 0.0f;
    for (int j = 0; j < 4; j++) {
        const float got = mlp_group(F, Bs, u, nrm, lu, lane, j);
        if (g == j) res = got;
    }
    return res;
}

/* probabilities of n queries, 64 per wave. RAW: inputs from isp/nrm/light
 * (n x 3 each); else from the feature matrix X[48][B] */
template <bool RAW>
__global__ void __launch_bounds__(256) k_mlp_infer(InferArgs ia, const uint16_t* __restrict__ X, int B,
                                                   const float* __restrict__ isp, const float* __restrict__ nrm,
                                                   const float* __restrict__ light, int n, float* __restrict__ pout)
{
    __shared__ float Bs[kLBN];
    load_biases_lds(ia, Bs);
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nw = (gridDim.x * blockDim.x) >> 6;
    const int r = lane & 15, g = lane >> 4;
    MlpFrags  F;
    load_frags(ia.W, lane, F);
    for (int q0 = wave * 64; q0 < n; q0 += nw * 64) {
        const int s = q0 + lane;
        float     logit;
        if (RAW) {
            float u[3] = {0.f, 0.f, 0.f}, nn[3] = {0.f, 0.f, 0.f}, lu[3] = {0.f, 0.f, 0.f};
            if (s < n) {
                const float p3[3] = {isp[3 * s], isp[3 * s + 1], isp[3 * s + 2]};
                const float l3[3] = {light[3 * s], light[3 * s + 1], light[3 * s + 2]};
                mlp_scale(p3, ia.bx, ia.by, ia.bz, ia.bw, u);
                mlp_scale(l3, ia.bx, ia.by, ia.bz, ia.bw, lu);
                nn[0] = nrm[3 * s], nn[1] = nrm[3 * s + 1], nn[2] = nrm[3 * s + 2];
            }
            logit = mlp_logit64(F, Bs, u, nn, lu, lane);
        } else {
            logit = 0.0f;
            for (int j = 0; j < 4; j++) {
                const int sj = q0 + 16 * j + r;
                bf16x4    xin[3];
                for (int kb = 0; kb < 3; kb++)
                    for (int e = 0; e < 4; e++)
                        xin[kb][e] = sj < n ? (short) X[(size_t) (16 * kb + 4 * g + e) * B + sj] : (short) 0;
                const float lg  = mlp_logit(F, Bs, xin, lane);
                const float got = __shfl(lg, r, 64);
                if (g == j) logit = got;
            }
        }
        if (s < n) pout[s] = 1.0f / (1.0f + expf(-logit));
    }
}

/* The decision pass: for each (shaded pixel, light) query, 64 per wave,
 * the predicted bit outside the band, LS_UNSURE inside it (those are then
 * compacted into the light's exact-trace list). Pixels of light l:
 * lists[l * lstride + i], i < *nsel. */
template <bool DUAL>
__global__ void __launch_bounds__(256) k_mlp_shadow(InferArgs ia, SceneDev s, LightsDev L,
                                                    const uint32_t* __restrict__ lists, size_t lstride,
                                                    const unsigned* __restrict__ nsel, const float4* __restrict__ isp,
                                                    const int* __restrict__ ms, const int* __restrict__ md,
                                                    uint8_t* __restrict__ lstat, size_t npix, float lo_logit,
                                                    float hi_logit,
                                                    unsigned long long* __restrict__ counters)
{
    __shared__ float Bs[kLBN];
    load_biases_lds(ia, Bs);
    const uint32_t nq     = *nsel;
    const uint32_t total  = nq * (uint32_t) L.n;
    const int      lane   = threadIdx.x & 63;
    const int      wave   = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int      nw     = (gridDim.x * blockDim.x) >> 6;
    unsigned       n_pred = 0;
    if ((uint32_t) wave * 64u >= total) return; /* uniform over the wave; no barrier follows */
    MlpFrags F;
    load_frags(ia.W, lane, F);
    /* The inputs of the wave's next 64 queries are gathered while the
     * current ones are evaluated: list entry before fragment group 0, hit
     * point and model index before group 1, normal before groups 2-3, so
     * that each dependent load is covered by a group's MFMA work. */
    const uint32_t step = (uint32_t) nw * 64u;
    uint32_t       q0   = (uint32_t) wave * 64u;
    bool           valid = q0 + (uint32_t) lane < total;
    int            l     = 0;
    uint32_t       gp    = 0;
    float          u[3] = {0.f, 0.f, 0.f}, nrm[3] = {0.f, 0.f, 0.f}, lu[3] = {0.f, 0.f, 0.f};
    if (valid) {
        const uint32_t qi = q0 + (uint32_t) lane;
        l                 = (int) (qi / nq);
        gp                = lists[(size_t) l * lstride + (qi - (uint32_t) l * nq)];
        const float4 ip    = isp[gp];
        const float  p3[3] = {ip.x, ip.y, ip.z};
        mlp_scale(p3, ia.bx, ia.by, ia.bz, ia.bw, u);
        float col[4], n4[4];
        fetch_attr(s.astat, ms[gp], col, n4);
        if (DUAL && md[gp] > 0) fetch_attr(s.adyn, md[gp], col, n4);
        nrm[0] = n4[0], nrm[1] = n4[1], nrm[2] = n4[2];
        mlp_scale(L.p[l], ia.bx, ia.by, ia.bz, ia.bw, lu);
    }
    for (; q0 < total; q0 += step) {
        const int      g      = lane >> 4;
        const uint32_t qn     = q0 + step + (uint32_t) lane;
        const bool     nvalid = qn < total;
        int            nl_    = 0;
        uint32_t       ngp    = 0;
        if (nvalid) {
            nl_ = (int) (qn / nq);
            ngp = lists[(size_t) nl_ * lstride + (qn - (uint32_t) nl_ * nq)];
        }
        float res = mlp_group(F, Bs, u, nrm, lu, lane, 0);
        float4 nip = make_float4(0.f, 0.f, 0.f, 0.f);
        int    nms = 0, nmd = 0;
        if (nvalid) {
            nip = isp[ngp];
            nms = ms[ngp];
            if (DUAL) nmd = md[ngp];
        }
        const float r1 = mlp_group(F, Bs, u, nrm, lu, lane, 1);
        float       ncol[4], n4[4] = {0.f, 0.f, 0.f, 0.f};
        if (nvalid) {
            fetch_attr(s.astat, nms, ncol, n4);
            if (DUAL && nmd > 0) fetch_attr(s.adyn, nmd, ncol, n4);
        }
        const float r2 = mlp_group(F, Bs, u, nrm, lu, lane, 2);
        const float r3 = mlp_group(F, Bs, u, nrm, lu, lane, 3);
        res = g == 1 ? r1 : res;
        res = g == 2 ? r2 : res;
        res = g == 3 ? r3 : res;
        if (valid) {
            const uint8_t d = res < lo_logit ? LS_PRED_OCC : res > hi_logit ? LS_PRED_LIT : LS_UNSURE;
            lstat[(size_t) l * npix + gp] = d;
            n_pred += d != LS_UNSURE ? 1u : 0u;
        }
        valid = nvalid;
        l     = nl_;
        gp    = ngp;
        if (nvalid) {
            const float p3[3] = {nip.x, nip.y, nip.z};
            mlp_scale(p3, ia.bx, ia.by, ia.bz, ia.bw, u);
            nrm[0] = n4[0], nrm[1] = n4[1], nrm[2] = n4[2];
            mlp_scale(L.p[nl_], ia.bx, ia.by, ia.bz, ia.bw, lu);
        }
    }
    n_pred = wave_sum(n_pred);
    if (lane == 0 && n_pred) atomicAdd(&counters[CNT_PRED], (unsigned long long) n_pred);
}

/* training samples of an exact render: queries stride * k + phase of the
 * (light, shaded pixel) sequence, label = the exact visibility */
template <bool DUAL>
__global__ void k_collect(SceneDev s, LightsDev L, const uint32_t* __restrict__ lists, size_t lstride,
                          const unsigned* __restrict__ nsel, const float4* __restrict__ isp,
                          const int* __restrict__ ms, const int* __restrict__ md, const uint8_t* __restrict__ lstat,
                          size_t npix, uint32_t stride, uint32_t phase, RawSample* __restrict__ out, size_t cap,
                          unsigned long long* __restrict__ ctr)
{
    const uint64_t nq    = *nsel;
    const uint64_t qi    = ((uint64_t) blockIdx.x * blockDim.x + threadIdx.x) * stride + phase;
    const int      lane  = threadIdx.x & 63;
    const bool     valid = qi < nq * (uint64_t) L.n;
    int            l     = 0;
    uint32_t       gp    = 0;
    uint8_t        ls    = LS_SKIPPED;
    if (valid) {
        l  = (int) (qi / nq);
        gp = lists[(size_t) l * lstride + (qi - (uint64_t) l * nq)];
        ls = lstat[(size_t) l * npix + gp];
        /* a light after one that discarded the fragment does not count (its
         * ray may have been traced: the lights share one launch) */
        for (int j = 0; j < l; j++) ls = lstat[(size_t) j * npix + gp] == LS_DISCARD ? (uint8_t) LS_SKIPPED : ls;
    }
    /* one slot reservation per wave (a per-sample atomic on one counter
     * would queue every sample behind the others) */
    const bool     keep = ls == LS_LIT || ls == LS_OCCLUDED;
    const uint64_t m    = __ballot(keep);
    if (m == 0ull) return; /* uniform over the wave */
    const int          leader = __ffsll((unsigned long long) m) - 1;
    unsigned long long base   = 0;
    if (lane == leader) base = atomicAdd(&ctr[0], (unsigned long long) __popcll(m));
    base                          = __shfl(base, leader, 64);
    const unsigned long long slot = base + (unsigned long long) __popcll(m & ((1ull << lane) - 1ull));
    const uint64_t           lit  = __ballot(keep && ls == LS_LIT && slot < cap);
    if (lane == leader && lit) atomicAdd(&ctr[1], (unsigned long long) __popcll(lit));
    if (!keep || slot >= cap) return;
    RawSample    rs;
    const float4 ip = isp[gp];
    rs.isp[0] = ip.x, rs.isp[1] = ip.y, rs.isp[2] = ip.z;
    float col[4], n4[4];
    fetch_attr(s.astat, ms[gp], col, n4);
    if (DUAL && md[gp] > 0) fetch_attr(s.adyn, md[gp], col, n4);
    rs.nrm[0] = n4[0], rs.nrm[1] = n4[1], rs.nrm[2] = n4[2];
    rs.light[0] = L.p[l][0], rs.light[1] = L.p[l][1], rs.light[2] = L.p[l][2];
    rs.label    = ls == LS_LIT ? 1.0f : 0.0f;
    out[slot]   = rs;
}

__global__ void k_unpack_samples(const RawSample* __restrict__ S, size_t start, size_t n, float* __restrict__ out)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const RawSample r = S[start + i];
    float*          o = out + 10 * i;
    for (int k = 0; k < 3; k++) o[k] = r.isp[k], o[3 + k] = r.nrm[k], o[6 + k] = r.light[k];
    o[9] = r.label;
}

} // namespace

/* ------------------------------------------------------------------------ */
/* host side                                                                */
/* ------------------------------------------------------------------------ */

void qbt_ctx::mlp_free()
{
    if (!mlp) return;
    MlpState* m = mlp;
    for (DevBuf* b : {&m->params, &m->m, &m->v, &m->grads, &m->packed, &m->ctr, &m->samples, &m->perm, &m->keys,
                      &m->vals, &m->sort_tmp, &m->X, &m->gpart, &m->lossbuf, &m->io_x, &m->io_p, &m->io_y,
                      &m->io_a, &m->io_c, &m->flat})
        dev_free(*b);
    delete mlp;
    mlp = nullptr;
}

bool qbt_surrogate_ready(const qbt_ctx* c) { return c && c->mlp && c->mlp->ready; }

static int mlp_state(qbt_ctx* c, MlpState** out)
{
    if (!c->mlp) {
        c->mlp = new (std::nothrow) MlpState();
        if (!c->mlp) return QBT_ENOMEM;
        MlpState* m = c->mlp;
        int       rc;
        if ((rc = c->dev_alloc(m->params, kMlpParams * 4)) || (rc = c->dev_alloc(m->m, kMlpParams * 4)) ||
            (rc = c->dev_alloc(m->v, kMlpParams * 4)) || (rc = c->dev_alloc(m->grads, kGradN * 4)) ||
            (rc = c->dev_alloc(m->packed, kPackN * 2)) || (rc = c->dev_alloc(m->lossbuf, 64)) ||
            (rc = c->dev_alloc(m->ctr, 16)) || (rc = c->dev_alloc(m->flat, kMlpParams * 4))) {
            c->mlp_free();
            return rc;
        }
        k_fill_u64<<<1, 64, 0, c->stream>>>((unsigned long long*) m->ctr.p, 2);
        HIPCHK(hipGetLastError());
    }
    *out = c->mlp;
    return QBT_OK;
}

static int mlp_repack(qbt_ctx* c)
{
    k_mlp_pack<<<nblk(kPackN), 256, 0, c->stream>>>((const float*) c->mlp->params.p, (uint16_t*) c->mlp->packed.p);
    HIPCHK(hipGetLastError());
    return QBT_OK;
}

static int mlp_reset_adam(qbt_ctx* c)
{
    MlpState* m = c->mlp;
    k_fill_f32<<<nblk(kMlpParams), 256, 0, c->stream>>>((float*) m->m.p, kMlpParams, 0.0f);
    k_fill_f32<<<nblk(kMlpParams), 256, 0, c->stream>>>((float*) m->v.p, kMlpParams, 0.0f);
    HIPCHK(hipGetLastError());
    m->t = 0;
    return QBT_OK;
}

extern "C" int qbt_surrogate_init(qbt_ctx* c, uint64_t seed)
{
    if (!c) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    MlpState* m;
    int       rc = mlp_state(c, &m);
    if (rc) return rc;
    k_mlp_init<<<nblk(kMlpParams), 256, 0, c->stream>>>((float*) m->params.p, seed);
    HIPCHK(hipGetLastError());
    if ((rc = mlp_reset_adam(c)) || (rc = mlp_repack(c))) return rc;
    m->ready = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

extern "C" int qbt_surrogate_set_params(qbt_ctx* c, const float* params, size_t n)
{
    if (!c || !params || n != (size_t) kMlpParams) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    MlpState* m;
    int       rc = mlp_state(c, &m);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(m->params.p, params, kMlpParams * 4, hipMemcpyHostToDevice, c->stream));
    if ((rc = mlp_reset_adam(c)) || (rc = mlp_repack(c))) return rc;
    m->ready = true;
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

extern "C" int qbt_surrogate_get_params(qbt_ctx* c, float* params, size_t n)
{
    if (!c || !params || n != (size_t) kMlpParams || !qbt_surrogate_ready(c)) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(params, c->mlp->params.p, kMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

extern "C" int qbt_surrogate_set_band(qbt_ctx* c, float lo, float hi)
{
    if (!c || !(lo > 0.0f) || !(hi < 1.0f) || !(lo <= hi)) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    MlpState* m;
    int       rc = mlp_state(c, &m);
    if (rc) return rc;
    m->band_lo = lo;
    m->band_hi = hi;
    return QBT_OK;
}

extern "C" int qbt_surrogate_reserve(qbt_ctx* c, size_t nsamples, uint32_t stride)
{
    if (!c || nsamples == 0 || nsamples > ((size_t) 1 << 31) || stride == 0) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    MlpState* m;
    int       rc = mlp_state(c, &m);
    if (rc) return rc;
    if ((rc = c->dev_alloc(m->samples, nsamples * sizeof(RawSample)))) return rc;
    m->cap    = nsamples;
    m->stride = stride;
    m->phase  = 0;
    m->perm_n = 0;
    k_fill_u64<<<1, 64, 0, c->stream>>>((unsigned long long*) m->ctr.p, 2);
    HIPCHK(hipGetLastError());
    return QBT_OK;
}

static int mlp_counts(qbt_ctx* c, size_t* n, size_t* lit)
{
    unsigned long long h[2] = {0ull, 0ull};
    MlpState*          m    = c->mlp;
    if (m && m->cap) {
        HIPCHK(hipMemcpyAsync(h, m->ctr.p, 16, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (n) *n = m ? (size_t) (h[0] < m->cap ? h[0] : m->cap) : 0;
    if (lit) *lit = (size_t) h[1];
    return QBT_OK;
}

extern "C" int qbt_surrogate_samples(qbt_ctx* c, size_t* n, size_t* lit)
{
    if (!c || !n) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    return mlp_counts(c, n, lit);
}

extern "C" int qbt_surrogate_read_samples(qbt_ctx* c, size_t start, size_t n, float* out)
{
    if (!c || (n && !out)) return QBT_EINVAL;
    if (n == 0) return QBT_OK;
    HIPCHK(hipSetDevice(c->device));
    size_t held = 0;
    int    rc   = mlp_counts(c, &held, nullptr);
    if (rc) return rc;
    if (start > held || n > held - start) return QBT_ERANGE;
    MlpState* m = c->mlp;
    if ((rc = c->dev_alloc(m->io_c, n * 40))) return rc;
    k_unpack_samples<<<nblk(n), 256, 0, c->stream>>>((const RawSample*) m->samples.p, start, n, (float*) m->io_c.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, m->io_c.p, n * 40, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

constexpr int kTrainBlocks = 256; /* blocks of k_mlp_train, rows of partial gradients, at most */

/* batch buffers for batch width B (multiple of 16) */
static int mlp_batch_bufs(qbt_ctx* c, int B)
{
    MlpState*    m      = c->mlp;
    const size_t b      = (size_t) B;
    const size_t chunks = (b + kChunk - 1) / kChunk;
    const size_t rows   = chunks < (size_t) kTrainBlocks ? chunks : (size_t) kTrainBlocks;
    int          rc;
    if ((rc = c->dev_alloc(m->X, kMlpInPad * b * 2)) || (rc = c->dev_alloc(m->gpart, rows * kGradN * 4)) ||
        (rc = c->dev_alloc(m->io_y, b * 4)) || (rc = c->dev_alloc(m->io_p, b * 4)))
        return rc;
    return QBT_OK;
}

/* Forward and backward of the B-column batch in X (n valid columns, labels
 * in io_y): the summed loss into lossbuf, the gradient (layout kGradN) into
 * grads. Three launches: loss reset, k_mlp_train, k_grad_rows. */
static int mlp_fwd_bwd(qbt_ctx* c, int B, int n, double* loss_out)
{
    MlpState*   m      = c->mlp;
    hipStream_t st     = c->stream;
    const int   chunks = (B + kChunk - 1) / kChunk;
    const int   rows   = chunks < kTrainBlocks ? chunks : kTrainBlocks;
    double*     lossd  = (double*) m->lossbuf.p;
    k_fill_f32<<<1, 64, 0, st>>>((float*) m->lossbuf.p, 2, 0.0f);
    k_mlp_train<<<rows, 256, 0, st>>>((const uint16_t*) m->packed.p, (const float*) m->params.p,
                                      (const uint16_t*) m->X.p, (const float*) m->io_y.p, B, n, (float*) m->gpart.p,
                                      lossd);
    k_grad_rows<<<nblk(kGradN), 256, 0, st>>>((const float*) m->gpart.p, rows, (float*) m->grads.p);
    HIPCHK(hipGetLastError());
    if (loss_out) {
        double l = 0.0;
        HIPCHK(hipMemcpyAsync(&l, lossd, 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        *loss_out = l / (n > 0 ? n : 1);
    }
    return QBT_OK;
}

static int mlp_adam_step(qbt_ctx* c, float lr)
{
    MlpState* m = c->mlp;
    m->t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float) m->t), c2 = 1.0f - powf(b2, (float) m->t);
    k_mlp_adam<<<nblk(kMlpParams), 256, 0, c->stream>>>((float*) m->params.p, (const float*) m->grads.p,
                                                         (float*) m->m.p, (float*) m->v.p, lr, b1, b2, eps, c1, c2);
    HIPCHK(hipGetLastError());
    return mlp_repack(c);
}

/* host feature rows (n x 42) -> X, labels -> io_y; B = n rounded up to 16 */
static int mlp_load_host(qbt_ctx* c, const float* x, const float* y, size_t n, int* Bout)
{
    MlpState* m = c->mlp;
    const int B = (int) ((n + 15) / 16) * 16;
    int       rc;
    if ((rc = mlp_batch_bufs(c, B)) || (rc = c->dev_alloc(m->io_x, n * kMlpIn * 4)) ||
        (rc = c->dev_alloc(m->io_a, n * 4)))
        return rc;
    HIPCHK(hipMemcpyAsync(m->io_x.p, x, n * kMlpIn * 4, hipMemcpyHostToDevice, c->stream));
    if (y) HIPCHK(hipMemcpyAsync(m->io_a.p, y, n * 4, hipMemcpyHostToDevice, c->stream));
    k_copy_features<<<nblk(B), 256, 0, c->stream>>>((const float*) m->io_x.p, B, (int) n, (uint16_t*) m->X.p,
                                                    y ? (const float*) m->io_a.p : nullptr, (float*) m->io_y.p);
    HIPCHK(hipGetLastError());
    *Bout = B;
    return QBT_OK;
}

static unsigned infer_grid(const qbt_ctx* c, size_t n)
{
    const size_t groups = (n + 15) / 16; /* one group of 16 queries per wave, 4 waves per block */
    const size_t blocks = (groups + 3) / 4;
    const size_t maxb   = (size_t) c->num_cus * 4;
    return (unsigned) (blocks < maxb ? (blocks ? blocks : 1) : maxb);
}

extern "C" int qbt_surrogate_forward(qbt_ctx* c, const float* x, size_t n, float* p)
{
    if (!c || (n && (!x || !p)) || n > ((size_t) 1 << 24) || !qbt_surrogate_ready(c)) return QBT_EINVAL;
    if (n == 0) return QBT_OK;
    HIPCHK(hipSetDevice(c->device));
    int B, rc;
    if ((rc = mlp_load_host(c, x, nullptr, n, &B))) return rc;
    MlpState* m = c->mlp;
    InferArgs ia{(const uint16_t*) m->packed.p, (const float*) m->params.p, 0.f, 0.f, 0.f, 1.f};
    k_mlp_infer<false><<<infer_grid(c, n), 256, 0, c->stream>>>(ia, (const uint16_t*) m->X.p, B, nullptr, nullptr,
                                                                 nullptr, (int) n, (float*) m->io_p.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(p, m->io_p.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

extern "C" int qbt_surrogate_predict(qbt_ctx* c, const float* isp, const float* nrm, const float* light, size_t n,
                                     const float basecube[4], float* p)
{
    if (!c || (n && (!isp || !nrm || !light || !p)) || !basecube || !(basecube[3] > 0.0f) ||
        n > ((size_t) 1 << 24) || !qbt_surrogate_ready(c))
        return QBT_EINVAL;
    if (n == 0) return QBT_OK;
    HIPCHK(hipSetDevice(c->device));
    MlpState* m = c->mlp;
    int       rc;
    if ((rc = c->dev_alloc(m->io_x, n * 12 * 3)) || (rc = c->dev_alloc(m->io_a, n * 4))) return rc;
    float* d = (float*) m->io_x.p;
    HIPCHK(hipMemcpyAsync(d, isp, n * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d + 3 * n, nrm, n * 12, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(d + 6 * n, light, n * 12, hipMemcpyHostToDevice, c->stream));
    InferArgs ia{(const uint16_t*) m->packed.p, (const float*) m->params.p, basecube[0], basecube[1], basecube[2],
                 basecube[3]};
    k_mlp_infer<true><<<infer_grid(c, n), 256, 0, c->stream>>>(ia, nullptr, 0, d, d + 3 * n, d + 6 * n, (int) n,
                                                                (float*) m->io_a.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(p, m->io_a.p, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

extern "C" int qbt_surrogate_grad(qbt_ctx* c, const float* x, const float* y, size_t n, float* grads, float* loss)
{
    if (!c || !x || !y || n == 0 || n > ((size_t) 1 << 22) || !qbt_surrogate_ready(c)) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    int B, rc;
    if ((rc = mlp_load_host(c, x, y, n, &B))) return rc;
    double l = 0.0;
    if ((rc = mlp_fwd_bwd(c, B, (int) n, &l))) return rc;
    if (grads) {
        k_grads_to_flat<<<nblk(kMlpParams), 256, 0, c->stream>>>((const float*) c->mlp->grads.p,
                                                                  (float*) c->mlp->flat.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(grads, c->mlp->flat.p, kMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    if (loss) *loss = (float) l;
    return QBT_OK;
}

extern "C" int qbt_surrogate_step(qbt_ctx* c, const float* x, const float* y, size_t n, float lr, float* loss)
{
    if (!(lr >= 0.0f)) return QBT_EINVAL;
    int rc = qbt_surrogate_grad(c, x, y, n, nullptr, loss);
    if (rc) return rc;
    if ((rc = mlp_adam_step(c, lr))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return QBT_OK;
}

int qbt_collect_batch(qbt_ctx* c, const SceneDev& s, bool dual, const LightsDev& L, const uint32_t* lists,
                      size_t list_stride, const unsigned* nsel, size_t npix)
{
    MlpState* m = c->mlp;
    if (!m || !m->cap) return QBT_OK;
    const uint32_t stride = m->stride, phase = m->phase;
    m->phase = (m->phase + 1) % stride; /* successive renders sample different queries */
    const size_t threads = (npix * (size_t) L.n) / stride + 1;
    auto         go      = [&](auto kern) {
        kern<<<nblk(threads), 256, 0, c->stream>>>(s, L, lists, list_stride, nsel, (const float4*) c->wb_isp.p,
                                                   (const int*) c->wb_ms.p, (const int*) c->wb_md.p,
                                                   (const uint8_t*) c->wb_lstat.p, npix, stride, phase,
                                                   (RawSample*) m->samples.p, m->cap, (unsigned long long*) m->ctr.p);
    };
    if (dual) go(k_collect<true>);
    else go(k_collect<false>);
    HIPCHK(hipGetLastError());
    return QBT_OK;
}

extern "C" int qbt_surrogate_train(qbt_ctx* c, const float basecube[4], int steps, int batch, float lr,
                                   uint64_t seed, float* loss_out)
{
    if (!c || !basecube || steps <= 0 || batch < 16 || batch % 16 || batch > (1 << 20) || !(lr >= 0.0f) ||
        !(basecube[3] > 0.0f) || !qbt_surrogate_ready(c))
        return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    MlpState* m = c->mlp;
    size_t    n = 0;
    int       rc;
    if ((rc = mlp_counts(c, &n, nullptr))) return rc;
    if (n < (size_t) batch) return QBT_EINVAL;
    if ((rc = mlp_batch_bufs(c, batch)) || (rc = c->dev_alloc(m->perm, n * 4)) || (rc = c->dev_alloc(m->keys, n * 8)) ||
        (rc = c->dev_alloc(m->vals, n * 4)))
        return rc;
    size_t tb = 0;
    HIPCHK(rocprim::radix_sort_pairs(nullptr, tb, (uint32_t*) m->keys.p, (uint32_t*) m->keys.p + n,
                                     (uint32_t*) m->vals.p, (uint32_t*) m->perm.p, n, 0, 32, c->stream));
    if ((rc = c->dev_alloc(m->sort_tmp, tb ? tb : 16))) return rc;
    auto shuffle = [&]() -> int {
        k_perm_keys<<<nblk(n), 256, 0, c->stream>>>((uint32_t*) m->keys.p, (uint32_t*) m->vals.p, n,
                                                    seed + 0x51ED270Bull * ++m->epoch);
        size_t b2 = tb;
        HIPCHK(rocprim::radix_sort_pairs(m->sort_tmp.p, b2, (uint32_t*) m->keys.p, (uint32_t*) m->keys.p + n,
                                         (uint32_t*) m->vals.p, (uint32_t*) m->perm.p, n, 0, 32, c->stream));
        m->pos    = 0;
        m->perm_n = n;
        return QBT_OK;
    };
    /* a new shuffle when the store changed size or the permutation is used up */
    if (m->perm_n != n && (rc = shuffle())) return rc;
    double last = 0.0;
    for (int it = 0; it < steps; it++) {
        if (m->pos + (size_t) batch > n && (rc = shuffle())) return rc;
        k_batch_features<<<nblk(batch), 256, 0, c->stream>>>((const RawSample*) m->samples.p,
                                                             (const uint32_t*) m->perm.p, m->pos, batch, basecube[0],
                                                             basecube[1], basecube[2], basecube[3],
                                                             (uint16_t*) m->X.p, (float*) m->io_y.p);
        HIPCHK(hipGetLastError());
        m->pos += (size_t) batch;
        if ((rc = mlp_fwd_bwd(c, batch, batch, it == steps - 1 ? &last : nullptr))) return rc;
        if ((rc = mlp_adam_step(c, lr))) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (loss_out) *loss_out = (float) last;
    return QBT_OK;
}

void qbt_surrogate_band(const qbt_ctx* c, float* lo, float* hi)
{
    *lo = c->mlp ? c->mlp->band_lo : 0.0f;
    *hi = c->mlp ? c->mlp->band_hi : 0.0f;
}

int qbt_surrogate_decide(qbt_ctx* c, const SceneDev& s, bool dual, const LightsDev& L, const uint32_t* lists,
                         size_t list_stride, const unsigned* nsel, size_t npix)
{
    MlpState* m = c->mlp;
    if (!m || !m->ready) return QBT_EINVAL;
    InferArgs ia{(const uint16_t*) m->packed.p, (const float*) m->params.p, s.bx, s.by, s.bz, s.bw};
    /* band edges as logits: p < lo  <=>  logit < log(lo / (1 - lo)) */
    const float lo   = logf(m->band_lo / (1.0f - m->band_lo));
    const float hi   = logf(m->band_hi / (1.0f - m->band_hi));
    const int   grid = c->num_cus * 2; /* two 4-wave blocks per CU: the occupancy at ~220 VGPRs */
    auto        go   = [&](auto kern) {
        kern<<<grid, 256, 0, c->stream>>>(ia, s, L, lists, list_stride, nsel, (const float4*) c->wb_isp.p,
                                          (const int*) c->wb_ms.p, (const int*) c->wb_md.p, (uint8_t*) c->wb_lstat.p,
                                          npix, lo, hi, (unsigned long long*) c->counters.p);
    };
    c->prof_mark_begin(QBT_PROF_SURROGATE);
    if (dual) go(k_mlp_shadow<true>);
    else go(k_mlp_shadow<false>);
    c->prof_mark_end();
    return QBT_OK;
}
