/*
 * mlp.hpp - the shadow-visibility MLP on gfx950 matrix cores.
 *
 * Network: 42 inputs (6-band sin/cos of the hit position scaled to the cube,
 * the unit normal, the light position scaled to the cube) -> 64 -> 64 -> 64
 * -> 1, ReLU, sigmoid. Weights and activations are bf16, accumulation fp32,
 * biases fp32 (no reference counterpart: the reference traces every shadow
 * ray).
 *
 * Matrix-core mapping: v_mfma_f32_16x16x16_bf16 computes D(16x16) +=
 * A(16x16) B(16x16) with, for lane l of a wave and e = 0..3,
 *     A[l % 16][4 (l / 16) + e],  B[4 (l / 16) + e][l % 16],
 *     D[4 (l / 16) + e][l % 16].
 * Layers are evaluated transposed, Z (features x samples) = W X^T, with 16
 * samples per lane column: the D fragment of feature tile t holds features
 * 16t + 4 (l/16) + e of sample l % 16, which is exactly the B fragment of
 * K-block t of the next layer, so bias + ReLU + bf16 packing happen in
 * registers between layers.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qbt {

constexpr int kMlpIn     = 42;
constexpr int kMlpInPad  = 48;
constexpr int kMlpH      = 64;
constexpr int kMlpParams = 64 * 42 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1;
/* offsets in the flat fp32 parameter array (row-major W[out][in]) */
constexpr int kW1 = 0, kB1 = 64 * 42, kW2 = kB1 + 64, kB2 = kW2 + 64 * 64, kW3 = kB2 + 64, kB3 = kW3 + 64 * 64,
              kW4 = kB3 + 64, kB4 = kW4 + 64;
/* packed bf16 weights: W1 64x48 (zero-padded inputs), W2 64x64, W3 64x64,
 * W4 16x64 (row 0 = the output layer, rows 1..15 zero) */
constexpr int kP1 = 0, kP2 = kP1 + 64 * 48, kP3 = kP2 + 64 * 64, kP4 = kP3 + 64 * 64, kPackN = kP4 + 16 * 64;

typedef short  bf16x4 __attribute__((ext_vector_type(4)));
typedef float  f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bfv4 __attribute__((ext_vector_type(4)));

/* fp32 -> bf16, round to nearest even (NaN quieted): gfx950's
 * v_cvt_pk_bf16_f32 */
__device__ __forceinline__ uint16_t f2bf(float x) { return __builtin_bit_cast(uint16_t, (__bf16) x); }
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t) b << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d)
{
    const f32x4 v = {a, b, c, d};
    return __builtin_bit_cast(bf16x4, __builtin_convertvector(v, bfv4)); /* two v_cvt_pk_bf16_f32 */
}

/* Feature f (0..47) of one query: f = 6k + 2a + {0 sin, 1 cos} for band k
 * (frequency 2^k pi) and axis a of the hit position scaled to [0,1]; 36..38
 * normal; 39..41 light scaled to [0,1]; 42..47 zero padding. Training and
 * inference compute them with the same code, so the network sees the same
 * features in both. */
/* element a (0..2) of a 3-vector, selected by value (an indexed access with
 * a lane-dependent index would put the vector in scratch memory) */
__device__ __forceinline__ float pick3(const float v[3], int a)
{
    const float x = v[0], y = v[1], z = v[2];
    return a == 0 ? x : (a == 1 ? y : z);
}

/* sin and cos of 2^k pi x (band k of coordinate x), by the hardware's
 * transcendental unit: v_sin_f32 / v_cos_f32 take revolutions, and
 * 2^k pi x = 2 pi (2^(k-1) x) with 2^(k-1) x exact */
__device__ __forceinline__ void band_sincos(int k, float x, float* s, float* c)
{
    const float rev = ldexpf(x, k - 1);
    *s              = __builtin_amdgcn_sinf(rev);
    *c              = __builtin_amdgcn_cosf(rev);
}

__device__ __forceinline__ float mlp_feature(int f, const float u[3], const float n[3], const float lu[3])
{
    if (f < 36) {
        int   k = f / 6, a = (f % 6) >> 1;
        float s, c;
        band_sincos(k, pick3(u, a), &s, &c);
        return (f & 1) ? c : s;
    }
    if (f < 39) return pick3(n, f - 36);
    if (f < 42) return pick3(lu, f - 39);
    return 0.0f;
}

/* scaled hit / light position: (p - lo) / edge with lo = (x, y - w, z - w) */
__device__ __forceinline__ void mlp_scale(const float p[3], float bx, float by, float bz, float bw, float u[3])
{
    u[0] = (p[0] - bx) / bw;
    u[1] = (p[1] - (by - bw)) / bw;
    u[2] = (p[2] - (bz - bw)) / bw;
}

} // namespace qbt
