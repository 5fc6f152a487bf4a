// Learned return map: the 10-64-64-3 GELU network on the gfx950 matrix cores.
//
// Parameters are one FP32 vector (master copy, updated by Adam):
//   W1[64][10], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]   (out x in)
// The kernels round weights and back-propagated gradients to bfloat16 (round
// to nearest even) where they enter v_mfma_f32_16x16x16_bf16, and accumulate
// in FP32.  Inputs and hidden activations enter as two bfloat16 halves,
// hi = bf16(x) and lo = bf16(x - hi), in two MFMA passes per layer: 16
// significant bits of every value, so that the outputs are not limited by
// the 8 bits of one bfloat16 (an output made of 64 terms would carry their
// rounding errors).
//
// Every product is computed "transposed", features x samples: a layer's
// output Z^T[o][b] = sum_i W[o][i] X^T[i][b] takes W as the A operand (16 rows
// of outputs x 16 inputs) and the activations of 16 samples as the B operand.
// The D fragment of one 16-output block (lane l holds rows 4(l/16)..+3 of
// column l%16) has exactly the B-operand layout of the next layer's 16-input
// block, so hidden activations go from one MFMA to the next in registers.
#pragma once
#include "mpmx_device.hpp"

namespace mpmx {

constexpr int kNI = 10, kNH = 64, kNO = 3;
constexpr int kOW1 = 0, kOB1 = kOW1 + kNH * kNI, kOW2 = kOB1 + kNH, kOB2 = kOW2 + kNH * kNH, kOW3 = kOB2 + kNH,
              kOB3 = kOW3 + kNO * kNH;
constexpr int kNParam = kOB3 + kNO;  // 5059

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(float f) { return __uint_as_float(((uint32_t)bf16_bits(f)) << 16); }

// the low half x - bf16(x) of a split input, as bfloat16 bits
__device__ __forceinline__ unsigned short bf16_lo_bits(float f) { return bf16_bits(f - bf16_value(f)); }

__device__ __forceinline__ bf16x4 pack_bf16(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)bf16_bits(a);
    r[1] = (short)bf16_bits(b);
    r[2] = (short)bf16_bits(c);
    r[3] = (short)bf16_bits(d);
    return r;
}

__device__ __forceinline__ bf16x4 pack_bf16_lo(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)bf16_lo_bits(a);
    r[1] = (short)bf16_lo_bits(b);
    r[2] = (short)bf16_lo_bits(c);
    r[3] = (short)bf16_lo_bits(d);
    return r;
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// tanh form of GELU (Hendrycks & Gimpel 2016), 0.5 x (1 + tanh u) with
// u = sqrt(2/pi) (x + 0.044715 x^3), and its derivative, written with the
// logistic function: 0.5 (1 + tanh u) = sig(2 u) = 1 / (1 + exp(-2 u)) and
// 1 - tanh^2 u = 4 sig (1 - sig).  One hardware exponential and one
// reciprocal (about 1 ulp each); exp(-2u) overflowing to inf gives sig = 0.
__device__ __forceinline__ float sig2u(float x) {
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    return __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}
__device__ __forceinline__ float gelu_f(float x) { return x * sig2u(x); }
__device__ __forceinline__ float gelu_grad_f(float x) {
    const float s = sig2u(x);
    const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
    return s + 2.0f * x * (s * (1.0f - s)) * du;
}

// Weights of the forward pass in LDS as bfloat16, zero-padded to 16-wide blocks.
// Rows are padded (16 -> 20 and 64 -> 68 entries) so that the 16 rows one A
// fragment reads fall in distinct LDS banks: a row of 16 or 64 bfloat16 is 8
// or 32 banks long, which puts rows r and r + 4 (or every row) in the same
// bank; rows of 40 or 136 bytes spread them over all 32 banks, and keep the
// 8-byte alignment of the fragments.
constexpr int kS16 = 20, kS64 = 68;
struct MlpWeights {
    unsigned short W1[kNH][kS16];  // [o][i], i >= 10 zero
    unsigned short W2[kNH][kS64];
    unsigned short W3[16][kS64];  // [o][i], o >= 3 zero
    float b1[kNH], b2[kNH], b3[16];
};

// block-cooperative fill of w from the FP32 parameter vector
__device__ __forceinline__ void load_weights(const float* __restrict__ prm, MlpWeights& w) {
    for (int k = threadIdx.x; k < kNH * 16; k += blockDim.x) {
        const int o = k / 16, i = k % 16;
        w.W1[o][i] = i < kNI ? bf16_bits(prm[kOW1 + o * kNI + i]) : 0;
    }
    for (int k = threadIdx.x; k < kNH * kNH; k += blockDim.x) w.W2[k / kNH][k % kNH] = bf16_bits(prm[kOW2 + k]);
    for (int k = threadIdx.x; k < 16 * kNH; k += blockDim.x) {
        const int o = k / kNH, i = k % kNH;
        w.W3[o][i] = o < kNO ? bf16_bits(prm[kOW3 + o * kNH + i]) : 0;
    }
    for (int k = threadIdx.x; k < kNH; k += blockDim.x) {
        w.b1[k] = prm[kOB1 + k];
        w.b2[k] = prm[kOB2 + k];
    }
    for (int k = threadIdx.x; k < 16; k += blockDim.x) w.b3[k] = k < kNO ? prm[kOB3 + k] : 0.0f;
}

// A fragment of a bfloat16 matrix with rows of COLS entries (the padded
// stride) for output block mb, input block kb
template <int COLS>
__device__ __forceinline__ bf16x4 a_frag(const unsigned short (*M)[COLS], int mb, int kb, int lane) {
    return *reinterpret_cast<const bf16x4*>(&M[mb * 16 + (lane & 15)][kb * 16 + 4 * (lane >> 4)]);
}

// Forward pass of one tile of 16 samples.  xb, xl are the B fragments of the
// inputs' high and low bfloat16 halves (lane l: features 4(l/16)..+3 of sample
// l%16).  Outputs: z1, z2 (pre-activations, D layout per 16-row block), h1b,
// h1l, h2b, h2l (high and low bfloat16 halves of the activations, B layout)
// and y (lanes 0..15 hold y[0..2] of sample l%16 in y[0..2]).
struct MlpTile {
    f32x4 z1[4], z2[4];
    bf16x4 h1b[4], h1l[4], h2b[4], h2l[4];
    f32x4 y;
};

__device__ __forceinline__ void mlp_tile_forward(const MlpWeights& w, bf16x4 xb, bf16x4 xl, int lane, MlpTile& t) {
    const int r0 = 4 * (lane >> 4);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        f32x4 c = {w.b1[mb * 16 + r0], w.b1[mb * 16 + r0 + 1], w.b1[mb * 16 + r0 + 2], w.b1[mb * 16 + r0 + 3]};
        const bf16x4 a = a_frag<kS16>(w.W1, mb, 0, lane);
        c = mfma16(a, xb, c);
        c = mfma16(a, xl, c);
        t.z1[mb] = c;
        const float g0 = gelu_f(c[0]), g1 = gelu_f(c[1]), g2 = gelu_f(c[2]), g3 = gelu_f(c[3]);
        t.h1b[mb] = pack_bf16(g0, g1, g2, g3);
        t.h1l[mb] = pack_bf16_lo(g0, g1, g2, g3);
    }
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
        f32x4 c = {w.b2[mb * 16 + r0], w.b2[mb * 16 + r0 + 1], w.b2[mb * 16 + r0 + 2], w.b2[mb * 16 + r0 + 3]};
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
            const bf16x4 a = a_frag<kS64>(w.W2, mb, kb, lane);
            c = mfma16(a, t.h1b[kb], c);
            c = mfma16(a, t.h1l[kb], c);
        }
        t.z2[mb] = c;
        const float g0 = gelu_f(c[0]), g1 = gelu_f(c[1]), g2 = gelu_f(c[2]), g3 = gelu_f(c[3]);
        t.h2b[mb] = pack_bf16(g0, g1, g2, g3);
        t.h2l[mb] = pack_bf16_lo(g0, g1, g2, g3);
    }
    f32x4 c = {w.b3[r0], w.b3[r0 + 1], w.b3[r0 + 2], w.b3[r0 + 3]};
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
        const bf16x4 a = a_frag<kS64>(w.W3, 0, kb, lane);
        c = mfma16(a, t.h2b[kb], c);
        c = mfma16(a, t.h2l[kb], c);
    }
    t.y = c;
}

}  // namespace mpmx
