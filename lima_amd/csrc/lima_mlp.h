// lima_mlp.h -- the per-atom step-scale MLP (32 inputs -> 64 -> 64 SiLU -> 1, output 4 sigmoid(z) in
// (0, 4)) as BF16 MFMA tiles with FP32 accumulation, shared by the batched minimiser (inference) and
// the training kernels.
//
// One wave evaluates 16 samples with v_mfma_f32_16x16x16_bf16 (operand lanes: A row = lane % 16,
// k = 4 (lane / 16) + i; B column = lane % 16, same k; D rows 4 (lane / 16) + i, column lane % 16).
// Weights are the A operand and activations the B operand, so each layer's D tile (hidden rows x
// 16 samples) is, after SiLU and BF16 rounding, exactly the next layer's B operand.
#pragma once

#include "../../include/lima_hip.h"
#include "lima_dev.h"

namespace lima {

constexpr int MLP_IN = 32;
constexpr int MLP_H = 64;
// parameter vector layout (FP32 master copy): W1 [32][64], b1 [64], W2 [64][64], b2 [64], w3 [64], b3
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {   // round to nearest even
    const unsigned u = __float_as_uint(f);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = 1.f / (1.f + __expf(-z));
    return s * (1.f + z * (1.f - s));
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a); r[1] = (short)f2bf(b); r[2] = (short)f2bf(c); r[3] = (short)f2bf(d);
    return r;
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Weights in the layouts the MFMA operands read (BF16) plus the FP32 vectors.
struct MlpWeights {
    const unsigned short* w1t;   // [64][32]: W1 transposed (hidden-major)
    const unsigned short* w2t;   // [64][64]: W2 transposed (output-major)
    const unsigned short* w2;    // [64][64]: W2 (input-major), for the backward pass
    const float* p;              // FP32 parameter vector (biases, w3, b3 read from here)
};

// Forward pass of 16 samples held by one wave.  xb[ks]: B operand of input block ks (this lane's
// sample lane % 16, inputs 16 ks + 4 (lane / 16) + 0..3).  Returns z3 of sample lane % 16 (in every
// lane of the sample's column) and keeps the pre-activations for a backward pass.
__device__ __forceinline__ float mlp_forward16(const MlpWeights& w, const bf16x4* xb, f32x4* z1, bf16x4* h1, f32x4* z2) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int hb = 0; hb < 4; hb++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(w.w1t + (hb * 16 + r) * MLP_IN + ks * 16 + 4 * g);
            acc = mfma16(a, xb[ks], acc);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) acc[i] += w.p[MLP_OFF_B1 + hb * 16 + 4 * g + i];
        z1[hb] = acc;
        h1[hb] = pack4(silu(acc[0]), silu(acc[1]), silu(acc[2]), silu(acc[3]));
    }
    float part = 0.f;
#pragma unroll
    for (int hb = 0; hb < 4; hb++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(w.w2t + (hb * 16 + r) * MLP_H + kb * 16 + 4 * g);
            acc = mfma16(a, h1[kb], acc);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            acc[i] += w.p[MLP_OFF_B2 + hb * 16 + 4 * g + i];
            part += w.p[MLP_OFF_W3 + hb * 16 + 4 * g + i] * silu(acc[i]);
        }
        z2[hb] = acc;
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    return part + w.p[MLP_OFF_B3];
}

__device__ __forceinline__ float mlp_scale(float z3) { return 4.f / (1.f + __expf(-z3)); }

// Device views of a lima_mlp handle (mlp.hip).
MlpWeights mlp_view(const lima_mlp* h);

}  // namespace lima
