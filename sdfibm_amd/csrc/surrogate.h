// surrogate.h -- learned volume-fraction surrogate: an MLP 11 -> 32 -> 32 -> 1
// (ReLU) over a cut cell's 8 corner signed distances / h and the unit SDF
// gradient at the cell centre, BF16 weights and activations with FP32
// accumulation, evaluated with MFMA (v_mfma_f32_16x16x16_bf16) by a whole wave.
//
// Parameter vector (FP32 master copy, 1473 values):
//   W1[32][11], b1[32], W2[32][32], b2[32], w3[32], b3
//
// Matrix layout.  A wave evaluates 16 cells per MFMA group in the transposed
// form Z^T = W . X^T: the weights are the A operand (16 x 16 tiles, lane l holds
// A[l % 16][4 (l / 16) + t], t = 0..3), the cells the B operand (lane l holds
// B[4 (l / 16) + t][l % 16]) and the result D (lane l holds D[4 (l / 16) + r]
// [l % 16], r = 0..3) is already the next layer's B operand for the same 16
// cells -- no data exchange between the layers.
#pragma once

#include <hip/hip_runtime.h>

#include "shapes.h"

namespace sdfibm_dev {

constexpr int kSgIn = 11, kSgPad = 16, kSgHid = 32;
constexpr int kSgParams = kSgHid * kSgIn + kSgHid + kSgHid * kSgHid + kSgHid + kSgHid + 1;  // 1473
constexpr int kSgOffB1 = kSgHid * kSgIn, kSgOffW2 = kSgOffB1 + kSgHid, kSgOffB2 = kSgOffW2 + kSgHid * kSgHid;
constexpr int kSgOffW3 = kSgOffB2 + kSgHid, kSgOffB3 = kSgOffW3 + kSgHid;

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// FP32 -> BF16 with round to nearest even (finite inputs), and back.
__host__ __device__ inline unsigned short to_bf16(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline float from_bf16(unsigned short h) {
    const unsigned u = (unsigned)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
__host__ __device__ inline float round_bf16(float f) { return from_bf16(to_bf16(f)); }

// Device copy of the network: BF16 weights (W1 padded to 16 inputs), FP32 biases and output layer.
struct SgNet {
    unsigned short w1[kSgHid * kSgPad];
    unsigned short w2[kSgHid * kSgHid];
    float b1[kSgHid], b2[kSgHid], w3[kSgHid], b3;
};

__host__ inline void sg_pack(const float* p, SgNet& n) {
    for (int i = 0; i < kSgHid; ++i) {
        for (int k = 0; k < kSgPad; ++k) n.w1[i * kSgPad + k] = k < kSgIn ? to_bf16(p[i * kSgIn + k]) : 0;
        for (int k = 0; k < kSgHid; ++k) n.w2[i * kSgHid + k] = to_bf16(p[kSgOffW2 + i * kSgHid + k]);
        n.b1[i] = p[kSgOffB1 + i];
        n.b2[i] = p[kSgOffB2 + i];
        n.w3[i] = p[kSgOffW3 + i];
    }
    n.b3 = p[kSgOffB3];
}

// Per-lane A operands of the weights.
struct SgRegs {
    bf16x4_t w1[2];
    bf16x4_t w2[2][2];
};

__device__ inline bf16x4_t ld4(const unsigned short* p) {
    bf16x4_t v;
    v[0] = (short)p[0];
    v[1] = (short)p[1];
    v[2] = (short)p[2];
    v[3] = (short)p[3];
    return v;
}

__device__ inline void sg_load(const SgNet& n, SgRegs& r, int lane) {
    const int i = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        r.w1[m] = ld4(n.w1 + (16 * m + i) * kSgPad + 4 * kg);
#pragma unroll
        for (int s = 0; s < 2; ++s) r.w2[m][s] = ld4(n.w2 + (16 * m + i) * kSgHid + 16 * s + 4 * kg);
    }
}

// Network output for the cell of every lane of the wave (all 64 lanes must be
// active).  feat: this wave's features in LDS, [64][16] BF16, row = lane.
__device__ inline float sg_infer(const SgNet& n, const SgRegs& r, const unsigned short* feat, int lane) {
    const int j = lane & 15, kg = lane >> 4;
    float mine = 0.0f;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const bf16x4_t x = ld4(feat + (16 * g + j) * kSgPad + 4 * kg);
        const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
        const f32x4_t z0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(r.w1[0], x, zero, 0, 0, 0);
        const f32x4_t z1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(r.w1[1], x, zero, 0, 0, 0);
        bf16x4_t h0, h1;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            h0[t] = (short)to_bf16(fmaxf(z0[t] + n.b1[4 * kg + t], 0.f));
            h1[t] = (short)to_bf16(fmaxf(z1[t] + n.b1[16 + 4 * kg + t], 0.f));
        }
        f32x4_t u0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(r.w2[0][0], h0, zero, 0, 0, 0);
        u0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(r.w2[0][1], h1, u0, 0, 0, 0);
        f32x4_t u1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(r.w2[1][0], h0, zero, 0, 0, 0);
        u1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(r.w2[1][1], h1, u1, 0, 0, 0);
        float part = 0.f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            part += round_bf16(fmaxf(u0[t] + n.b2[4 * kg + t], 0.f)) * n.w3[4 * kg + t];
            part += round_bf16(fmaxf(u1[t] + n.b2[16 + 4 * kg + t], 0.f)) * n.w3[16 + 4 * kg + t];
        }
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (kg == g) mine = part + n.b3;
    }
    return mine;
}

// Features of a cell: 8 corner distances / h (cellPoints order), the unit SDF
// gradient at the centre; the rest zero.
__device__ inline void sg_features(const float* phi8, float inv_h, const float* grad, unsigned short* out16) {
#pragma unroll
    for (int v = 0; v < 8; ++v) out16[v] = to_bf16(phi8[v] * inv_h);
    out16[8] = to_bf16(grad[0]);
    out16[9] = to_bf16(grad[1]);
    out16[10] = to_bf16(grad[2]);
#pragma unroll
    for (int k = 11; k < kSgPad; ++k) out16[k] = 0;
}

// Unit gradient of a quadric shape's SDF at local point l, returned in world
// axes (R l_grad with the rotation rt = R^T rows): sphere/circle (x + com),
// ellipsoid/ellipse x_i / a_i^2.  Other shapes: 0 (the surrogate is not used).
__device__ inline bool quadric_grad(const SolidK& K, d3 l, float* g) {
    double gx, gy, gz;
    switch (K.shape) {
    case SDFIBM_SHAPE_SPHERE: gx = l.x + K.com[0], gy = l.y + K.com[1], gz = l.z + K.com[2]; break;
    case SDFIBM_SHAPE_CIRCLE: gx = l.x + K.com[0], gy = l.y + K.com[1], gz = 0.0; break;
    case SDFIBM_SHAPE_ELLIPSOID: gx = l.x * K.k[0], gy = l.y * K.k[1], gz = l.z * K.k[2]; break;
    case SDFIBM_SHAPE_ELLIPSE: gx = (l.x + K.com[0]) * K.k[0], gy = (l.y + K.com[1]) * K.k[1], gz = 0.0; break;
    default: g[0] = g[1] = g[2] = 0.f; return false;
    }
    // world = R local, R = rt^T
    const double wx = K.rt[0] * gx + K.rt[3] * gy + K.rt[6] * gz;
    const double wy = K.rt[1] * gx + K.rt[4] * gy + K.rt[7] * gz;
    const double wz = K.rt[2] * gx + K.rt[5] * gy + K.rt[8] * gz;
    const double m = sqrt(wx * wx + wy * wy + wz * wz);
    if (!(m > 0)) {
        g[0] = g[1] = g[2] = 0.f;
        return false;
    }
    g[0] = (float)(wx / m), g[1] = (float)(wy / m), g[2] = (float)(wz / m);
    return true;
}

// cos of the angle between the centre gradient g and the corner-difference
// gradient of the 8 distances (the fallback test); ih = 1 / cell size per axis.
__device__ inline float corner_alignment(const float* phi8, const float* g, float ihx, float ihy, float ihz) {
    const float dx = ((phi8[1] + phi8[3] + phi8[5] + phi8[7]) - (phi8[0] + phi8[2] + phi8[4] + phi8[6])) * ihx;
    const float dy = ((phi8[2] + phi8[3] + phi8[6] + phi8[7]) - (phi8[0] + phi8[1] + phi8[4] + phi8[5])) * ihy;
    const float dz = ((phi8[4] + phi8[5] + phi8[6] + phi8[7]) - (phi8[0] + phi8[1] + phi8[2] + phi8[3])) * ihz;
    const float m = sqrtf(dx * dx + dy * dy + dz * dz);
    if (!(m > 0)) return -1.f;
    return (dx * g[0] + dy * g[1] + dz * g[2]) / m;
}

}  // namespace sdfibm_dev
