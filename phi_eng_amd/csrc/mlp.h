// mlp.h — the surrogate's 2-layer MLP (16 -> 128 GELU -> 16) on MFMA.
//
// BF16 weights and activations feed MFMA tiles with FP32 accumulation;
// biases, GELU and the loss are FP32.  A block of 256 threads evaluates 256
// rows: wave w owns rows [64 w, 64 w + 64) as four 16-row tiles.  Products
// that reduce over the 128 hidden units (the forward's second layer) use
// gfx950's v_mfma_f32_16x16x32_bf16; those that reduce over 16 (the first
// layer's 16 inputs, and the backward's products over 16 outputs or a
// 16-row tile) use v_mfma_f32_16x16x16_bf16.  Lane layouts (lane l,
// r = l % 16, q = l / 16): K = 16: A[r][4q + i], B[4q + i][r], i = 0..3;
// K = 32: A[r][8q + e], B[8q + e][r], e = 0..7; both: D[4q + i][r].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MLP_IN 16
#define MLP_HID 128
#define MLP_OUT 16
#define MLP_ROWS 256
#define MLP_HPAD 136  // LDS row pitch (bf16) of a hidden tile
// fp32 parameter vector: W1 [IN][HID], b1 [HID], W2 [HID][OUT], b2 [OUT]
#define MLP_OFF_W1 0
#define MLP_OFF_B1 (MLP_IN * MLP_HID)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_HID)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_HID * MLP_OUT)
#define MLP_NPARAM (MLP_OFF_B2 + MLP_OUT)

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef short mlp_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

struct MlpDev {
    const uint16_t *w1;  // bf16 [IN][HID]
    const float *b1;     // [HID]
    const uint16_t *w2;  // bf16 [HID][OUT]
    const float *b2;     // [OUT]
};

// fp32 -> bf16, round to nearest even (finite inputs).
__host__ __device__ __forceinline__ uint16_t mlp_bf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
// The same rounding by the hardware converter (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t mlp_bf16_hw(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__host__ __device__ __forceinline__ float mlp_f32(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// tanh-form GELU, 0.5 z (1 + tanh(u)) with u = k0 (z + k1 z^3), written as
// z * sigmoid(2u) with the hardware exp2 and reciprocal (about 1e-6
// relative, far below the bf16 rounding of the hidden layer), and its
// derivative sigmoid + z sigmoid (1 - sigmoid) d(2u)/dz.
__device__ __forceinline__ float mlp_gelu(float z, float *dgelu) {
    const float c0 = 2.0f * 0.7978845608028654f, c1 = 2.0f * 0.7978845608028654f * 0.044715f;
    const float l2e = 1.4426950408889634f;  // exp(-v) = exp2(-v log2 e), folded into the constants
    const float z2 = z * z;
    const float sg = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(-l2e * c1, z2, -l2e * c0)));
    if (dgelu) *dgelu = fmaf(z * sg * (1.0f - sg), fmaf(3.0f * c1, z2, c0), sg);
    return z * sg;
}

__device__ __forceinline__ mlp_f32x4 mlp_mfma(mlp_bf16x4 a, mlp_bf16x4 b, mlp_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
typedef __bf16 mlp_bf16v8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ mlp_f32x4 mlp_mfma32(mlp_bf16x8 a, mlp_bf16x8 b, mlp_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_bf16v8, a),
                                                   __builtin_bit_cast(mlp_bf16v8, b), c, 0, 0, 0);
}

// Forward of the block's 256 rows: inputs in xs (bf16), outputs to ys
// (fp32).  All 256 threads must call it.
__device__ __forceinline__ void mlp_forward_block(const MlpDev &m, uint16_t (*xs)[MLP_IN],
                                                  uint16_t (*hs)[16][MLP_HPAD], float (*ys)[MLP_OUT]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
    mlp_bf16x4 b1v[8];
    mlp_bf16x8 b2v[4];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) b1v[nt][i] = (short)m.w1[(4 * q + i) * MLP_HID + nt * 16 + r];
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int e = 0; e < 8; ++e) b2v[kt][e] = (short)m.w2[(kt * 32 + 8 * q + e) * MLP_OUT + r];
    __syncthreads();
    for (int mt = 0; mt < 4; ++mt) {
        const int r0 = w * 64 + mt * 16;
        mlp_bf16x4 a;
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = (short)xs[r0 + r][4 * q + i];
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            mlp_f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
            z = mlp_mfma(a, b1v[nt], z);
            const float bias = m.b1[nt * 16 + r];
#pragma unroll
            for (int i = 0; i < 4; ++i) hs[w][4 * q + i][nt * 16 + r] = mlp_bf16(mlp_gelu(z[i] + bias, nullptr));
        }
        __syncthreads();
        mlp_f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            mlp_bf16x8 a2;
#pragma unroll
            for (int e = 0; e < 8; ++e) a2[e] = (short)hs[w][r][kt * 32 + 8 * q + e];
            y = mlp_mfma32(a2, b2v[kt], y);
        }
        const float bias2 = m.b2[r];
#pragma unroll
        for (int i = 0; i < 4; ++i) ys[r0 + 4 * q + i][r] = y[i] + bias2;
        __syncthreads();
    }
}

// Backward of the block's rows: recomputes the forward, takes dY (bf16 in
// dys; the b2 gradient is the column sum of these bf16 values) and adds
// the block's parameter gradients into grad (fp32, global, atomics).  Rows
// without a sample must have dY = 0.
__device__ __forceinline__ void mlp_backward_block(const MlpDev &m, uint16_t (*xs)[MLP_IN],
                                                   uint16_t (*hs)[16][MLP_HPAD], uint16_t (*dzs)[16][MLP_HPAD],
                                                   uint16_t (*dys)[MLP_OUT], float *__restrict__ grad) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, q = lane >> 4;
    mlp_bf16x4 b1v[8], btv[8];
#pragma unroll
    for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            b1v[nt][i] = (short)m.w1[(4 * q + i) * MLP_HID + nt * 16 + r];
            btv[nt][i] = (short)m.w2[(nt * 16 + r) * MLP_OUT + 4 * q + i];  // W2^T as [out][hid]
        }
    mlp_f32x4 gw1[8], gw2[8];
    float gb1[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        gw1[t] = mlp_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        gw2[t] = mlp_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        gb1[t] = 0.0f;
    }
    float gb2 = 0.0f;
    __syncthreads();
    for (int mt = 0; mt < 4; ++mt) {
        const int r0 = w * 64 + mt * 16;
        mlp_bf16x4 a, ady;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = (short)xs[r0 + r][4 * q + i];
            ady[i] = (short)dys[r0 + r][4 * q + i];
        }
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            mlp_f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
            z = mlp_mfma(a, b1v[nt], z);
            mlp_f32x4 dh = {0.0f, 0.0f, 0.0f, 0.0f};
            dh = mlp_mfma(ady, btv[nt], dh);  // dH = dY W2^T
            const float bias = m.b1[nt * 16 + r];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float dg;
                const float hv = mlp_gelu(z[i] + bias, &dg);
                const float dzv = dh[i] * dg;
                hs[w][4 * q + i][nt * 16 + r] = mlp_bf16(hv);
                dzs[w][4 * q + i][nt * 16 + r] = mlp_bf16(dzv);
                gb1[nt] += dzv;
            }
        }
        __syncthreads();
        // dW2 += H^T dY (hidden tile ht rows, K = the 16 rows of this tile)
        mlp_bf16x4 bdy;
#pragma unroll
        for (int i = 0; i < 4; ++i) bdy[i] = (short)dys[r0 + 4 * q + i][r];
#pragma unroll
        for (int ht = 0; ht < 8; ++ht) {
            mlp_bf16x4 ah;
#pragma unroll
            for (int i = 0; i < 4; ++i) ah[i] = (short)hs[w][4 * q + i][ht * 16 + r];
            gw2[ht] = mlp_mfma(ah, bdy, gw2[ht]);
        }
        // dW1 += X^T dZ (features x hidden, K = the 16 rows)
        mlp_bf16x4 axt;
#pragma unroll
        for (int i = 0; i < 4; ++i) axt[i] = (short)xs[r0 + 4 * q + i][r];
#pragma unroll
        for (int nt = 0; nt < 8; ++nt) {
            mlp_bf16x4 bdz;
#pragma unroll
            for (int i = 0; i < 4; ++i) bdz[i] = (short)dzs[w][4 * q + i][nt * 16 + r];
            gw1[nt] = mlp_mfma(axt, bdz, gw1[nt]);
        }
        if (q == 0)
            for (int i = 0; i < 16; ++i) gb2 += mlp_f32(dys[r0 + i][r]);
        __syncthreads();
    }
    // gw2[ht][i] = dW2[ht*16 + 4q + i][r]; gw1[nt][i] = dW1[4q + i][nt*16 + r]
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            atomicAdd(&grad[MLP_OFF_W2 + (t * 16 + 4 * q + i) * MLP_OUT + r], gw2[t][i]);
            atomicAdd(&grad[MLP_OFF_W1 + (4 * q + i) * MLP_HID + t * 16 + r], gw1[t][i]);
        }
    // bias gradients: gb1[nt] sums 4 rows per lane; reduce over q (lanes r, r+16, r+32, r+48)
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        float v = gb1[t];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        if (q == 0) atomicAdd(&grad[MLP_OFF_B1 + t * 16 + r], v);
    }
    if (q == 0) atomicAdd(&grad[MLP_OFF_B2 + r], gb2);
}
