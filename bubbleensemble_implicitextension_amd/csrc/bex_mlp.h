// Warm-start network of the BiCG coupling solve: shared definitions (host and device).
//
// The network maps five per-unit features to a factor y, and the BiCG initial guess is
// x0 = b * y (b = right-hand side of the coupling system, the converged solution divided by b is
// the training target). No reference counterpart: the reference always starts BiCG from the
// uncoupled acceleration (CoupledSystems_PerBlock_ExplicitRungeKutta_Steppers.cuh:494-499).
//
// Architecture and arithmetic (identical in training, inference and the host oracle):
//   in_k  = bf16(feature_k), k = 0..4;  in_5 = 1 (bias input)
//   z1_i  = sum_k bf16(W1[k][i]) * in_k          (FP32 accumulation),   i = 0..63
//   h1_i  = bf16(max(z1_i, 0))
//   z2_i  = b2_i + sum_k bf16(W2[k][i]) * h1_k    (FP32 accumulation)
//   h2_i  = bf16(max(z2_i, 0))
//   y     = b3 + sum_i bf16(w3_i) * h2_i          (FP32 accumulation)
// bf16() rounds an FP32 value to the nearest bfloat16, ties to even. Inside the solver the two
// hidden layers run on MFMA (v_mfma_f32_32x32x16_bf16) and the output layer on the VALU.
#pragma once

#include <stdint.h>

#define BEX_MLP_IN 5        // features per unit
#define BEX_MLP_IN_PAD 8    // K of the first MFMA: 5 features, the bias input, 2 zeros
#define BEX_MLP_HIDDEN 64
// Flat FP32 parameter vector: W1[6][64] (row 5 = first-layer bias), W2[64][64] (row = input),
// b2[64], w3[64], b3.
#define BEX_MLP_OFF_W1 0
#define BEX_MLP_OFF_W2 (6 * BEX_MLP_HIDDEN)
#define BEX_MLP_OFF_B2 (BEX_MLP_OFF_W2 + BEX_MLP_HIDDEN * BEX_MLP_HIDDEN)
#define BEX_MLP_OFF_W3 (BEX_MLP_OFF_B2 + BEX_MLP_HIDDEN)
#define BEX_MLP_OFF_B3 (BEX_MLP_OFF_W3 + BEX_MLP_HIDDEN)
#define BEX_MLP_NPARAM (BEX_MLP_OFF_B3 + 1)  // 4609
// Clamp of the training target x / b (and of the guess factor).
#define BEX_MLP_TARGET_CLAMP 5.0f
// Training-sample row: 5 features, target x / b (clamped), x, b, uncoupled acceleration, 0.
#define BEX_SAMPLE_FLOATS 10

#if defined(__HIP__) || defined(__HIPCC__)
#define BEX_MLP_HD __host__ __device__ inline
#else
#define BEX_MLP_HD inline
#endif

// Nearest bfloat16 (ties to even) of a finite FP32 value, returned as its 16 bits.
BEX_MLP_HD uint16_t bex_bf16_bits(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
BEX_MLP_HD float bex_bf16_value(uint16_t b) {
  return __builtin_bit_cast(float, (uint32_t)b << 16);
}
BEX_MLP_HD float bex_bf16_round(float f) { return bex_bf16_value(bex_bf16_bits(f)); }

// Device-side copy of a trained network in the solver's operand layout (built on the host by
// bex_set_warm_start_network):
//   a1[t][c]         : MFMA A operand of layer 1 for the primal-half lane c (c = 0..31) of tile t
//                      (hidden rows 32t..32t+31), 8 bf16; the shadow half supplies zeros
//   a2[t][s][lane]   : MFMA A operand of layer 2, tile t, k-step s, 8 bf16
//   b2, w3 (bf16-rounded, stored as FP32), b3
struct BexMlpDevice {
  uint64_t a1[2][32][2];
  uint64_t a2[2][4][64][2];
  float b2[BEX_MLP_HIDDEN];
  float w3[BEX_MLP_HIDDEN];
  float b3;
  float pad[3];
};

// MFMA operand layout of v_mfma_f32_32x32x16_bf16 (one block): the k-slot (g, e), g = 0..1,
// e = 0..7, pairs element e of lane i + 32 g of the A operand (row i) with element e of lane
// j + 32 g of the B operand (column j); lane l holds D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// in accumulator r = 0..15.
//
// Layer 1 computes H1^T (hidden x units) = W1^T (hidden x inputs) * X^T with the five features,
// the bias input 1 and two zeros in the slots (0, 0..7) and zeros in the slots (1, 0..7).
// Layer 2 computes H2^T = W2^T * H1^T in four k-steps; k-step s feeds element e of lane l's B
// operand from layer-1 accumulator 8 (s % 2) + e of tile s / 2, i.e. hidden input
// bex_mlp_k2(s, l / 32, e), so activations never leave the registers. A2 is packed with the
// W2 rows of the same hidden inputs.
BEX_MLP_HD int bex_mlp_k2(int s, int g, int e) {
  const int r = 8 * (s & 1) + e;
  return 32 * (s >> 1) + 8 * (r >> 2) + 4 * g + (r & 3);
}

inline void bex_mlp_pack_device(const float* p, BexMlpDevice* d) {
  for (int t = 0; t < 2; t++)
    for (int c = 0; c < 32; c++) {
      const int i = 32 * t + c;
      for (int w = 0; w < 2; w++) {
        uint64_t v = 0;
        for (int e = 0; e < 4; e++) {
          const int k = 4 * w + e;
          const float x = (k < 6) ? p[BEX_MLP_OFF_W1 + k * BEX_MLP_HIDDEN + i] : 0.0f;
          v |= (uint64_t)bex_bf16_bits(x) << (16 * e);
        }
        d->a1[t][c][w] = v;
      }
    }
  for (int t = 0; t < 2; t++)
    for (int s = 0; s < 4; s++)
      for (int lane = 0; lane < 64; lane++) {
        const int i = 32 * t + (lane & 31);
        for (int w = 0; w < 2; w++) {
          uint64_t v = 0;
          for (int e = 0; e < 4; e++) {
            const int k = bex_mlp_k2(s, lane >> 5, 4 * w + e);
            v |= (uint64_t)bex_bf16_bits(p[BEX_MLP_OFF_W2 + k * BEX_MLP_HIDDEN + i]) << (16 * e);
          }
          d->a2[t][s][lane][w] = v;
        }
      }
  for (int i = 0; i < BEX_MLP_HIDDEN; i++) {
    d->b2[i] = p[BEX_MLP_OFF_B2 + i];
    d->w3[i] = bex_bf16_round(p[BEX_MLP_OFF_W3 + i]);
  }
  d->b3 = p[BEX_MLP_OFF_B3];
  d->pad[0] = d->pad[1] = d->pad[2] = 0.0f;
}
