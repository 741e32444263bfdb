// Learned surrogate of the three-centre table interpolation (no counterpart in
// the reference).  One network per 3c group (species triple, charge subtype):
//   inputs  u = (2x/xmax - 1, 2y/ymax - 1, cos theta)
//   hidden  two layers of 64 units, SiLU
//   outputs 64 slots holding the 5 Legendre coefficients of every matrix
//           element, i.e. the theta cells of getMEs_Fdata_3c
//           (e.FDATA/M_Fdata_3c.f90:391-489) before the fold, divided by a
//           fitted envelope exp(c - kx x - ky y) of the group and scaled per
//           slot, so that the network's relative accuracy carries over to
//           the small far-field values.
// Training runs in FP32; inference runs on BF16 MFMA with FP32 accumulation.
//
// Inference layout (one wavefront, 16 samples, v_mfma_f32_16x16x16_bf16): each
// layer is evaluated transposed, Z^T = W * A^T, so that the accumulator tile
// of one layer (lane l: units 4*(l/16)+j, sample l%16) is already the B
// operand of the next.  Output slot o = 16*mt + 4*q + j sits in lane group
// q = l/16, register mt*4 + j; the slot map below puts all five theta values
// of a matrix element in one lane group, so each lane folds its own elements.
#pragma once

#include "fb_internal.h"

#define MLP_IN 3
#define MLP_H 64
#define MLP_O 64
#define MLP_P_W1 0
#define MLP_P_B1 (MLP_P_W1 + MLP_H * MLP_IN)
#define MLP_P_W2 (MLP_P_B1 + MLP_H)
#define MLP_P_B2 (MLP_P_W2 + MLP_H * MLP_H)
#define MLP_P_W3 (MLP_P_B2 + MLP_H)
#define MLP_P_B3 (MLP_P_W3 + MLP_O * MLP_H)
#define MLP_NP (MLP_P_B3 + MLP_O)   // 8576 FP32 parameters

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// slot o -> (theta, matrix element); false for padding slots
__host__ __device__ inline bool mlp_slot(int o, int nme, int *th, int *m) {
  const int mt = o >> 4, q = (o >> 2) & 3, j = o & 3;
  const int r = mt * 4 + j;
  const int ml = r / 5;
  *th = r - ml * 5;
  *m = q + 4 * ml;
  return ml < 3 && *m < nme;
}

// device image of one trained network
struct DMlp {
  float w1[MLP_H * MLP_IN], b1[MLP_H], b2[MLP_H], b3[MLP_O];
  float oscale[MLP_O];        // slot scale: coefficient (eV) = output x oscale x envelope
  float xs, ys;               // 2/xmax, 2/ymax
  float xmax, ymax;
  float ec, ekx, eky;         // envelope exp(ec - ekx x - eky y), x and y in Angstrom
  int nme, nx, ny, ready;
  int64_t err_off;            // per-cell calibrated error, floats [ny][nx] (1-based ix, iy - 1)
  int pad;
  unsigned short w2op[4][4][64][4];   // BF16 A-operand tiles [mt][kb][lane][j]
  unsigned short w3op[4][4][64][4];
};

__device__ __forceinline__ float mlp_envelope(const DMlp *w, float x, float y) {
  return __expf(w->ec - w->ekx * x - w->eky * y);
}

// SiLU and its derivative (FP32, hardware exp and reciprocal); training and
// inference use the same function
__device__ __forceinline__ float mlp_sigmoid(float z) { return __frcp_rn(1.0f + __expf(-z)); }
__device__ __forceinline__ float mlp_silu(float z) { return z * mlp_sigmoid(z); }
__device__ __forceinline__ float mlp_dsilu(float z) {
  const float s = mlp_sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__host__ __device__ inline unsigned short mlp_bf16(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// Evaluate the network for sample lane%16 of the wavefront (inputs already
// normalised); out[mt*4 + j] receives slot 16*mt + 4*(lane/16) + j.  Every lane
// of the wavefront must call it (MFMA).
__device__ inline void mlp16(const DMlp *w, float u0, float u1, float u2, float out[16]) {
  const int l = threadIdx.x & 63, q = l >> 4;
  mlp_s4 h[4];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int u = kb * 16 + 4 * q + j;
      float z = w->w1[u * 3 + 0] * u0;
      z = __builtin_fmaf(w->w1[u * 3 + 1], u1, z);
      z = __builtin_fmaf(w->w1[u * 3 + 2], u2, z);
      z = z + w->b1[u];
      h[kb][j] = (short)mlp_bf16(mlp_silu(z));
    }
  }
  mlp_s4 h2[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const mlp_s4 a = *reinterpret_cast<const mlp_s4 *>(&w->w2op[mt][kb][l][0]);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h[kb], acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) h2[mt][j] = (short)mlp_bf16(mlp_silu(acc[j] + w->b2[mt * 16 + 4 * q + j]));
  }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const mlp_s4 a = *reinterpret_cast<const mlp_s4 *>(&w->w3op[mt][kb][l][0]);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h2[kb], acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) out[mt * 4 + j] = acc[j] + w->b3[mt * 16 + 4 * q + j];
  }
}
