// Learned surrogate of the three-centre table interpolation (no counterpart in
// the reference).  One network per 3c group (species triple, charge subtype):
//   inputs  u = (2x/xmax - 1, 2y/ymax - 1, cos theta)
//   hidden  two layers of 64 units, activation a(z) = z * s(z)
//   outputs 64 slots holding the 5 Legendre coefficients of every matrix
//           element, i.e. the theta cells of getMEs_Fdata_3c
//           (e.FDATA/M_Fdata_3c.f90:391-489) before the fold, divided by a
//           fitted envelope exp(c - kx x - ky y) of the group and scaled per
//           slot, so that the network's relative accuracy carries over to
//           the small far-field values.
// The activation is SiLU's shape with a polynomial sigmoid: s(z) = 1/2 +
// z/4 - z|z|/32 on [-4, 4] (s(0) = 1/2 and s'(0) = 1/4 as for the logistic
// function, s = 0 or 1 with zero slope at the ends) and constant outside, a
// C1 function.  It costs four VALU operations and no transcendental (gfx950
// issues exp and reciprocal at half the FMA rate, and SiLU needs one of each
// per unit).
// Training runs in FP32; inference runs on BF16 MFMA with FP32 accumulation.
//
// Inference layout (one wavefront, 16 samples; sample n = lane % 16, lane
// group q = lane / 16): every layer is evaluated transposed, Z^T = W A^T, so
// the FP32 accumulator tile of a layer (lane: rows 4q + j of a 16-row tile,
// column n) becomes the B operand of the next layer in place.  The K index of
// an MFMA is a contraction index, so the hidden units are assigned to its
// (q, i) operand slots in whatever order the accumulator tiles hold them, and
// the weights are packed with the same assignment:
//   layer 1  v_mfma_f32_16x16x16_bf16, K = 16: the inputs and the weights
//            split into BF16 high and low parts, u.w = uh.wh + ul.wh + uh.wl
//            (9 slots), the bias as the FP32 accumulator start;
//   layer 2  v_mfma_f32_16x16x32_bf16 over two K blocks: K block kb, slots
//            (q, i < 4) hold unit 32kb + 4q + i, slots (q, i >= 4) unit
//            32kb + 16 + 4q + i - 4 -- the accumulator tiles 2kb and 2kb+1;
//   layer 3  the same, with activations and weights split into high and low
//            BF16 parts (hh.wh + hl.wh + hh.wl, three MFMAs per tile and K
//            block), the slot scale folded into the weights and the bias.
// Output slot o = 16 mt + 4 q + j sits in lane group q, accumulator tile mt,
// element j; the slot map below puts all five theta values of a matrix element
// in one lane group, so each lane folds its own elements.
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
#define MLP_CSUB 4                  // calibration sub-cells per table cell and direction

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef __bf16 mlp_bf8 __attribute__((ext_vector_type(8)));
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

// MFMA operand fragments of one network, BF16 bits; lane-major so that a
// wavefront reads one fragment with one 8- or 16-byte access per lane
struct MlpW {
  uint2 a1[4][64];          // layer 1: [tile][lane], 4 BF16
  uint4 a2[4][2][64];       // layer 2: [tile][K block][lane], 8 BF16
  uint4 a3h[4][2][64];      // layer 3, high parts of (weight x slot scale)
  uint4 a3l[4][2][64];      // layer 3, low parts
  float b1[MLP_H], b2[MLP_H], b3[MLP_O];   // FP32 biases (b3 times the slot scale)
};

// device image of one trained network
struct DMlp {
  MlpW w;
  float oscale[MLP_O];        // slot scale: coefficient (eV) = output x envelope (scale folded into layer 3)
  float xmax, ymax;
  float ec, ekx, eky;         // envelope exp(ec - ekx x - eky y), x and y in Angstrom
  int nme, nx, ny, ready;
  int64_t err_off;            // calibrated error, floats [(ny-1) CSUB][(nx-1) CSUB] sub-cells
};

__device__ __forceinline__ float mlp_envelope(const DMlp *w, float x, float y) {
  return __expf(w->ec - w->ekx * x - w->eky * y);
}

// sub-cell of (x, y) on the calibration grid (MLP_CSUB per table cell)
__host__ __device__ inline int64_t mlp_subcell(double x, double y, double xmax, double ymax, int nx, int ny) {
  const int sx = (nx - 1) * MLP_CSUB, sy = (ny - 1) * MLP_CSUB;
  int ix = (int)(x / xmax * sx), iy = (int)(y / ymax * sy);
  ix = ix < 0 ? 0 : (ix >= sx ? sx - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= sy ? sy - 1 : iy);
  return (int64_t)iy * sx + ix;
}

// the activation and its derivative (FP32; training and inference use the
// same expressions)
#define MLP_SCLAMP 4.0f
__host__ __device__ __forceinline__ float mlp_sig(float z) {
  const float zc = fminf(fmaxf(z, -MLP_SCLAMP), MLP_SCLAMP);
  return __builtin_fmaf(zc, __builtin_fmaf(fabsf(zc), -0.03125f, 0.25f), 0.5f);
}
__host__ __device__ __forceinline__ float mlp_act(float z) { return z * mlp_sig(z); }
__host__ __device__ __forceinline__ float mlp_dact(float z) {
  // d/dz [z s(z)] = s(z) + z s'(z), s' = 1/4 - |z|/16 inside the clamp, 0 outside
  const float ds = (fabsf(z) < MLP_SCLAMP) ? __builtin_fmaf(fabsf(z), -0.0625f, 0.25f) : 0.0f;
  return __builtin_fmaf(z, ds, mlp_sig(z));
}

// FP32 -> BF16 bits, round to nearest even (finite inputs)
__host__ __device__ inline unsigned short mlp_bf16(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__host__ __device__ inline float mlp_bf16f(unsigned short h) {
  const unsigned u = (unsigned)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ mlp_s4 mlp_s4_of(uint2 v) { return __builtin_bit_cast(mlp_s4, v); }
__device__ __forceinline__ mlp_bf8 mlp_bf8_of(uint4 v) { return __builtin_bit_cast(mlp_bf8, v); }

// this lane's four accumulator rows of bias tile mt
__device__ __forceinline__ mlp_f4 mlp_bias4(const float *b, int mt) {
  const int q = (threadIdx.x & 63) >> 4;
  return *reinterpret_cast<const mlp_f4 *>(b + 16 * mt + 4 * q);
}

// Evaluate the network for sample lane % 16 of the wavefront (inputs already
// normalised); out[4 mt + j] receives slot 16 mt + 4 (lane / 16) + j times its
// slot scale.  Every lane of the wavefront must call it (MFMA).  Each weight
// fragment is read from w just before its MFMA: holding them in registers
// (104 VGPRs) or LDS (27 KB) would halve the occupancy of the 3c kernel that
// calls this, which costs more than the L1 reads.
__device__ __forceinline__ void mlp_eval16(const MlpW *w, float u0, float u1, float u2, float out[16]) {
  const int l = threadIdx.x & 63, q = l >> 4;
  // layer 1 operand: slots q0 (u0h, u0l, u1h, u1l), q1 (u2h, u2l, u0h, u1h),
  // q2 (u2h, 0, 0, 0), q3 zero -- against weights (w0h, w0h, w1h, w1h),
  // (w2h, w2h, w0l, w1l), (w2l, 0, 0, 0)
  const __bf16 h0 = (__bf16)u0, h1 = (__bf16)u1, h2 = (__bf16)u2;
  const __bf16 l0 = (__bf16)(u0 - (float)h0), l1 = (__bf16)(u1 - (float)h1), l2 = (__bf16)(u2 - (float)h2);
  const __bf16 zb = (__bf16)0.0f;
  __bf16 e0 = zb, e1 = zb, e2 = zb, e3 = zb;
  if (q == 0) { e0 = h0; e1 = l0; e2 = h1; e3 = l1; }
  else if (q == 1) { e0 = h2; e1 = l2; e2 = h0; e3 = h1; }
  else if (q == 2) { e0 = h2; }
  const mlp_s4 b1op = {__builtin_bit_cast(short, e0), __builtin_bit_cast(short, e1), __builtin_bit_cast(short, e2),
                       __builtin_bit_cast(short, e3)};
  mlp_f4 z[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    z[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(mlp_s4_of(w->a1[mt][l]), b1op, mlp_bias4(w->b1, mt), 0, 0, 0);
  }
  // activation; K block kb of layer 2 = accumulator tiles 2kb and 2kb + 1
  mlp_bf8 h[2];
#pragma unroll
  for (int kb = 0; kb < 2; kb++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      h[kb][i] = (__bf16)mlp_act(z[2 * kb][i]);
      h[kb][4 + i] = (__bf16)mlp_act(z[2 * kb + 1][i]);
    }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    mlp_f4 c = mlp_bias4(w->b2, mt);
#pragma unroll
    for (int kb = 0; kb < 2; kb++)
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(mlp_bf8_of(w->a2[mt][kb][l]), h[kb], c, 0, 0, 0);
    z[mt] = c;
  }
  // activation split into high and low BF16 parts for layer 3
  mlp_bf8 hh[2], hl[2];
#pragma unroll
  for (int kb = 0; kb < 2; kb++)
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const float a = mlp_act(z[2 * kb + (i >> 2)][i & 3]);
      const __bf16 ah = (__bf16)a;
      hh[kb][i] = ah;
      hl[kb][i] = (__bf16)(a - (float)ah);
    }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    mlp_f4 c = mlp_bias4(w->b3, mt);
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
      const mlp_bf8 a3h = mlp_bf8_of(w->a3h[mt][kb][l]);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3h, hh[kb], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3h, hl[kb], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(mlp_bf8_of(w->a3l[mt][kb][l]), hh[kb], c, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) out[mt * 4 + j] = c[j];
  }
}

