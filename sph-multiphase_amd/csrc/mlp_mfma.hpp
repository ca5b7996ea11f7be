// The learned warm start's network on MFMA: the weights' operand layout,
// staged in LDS by every workgroup that evaluates it, and the forward pass
// of one wave's 64 particles.  Shared by the inference kernels and the a_ii
// pass, which evaluates the network on its own features.
#pragma once

#include "mlp_feat.hpp"

#include <hip/hip_fp16.h>

namespace sphx {

constexpr int MLP_W1 = 0, MLP_B1 = MLP_NF * MLP_H, MLP_W2 = MLP_B1 + MLP_H,
              MLP_B2 = MLP_W2 + MLP_H * MLP_H, MLP_W3 = MLP_B2 + MLP_H,
              MLP_B3 = MLP_W3 + MLP_H * MLP_NO, MLP_NP = MLP_B3 + MLP_NO;   // 6020

// ------------------------------------------------- inference on MFMA --
// One wave runs 64 particles as four 16-particle column tiles on
// v_mfma_f32_16x16x32_f16 with every layer transposed (H^T = W^T X^T).  A
// layer's accumulators (rows 4 (lane / 16) + i, column lane % 16) are then
// the next layer's B operand (k = 8 (lane / 16) .. +8, same column) once the
// K index within each 32-unit chunk is read as u(g, s) = s < 4 ? 4 g + s :
// 16 + 4 g + s - 4; the weights are staged in that order.  Weights are the
// network's FP16 values exactly; activations enter as two FP16 pieces
// (x = hi + lo, two products each), so a layer's result is its FP32-input
// value to about 2^-22, accumulated in FP32.
typedef __attribute__((ext_vector_type(8))) _Float16 h8;
typedef __attribute__((ext_vector_type(4))) float f4v;

struct MlpMfmaShared {
  uint4 a1[4][64];      // W1^T, M tile m: lane (r, g) holds W1[8 g + s][16 m + r]
  uint4 a2[4][2][64];   // W2^T, M tile m, K chunk c: W2[32 c + u(g, s)][16 m + r]
  uint4 a3[2][64];      // W3^T (rows 4..15 zero), K chunk c: W3[32 c + u(g, s)][r]
  float b1[MLP_H], b2[MLP_H], b3[MLP_NO];
};

__device__ __forceinline__ int mfma_unit(int g, int s) { return s < 4 ? 4 * g + s : 16 + 4 * g + s - 4; }

__device__ inline void stage_mfma(MlpMfmaShared& S, const __half* __restrict__ w16) {
  const uint16_t* w = reinterpret_cast<const uint16_t*>(w16);
  uint16_t* a1 = reinterpret_cast<uint16_t*>(&S.a1[0][0]);
  uint16_t* a2 = reinterpret_cast<uint16_t*>(&S.a2[0][0][0]);
  uint16_t* a3 = reinterpret_cast<uint16_t*>(&S.a3[0][0]);
  for (int t = threadIdx.x; t < 4 * 64 * 8; t += blockDim.x) {
    const int sl = t & 7, lane = (t >> 3) & 63, m = t >> 9;
    const int r = lane & 15, g = lane >> 4, f = 8 * g + sl;
    a1[t] = f < MLP_NF ? w[MLP_W1 + f * MLP_H + 16 * m + r] : (uint16_t)0;
  }
  for (int t = threadIdx.x; t < 4 * 2 * 64 * 8; t += blockDim.x) {
    const int sl = t & 7, lane = (t >> 3) & 63, c = (t >> 9) & 1, m = t >> 10;
    const int r = lane & 15, g = lane >> 4, u = 32 * c + mfma_unit(g, sl);
    a2[t] = w[MLP_W2 + u * MLP_H + 16 * m + r];
  }
  for (int t = threadIdx.x; t < 2 * 64 * 8; t += blockDim.x) {
    const int sl = t & 7, lane = (t >> 3) & 63, c = t >> 9;
    const int r = lane & 15, g = lane >> 4, u = 32 * c + mfma_unit(g, sl);
    a3[t] = r < MLP_NO ? w[MLP_W3 + u * MLP_NO + r] : (uint16_t)0;
  }
  for (int t = threadIdx.x; t < MLP_H; t += blockDim.x) {
    S.b1[t] = __half2float(w16[MLP_B1 + t]);
    S.b2[t] = __half2float(w16[MLP_B2 + t]);
    if (t < MLP_NO) S.b3[t] = __half2float(w16[MLP_B3 + t]);
  }
}

__device__ __forceinline__ h8 as_h8(uint4 v) { return __builtin_bit_cast(h8, v); }

constexpr int MLP_MFMA_U4 = (int)(sizeof(MlpMfmaShared) / sizeof(uint4));
static_assert(sizeof(MlpMfmaShared) % sizeof(uint4) == 0, "MFMA weight block is whole uint4s");

__device__ __forceinline__ void load_mfma(MlpMfmaShared& S, const uint4* __restrict__ wmf) {
  uint4* dst = reinterpret_cast<uint4*>(&S);
  for (int t = threadIdx.x; t < MLP_MFMA_U4; t += blockDim.x) dst[t] = wmf[t];
}

// x = hi + lo, both FP16
__device__ __forceinline__ void split_h8(const float v[8], h8& hi, h8& lo) {
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const _Float16 h = (_Float16)v[s];
    hi[s] = h;
    lo[s] = (_Float16)(v[s] - (float)h);
  }
}

// Column tiles (of 16 particles) per pass of the forward: four passes of
// one keep the live accumulators and operands at a quarter (142 VGPRs, three
// waves per SIMD; two tiles per pass need 216, two waves).
constexpr int MLP_TPP = 1;

// Next layer's B operands from this layer's accumulators acc[m][t] (4 M
// tiles of 16 units), with bias and ReLU.
__device__ __forceinline__ void relu_to_b(const f4v acc[4][MLP_TPP], const float* bias, int g,
                                          h8 bh[2][MLP_TPP], h8 bl[2][MLP_TPP]) {
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int t = 0; t < MLP_TPP; ++t) {
      float v[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int m = 2 * c + (s < 4 ? 0 : 1), i = s & 3;
        v[s] = fmaxf(acc[m][t][i] + bias[32 * c + mfma_unit(g, s)], 0.f);
      }
      split_h8(v, bh[c][t], bl[c][t]);
    }
}

// Network outputs of one wave's 64 particles into y[64][4] (LDS): rows
// points at the first particle's 24 features (rows of 24 floats, global
// memory or LDS); rows past `valid` repeat row valid - 1.  Every lane of
// the wave must take part (MFMA).
__device__ __forceinline__ void mlp_forward_mfma(const MlpMfmaShared& S, const float* rows,
                                                 int valid, float* y) {
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  const f4v zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (int t0 = 0; t0 < 4; t0 += MLP_TPP) {
    h8 xh[MLP_TPP], xl[MLP_TPP];
#pragma unroll
    for (int t = 0; t < MLP_TPP; ++t) {
      const int p = min(16 * (t0 + t) + r, valid - 1);
      float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (8 * g < MLP_NF) {
        const float4* src = reinterpret_cast<const float4*>(rows + p * MLP_NF + 8 * g);
        const float4 u0 = src[0], u1 = src[1];
        v[0] = u0.x; v[1] = u0.y; v[2] = u0.z; v[3] = u0.w;
        v[4] = u1.x; v[5] = u1.y; v[6] = u1.z; v[7] = u1.w;
      }
      split_h8(v, xh[t], xl[t]);
    }
    f4v acc[4][MLP_TPP];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const h8 a = as_h8(S.a1[m][lane]);
#pragma unroll
      for (int t = 0; t < MLP_TPP; ++t) {
        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, xh[t], zero, 0, 0, 0);
        acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, xl[t], acc[m][t], 0, 0, 0);
      }
    }
    h8 bh[2][MLP_TPP], bl[2][MLP_TPP];
    relu_to_b(acc, S.b1, g, bh, bl);
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int t = 0; t < MLP_TPP; ++t) {
        f4v e = zero;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const h8 a = as_h8(S.a2[m][c][lane]);
          e = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bh[c][t], e, 0, 0, 0);
          e = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bl[c][t], e, 0, 0, 0);
        }
        acc[m][t] = e;
      }
    relu_to_b(acc, S.b2, g, bh, bl);
#pragma unroll
    for (int t = 0; t < MLP_TPP; ++t) {
      f4v o = zero;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const h8 a = as_h8(S.a3[c][lane]);
        o = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bh[c][t], o, 0, 0, 0);
        o = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, bl[c][t], o, 0, 0, 0);
      }
      if (g == 0)   // rows 0..3 are the four outputs
#pragma unroll
        for (int i = 0; i < MLP_NO; ++i) y[(16 * (t0 + t) + r) * MLP_NO + i] = o[i] + S.b3[i];
    }
  }
}

// The learned start of fluid particle i from the network's outputs y[4]:
// p0 = max(0, p(t-1) + P y0 / 10) (the plain start warm p(t-1) when the
// network returns NaN) into press_l, and s = m p0 / rho^2 into ps.w, when
// `warm`; the drift correction for the next step into dcorr when non-null.
// qi = phys[i] = {m, rho, p(t-1), rho0}.
__device__ __forceinline__ void mlp_start_out(const KP& k, const MLPP& mp, int warm, int i,
                                              const float* y, const float4& qi,
                                              float* __restrict__ press_l,
                                              float4* __restrict__ ps,
                                              float4* __restrict__ dcorr) {
  if (warm) {
    float p0 = fmaxf(0.f, fmaxf(0.f, qi.z) + y[0] * (mp.p_scale / MLP_TSCALE));
    if (!(p0 == p0)) p0 = k.warm * qi.z;   // NaN from the network: plain start
    press_l[i] = p0;
    ps[i].w = qi.x * p0 / (qi.y * qi.y);
  }
  if (dcorr) {
    float sc = mp.drift_blend * mp.u_scale / MLP_TSCALE;
    float4 d = make_float4(y[1] * sc, y[2] * sc, y[3] * sc, dcorr[i].w);
    if (!(d.x == d.x && d.y == d.y && d.z == d.z)) d = make_float4(0.f, 0.f, 0.f, d.w);
    dcorr[i] = d;
  }
}

}  // namespace sphx
