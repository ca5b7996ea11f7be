// Learned warm start and drift closure: a 24-64-64-4 MLP (6,020 weights)
// evaluated per fluid particle after the a_ii pass.  Output 0 predicts the
// pressure increment over the previous step, p0 = max(0, p(t-1) + P*y0/10),
// replacing the warm start p0 = 0.5 p(t-1) of fluid_system_kern.cu:899;
// outputs 1-3 predict the change of the dispersed phase's drift velocity
// when it is evaluated with the converged instead of the previous pressure
// (the lag of multiphase.cu:62, which reads the last step's pressure), and
// are added to that drift at the next step.  The plain solve stays the
// fallback: a learned start that does not reach the tolerance is discarded
// and the step is solved again from the reference warm start.
//
// Training is on the GPU with these kernels: the solver records (features,
// converged increments) pairs, k_mlp_train_mfma runs forward, backward and
// the weight gradients of a batch on FP32 MFMA (each wave one row of
// partial gradients), k_mlp_gsum adds the rows in a fixed order, and
// k_mlp_adam applies Adam and refreshes the FP16 copy the inference reads
// (FP16 weights, FP32 accumulation).
#include "mlp_mfma.hpp"

#include <hip/hip_fp16.h>

#include <cmath>
#include <random>

namespace sphx {

// ------------------------------------------------------------- features --
// Stand-alone feature pass (the a_ii pass computes the same map in its own
// neighbour loop when the network is in use).
__global__ __launch_bounds__(256) void k_mlp_features(
    KP k, NbrList L, MLPP mp, const uint32_t* __restrict__ key, const int* __restrict__ cs,
    const float4* __restrict__ pos, const float4* __restrict__ veleval,
    const float4* __restrict__ phys, const float4* __restrict__ dii,
    const float4* __restrict__ cst, const float* __restrict__ rho_adv,
    const float4* __restrict__ alpha, const float4* __restrict__ hist, float* __restrict__ X,
    float* __restrict__ pprev) {
  int i = xcd_index();
  if (i >= k.n) return;
  float* x = X + (size_t)i * MLP_NF;
  float4 pi = pos[i];
  uint32_t gc = key[i];
  pprev[i] = phys[i].z;
  if (type_of(pi) == SPHX_TYPE_BOUNDARY || (int)gc == k.gridtotal) {
    for (int q = 0; q < MLP_NF; ++q) x[q] = 0.f;
    return;
  }
  const float4 qi = phys[i];
  const float4 di = dii[i];
  const float p_i = fmaxf(0.f, qi.z);
  FeatAcc fa(p_i);
  struct Rec { float4 p, q; };
  for_each_listed_b(k, L, cs, gc, i, [&](int j) { return Rec{pos[j], phys[j]}; },
                    [&](int, const Rec& o) {
    const float4 pj = o.p;
    float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
    float dsq = dx * dx + dy * dy + dz * dz;
    float r, ir;
    dist_rinv(dsq, r, ir);
    fa.add(k, dx, dy, dz, r, ir, pj, o.q, p_i);
  });
  feat_write(k, mp, fa, qi, rho_adv[i], di, cst[i].x, veleval[i], alpha, hist, i, x);
}

// ------------------------------------------------- inference on MFMA --
// The MFMA operand layout of the current FP16 weights, built once per call
// of the network (one workgroup) and copied whole by every inference
// workgroup.
__global__ __launch_bounds__(256) void k_mlp_mfma_pack(const __half* __restrict__ w16,
                                                       MlpMfmaShared* __restrict__ out) {
  stage_mfma(*out, w16);
}

// The learned start of slot i (mlp_start_out); a boundary or out-of-grid
// slot only gets a zero drift correction.
__device__ __forceinline__ void mlp_infer_out(const KP& k, const MLPP& mp, int warm, int i,
                                              const float* y, const float4* __restrict__ pos,
                                              const uint32_t* __restrict__ key,
                                              const float4* __restrict__ phys,
                                              float* __restrict__ press_l,
                                              float4* __restrict__ ps,
                                              float4* __restrict__ dcorr) {
  float4 pi = pos[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY || (int)key[i] == k.gridtotal) {
    if (dcorr) dcorr[i] = make_float4(0.f, 0.f, 0.f, dcorr[i].w);
    return;
  }
  mlp_start_out(k, mp, warm, i, y, phys[i], press_l, ps, dcorr);
}

// The network on the feature rows X of every slot.  Workgroups of up to
// four waves of 64 particles walk the 256-particle tiles grid-stride, so
// each stages the weights in LDS once for several tiles.
__global__ __launch_bounds__(256, 2) void k_mlp_infer(
    KP k, MLPP mp, int warm, const float* __restrict__ X, const uint4* __restrict__ wmf,
    const float4* __restrict__ pos, const uint32_t* __restrict__ key,
    const float4* __restrict__ phys, float* __restrict__ press_l, float4* __restrict__ ps,
    float4* __restrict__ dcorr) {
  __shared__ MlpMfmaShared S;
  __shared__ float ys[4][64 * MLP_NO];
  load_mfma(S, wmf);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tiles = (k.n + (int)blockDim.x - 1) / (int)blockDim.x;
  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const int base = tile * (int)blockDim.x + 64 * wave;
    if (base >= k.n) continue;   // whole wave past the end
    mlp_forward_mfma(S, X + (size_t)base * MLP_NF, k.n - base, ys[wave]);
    // the wave's LDS writes above reach LDS before its reads below, and
    // its reads before the next tile's writes
    asm volatile("" ::: "memory");
    const int i = base + lane;
    if (i < k.n)
      mlp_infer_out(k, mp, warm, i, ys[wave] + lane * MLP_NO, pos, key, phys, press_l, ps, dcorr);
    asm volatile("" ::: "memory");
  }
}

// Workgroups of the inference: one per tile up to four per CU.
static int mlp_infer_blocks(int n, int threads) {
  return std::min(blocks_for(n, threads), 4 * 256);
}

// The network on stored feature rows (tests, evaluation): the same MFMA
// path as the solver's warm start.
__global__ __launch_bounds__(256, 2) void k_mlp_eval(const float* __restrict__ X, int n,
                                                     const uint4* __restrict__ wmf,
                                                     float* __restrict__ Y) {
  __shared__ MlpMfmaShared S;
  __shared__ float ys[4][64 * MLP_NO];
  load_mfma(S, wmf);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int base = blockIdx.x * (int)blockDim.x + 64 * wave;
  if (base >= n) return;
  mlp_forward_mfma(S, X + (size_t)base * MLP_NF, n - base, ys[wave]);
  asm volatile("" ::: "memory");
  const int i = base + lane;
  if (i >= n) return;
  for (int o = 0; o < MLP_NO; ++o) Y[(size_t)i * MLP_NO + o] = ys[wave][lane * MLP_NO + o];
}


// The weights in the MFMA operand layout.  Launched on the context stream
// right after every write of the FP16 weights (so the inference, eager or
// in a step graph, reads a layout that matches them in stream order).
static int mlp_pack_mfma(sphx_context* c) {
  auto& m = c->mlp;
  SPHX_TRY(m.wmf.alloc(MLP_MFMA_U4));
  k_mlp_mfma_pack<<<1, 256, 0, c->stream>>>(reinterpret_cast<const __half*>(m.w16.p),
                                            reinterpret_cast<MlpMfmaShared*>(m.wmf.p));
  SPHX_HIP(hipGetLastError());
  m.wmf_ok = true;
  return 0;
}

// Plain warm start p0 = warm * p(t-1) (fluid_system_kern.cu:899), used when
// a solve from the learned start does not converge.
__global__ void k_mlp_plain_start(KP k, const float4* __restrict__ pos,
                                  const uint32_t* __restrict__ key, float4* __restrict__ phys,
                                  const float* __restrict__ pprev, float* __restrict__ press_l,
                                  float4* __restrict__ ps, SceneCtl sc, float tol) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.n) return;
  const float4 pp = pos[i];
  if (type_of(pp) == SPHX_TYPE_BOUNDARY || (int)key[i] == k.gridtotal) return;
  // ensemble: only the scenes whose solve missed the tolerance restart
  if (sc.n && sc.err[scene_of(pp)] <= (double)tol) return;
  float4 q = phys[i];
  q.z = pprev[i];
  phys[i] = q;
  float p0 = k.warm * q.z;
  press_l[i] = p0;
  ps[i].w = q.x * p0 / (q.y * q.y);
}

// grad[e] = sum over the training waves' rows gpart[b][e] in a fixed order:
// a workgroup takes 64 entries, its four waves sum every fourth row each
// (b = wave, wave + 4, ...), and the four partial sums are added in wave
// order.
constexpr int GS_ENTRIES = 64;
__global__ __launch_bounds__(256) void k_mlp_gsum(const float* __restrict__ gpart, int blocks,
                                                  float* __restrict__ grad) {
  __shared__ float part[4][GS_ENTRIES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int e = blockIdx.x * GS_ENTRIES + lane;
  float s = 0.f;
  if (e < MLP_NP) {
#pragma unroll 8
    for (int b = wave; b < blocks; b += 4) s += gpart[(size_t)b * MLP_NP + e];
  }
  part[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && e < MLP_NP) grad[e] = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
}

// ----------------------------------------------------- training on MFMA --
// The loss gradient of one Adam step on v_mfma_f32_16x16x4f32: FP32
// operands and FP32 accumulation, so every product is the FP32 product of
// the VALU form and only the order of the sums differs.  A wave takes tiles
// of 16 samples.  Forward and backward keep the inference's transposed
// layout: units on the rows of a 16x16 tile, samples on its columns, lane
// (r = lane % 16, g = lane / 16) holding rows 4 g + i (i = 0..3) of column r.
// A layer's accumulators are then the next product's B operand for the K
// step (t, i) that takes units 16 t + 4 g + i, and the weights are read
// from LDS in that K order.  The weight gradients sum over samples: their
// operands need a unit on each lane and samples on the lane groups, so the
// wave stores H1, dH1, H2 and dH2 of its tile in LDS as [unit][sample] and
// reads them back transposed.  Each wave writes its partial gradient as one
// row of gpart (k_mlp_gsum adds the rows in a fixed order) and adds its
// loss to *loss (FP64, for reporting).
constexpr int TR_TS = 16;               // samples per tile
constexpr int TR_WAVES = 4;             // waves per workgroup
constexpr int TR_LD = TR_TS + 1;        // [unit][sample] row stride (LDS banks)
constexpr int TR_ROWS_MAX = 512;        // gpart rows (waves) per step
struct MlpTrainLds {
  float w[MLP_NP];                               // FP32 master weights
  float act[TR_WAVES][4][MLP_H * TR_LD];         // per wave: H1, dH1, H2, dH2
  float xs[TR_WAVES][TR_TS * 32];                // per wave: features [sample][feature]
  float dys[TR_WAVES][TR_TS * MLP_NO];           // per wave: dY [sample][output]
};

// Sample k of an epoch: valid[perm(k)], perm a bijection of [0, n) keyed
// per epoch: a four-round Feistel network on 2h bits (2^2h >= n, the round
// function a SplitMix64 finaliser of key, round and half), walked until it
// lands below n (it does: the walk stays on the cycle through k).  key 0:
// the identity (order[k] as stored).
struct SamplePerm {
  uint64_t key;
  uint32_t n;
  int h;
};
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint32_t perm_index(const SamplePerm& sp, uint32_t k) {
  if (sp.key == 0) return k;
  const uint32_t mask = (1u << sp.h) - 1u;
  uint32_t x = k;
  do {
    uint32_t L = x >> sp.h, R = x & mask;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t F = (uint32_t)mix64(sp.key + ((uint64_t)r << 40) + R) & mask;
      const uint32_t t = R;
      R = L ^ F;
      L = t;
    }
    x = (L << sp.h) | R;
  } while (x >= sp.n);
  return x;
}

__device__ __forceinline__ f4v mfma4(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// sum over the 16 lanes r of each lane group (fixed order)
__device__ __forceinline__ float group_sum16(float v) {
#pragma unroll
  for (int o = 8; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(TR_WAVES * 64) void k_mlp_train_mfma(
    const float* __restrict__ Xall, const float* __restrict__ Tall, const float* __restrict__ Mall,
    const int* __restrict__ perm, SamplePerm sp, int base, int nb, int nvalid_total,
    const float* __restrict__ w, float4 lw, int tpw, float* __restrict__ gpart,
    double* __restrict__ loss) {
  __shared__ MlpTrainLds S;
  for (int t = threadIdx.x; t < MLP_NP; t += blockDim.x) S.w[t] = w[t];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int row = blockIdx.x * TR_WAVES + wave;
  const int ntiles = (nb + TR_TS - 1) / TR_TS;
  const f4v z = {0.f, 0.f, 0.f, 0.f};
  const float lwv[4] = {lw.x, lw.y, lw.z, lw.w};
  f4v gw1[4][2], gw2[4][4], gw3[4];   // dW1^T (a x q), dW2 (a x b), dW3 (b x o) tiles
  float sb1[4][4], sb2[4][4], sb3[4];  // bias sums of the lane's sample column
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    gw1[t][0] = z; gw1[t][1] = z; gw3[t] = z; sb3[t] = 0.f;
#pragma unroll
    for (int u = 0; u < 4; ++u) { gw2[t][u] = z; sb1[t][u] = 0.f; sb2[t][u] = 0.f; }
  }
  double my_loss = 0.0;
  float* H1s = S.act[wave][0];
  float* D1s = S.act[wave][1];
  float* H2s = S.act[wave][2];
  float* D2s = S.act[wave][3];
  float* xs = S.xs[wave];
  float* dys = S.dys[wave];
  for (int k = 0; k < tpw; ++k) {
    const int tile = row * tpw + k;
    if (tile >= ntiles) break;   // the same for the whole wave
    const int s0 = tile * TR_TS;
    // lane r < 16 finds sample r's slot once; every lane loads 6 features
    int src_r = -1;
    if (lane < TR_TS && s0 + lane < nb) src_r = perm[perm_index(sp, base + s0 + lane)];
    for (int e = lane; e < TR_TS * MLP_NF; e += 64) {
      const int sl = e / MLP_NF, q = e - sl * MLP_NF;
      const int src = __shfl(src_r, sl, 64);
      xs[sl * 32 + q] = src >= 0 ? Xall[(size_t)src * MLP_NF + q] : 0.f;
    }
    // one wave's LDS accesses complete in order; keep the compiler's too
    asm volatile("" ::: "memory");
    // layer 1: H1^T = W1^T X^T (K = 24 features in steps of 4)
    f4v h1[4] = {z, z, z, z};
#pragma unroll
    for (int ks = 0; ks < MLP_NF / 4; ++ks) {
      const float b = xs[r * 32 + 4 * ks + g];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
        h1[mt] = mfma4(S.w[MLP_W1 + (4 * ks + g) * MLP_H + 16 * mt + r], b, h1[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) h1[mt][i] = fmaxf(h1[mt][i] + S.w[MLP_B1 + 16 * mt + 4 * g + i], 0.f);
    // layer 2: H2^T = W2^T H1^T (K step (t, i): units 16 t + 4 g + i)
    f4v h2[4] = {z, z, z, z};
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
          h2[mt] = mfma4(S.w[MLP_W2 + (16 * t + 4 * g + i) * MLP_H + 16 * mt + r], h1[t][i], h2[mt]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) h2[mt][i] = fmaxf(h2[mt][i] + S.w[MLP_B2 + 16 * mt + 4 * g + i], 0.f);
    // layer 3: Y^T = W3^T H2^T (rows 0..3: lanes of group 0 hold y_o = row i)
    f4v y = z;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        y = mfma4(r < MLP_NO ? S.w[MLP_W3 + (16 * t + 4 * g + i) * MLP_NO + r] : 0.f, h2[t][i], y);
    if (g == 0) {
      float dy[MLP_NO] = {0.f, 0.f, 0.f, 0.f};
      if (src_r >= 0) {   // lane r of group 0: sample r
        const int src = src_r;
        const float4 tv = reinterpret_cast<const float4*>(Tall)[src];
        const float msk = Mall[src];
        const float tt[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
        for (int o = 0; o < MLP_NO; ++o) {
          const float e = y[o] + S.w[MLP_B3 + o] - tt[o];
          my_loss += (double)msk * lwv[o] * e * e;
          dy[o] = 2.f * msk * lwv[o] * e / (float)nvalid_total;
        }
      }
#pragma unroll
      for (int o = 0; o < MLP_NO; ++o) {
        dys[r * MLP_NO + o] = dy[o];
        sb3[o] += dy[o];
      }
    }
    asm volatile("" ::: "memory");
    // backward, layer 3: dH2^T = W3 dY^T (K = the 4 outputs), masked by H2 > 0
    f4v d2[4];
    const float bdy = dys[r * MLP_NO + g];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      d2[mt] = mfma4(S.w[MLP_W3 + (16 * mt + r) * MLP_NO + g], bdy, z);
#pragma unroll
      for (int i = 0; i < 4; ++i) d2[mt][i] = h2[mt][i] > 0.f ? d2[mt][i] : 0.f;
    }
    // backward, layer 2: dH1^T = W2 dH2^T, masked by H1 > 0
    f4v g1[4] = {z, z, z, z};
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
          g1[mt] = mfma4(S.w[MLP_W2 + (16 * mt + r) * MLP_H + 16 * t + 4 * g + i], d2[t][i], g1[mt]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        g1[mt][i] = h1[mt][i] > 0.f ? g1[mt][i] : 0.f;
        sb1[mt][i] += g1[mt][i];
        sb2[mt][i] += d2[mt][i];
        const int u = (16 * mt + 4 * g + i) * TR_LD + r;
        H1s[u] = h1[mt][i];
        D1s[u] = g1[mt][i];
        H2s[u] = h2[mt][i];
        D2s[u] = d2[mt][i];
      }
    asm volatile("" ::: "memory");
    // weight gradients, K = the tile's 16 samples in steps of 4 (sample 4 ks + g)
#pragma unroll
    for (int ks = 0; ks < TR_TS / 4; ++ks) {
      const int sk = 4 * ks + g;
      float ah1[4], ad1[4], ah2[4], bd2[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int u = (16 * t + r) * TR_LD + sk;
        ah1[t] = H1s[u];
        ad1[t] = D1s[u];
        ah2[t] = H2s[u];
        bd2[t] = D2s[u];
      }
#pragma unroll
      for (int ma = 0; ma < 4; ++ma)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) gw2[ma][nt] = mfma4(ah1[ma], bd2[nt], gw2[ma][nt]);
#pragma unroll
      for (int nq = 0; nq < 2; ++nq) {
        const int q = 16 * nq + r;
        const float bx = q < MLP_NF ? xs[sk * 32 + q] : 0.f;
#pragma unroll
        for (int ma = 0; ma < 4; ++ma) gw1[ma][nq] = mfma4(ad1[ma], bx, gw1[ma][nq]);
      }
      const float bo = r < MLP_NO ? dys[sk * MLP_NO + r] : 0.f;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) gw3[mb] = mfma4(ah2[mb], bo, gw3[mb]);
    }
    asm volatile("" ::: "memory");   // this tile's reads before the next tile's writes
  }
  // this wave's row of partial gradients, in the weight order of k_mlp_adam
  float* out = gpart + (size_t)row * MLP_NP;
#pragma unroll
  for (int ma = 0; ma < 4; ++ma)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int a = 16 * ma + 4 * g + i;
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) out[MLP_W2 + a * MLP_H + 16 * nt + r] = gw2[ma][nt][i];
#pragma unroll
      for (int nq = 0; nq < 2; ++nq) {
        const int q = 16 * nq + r;
        if (q < MLP_NF) out[MLP_W1 + q * MLP_H + a] = gw1[ma][nq][i];
      }
      if (r < MLP_NO) out[MLP_W3 + a * MLP_NO + r] = gw3[ma][i];
      const float b1 = group_sum16(sb1[ma][i]), b2 = group_sum16(sb2[ma][i]);
      if (r == 0) {
        out[MLP_B1 + a] = b1;
        out[MLP_B2 + a] = b2;
      }
    }
#pragma unroll
  for (int o = 0; o < MLP_NO; ++o) {
    const float b3 = group_sum16(sb3[o]);   // lanes of group 0 hold the sums
    if (lane == 0) out[MLP_B3 + o] = b3;
  }
  for (int off = 32; off > 0; off >>= 1) my_loss += __shfl_down(my_loss, off, 64);
  if (lane == 0 && my_loss != 0.0) atomicAdd(loss, my_loss);
}

__global__ void k_mlp_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, __half* __restrict__ w16, int np,
                           float lr, float b1, float b2, float eps, float bc1, float bc2) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np) return;
  float gg = g[e];
  float mm = b1 * m[e] + (1.f - b1) * gg;
  float vv = b2 * v[e] + (1.f - b2) * gg * gg;
  m[e] = mm;
  v[e] = vv;
  float mh = mm / bc1, vh = vv / bc2;
  float nw = w[e] - lr * mh / (sqrtf(vh) + eps);
  w[e] = nw;
  w16[e] = __float2half(nw);
}

__global__ void k_f32_to_f16(const float* __restrict__ a, __half* __restrict__ b, int n) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) b[e] = __float2half(a[e]);
}

// Record one step's samples at offset `base` of the sample store: the
// features computed before the solve, the converged pressure increment and
// (multiphase) the change of the phase-1 drift when it is evaluated with the
// converged instead of the previous pressure.
// Sample t of the call is slot i = phase + t * stride (stride 1: every slot).
// One recorded sample: the feature row of slot i (as the a_ii pass left
// it), the targets and the mask into row t of the sample store.  `d` is
// slot i's pressure increment p - max(0, p(t-1)) (0 off the fluid).
__device__ __forceinline__ void record_sample(const MLPP& mp, int i, int t, bool ok, float d,
                                              const float* __restrict__ X,
                                              const float4* __restrict__ u_old,
                                              const float4* __restrict__ u_new,
                                              float* __restrict__ Xs, float* __restrict__ Ts,
                                              float* __restrict__ Ms) {
  const float4* xs = reinterpret_cast<const float4*>(X + (size_t)i * MLP_NF);
  float4* xd = reinterpret_cast<float4*>(Xs + (size_t)t * MLP_NF);
#pragma unroll
  for (int q = 0; q < MLP_NF / 4; ++q) xd[q] = xs[q];
  float t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
  if (ok) {
    t0 = d / mp.p_scale * MLP_TSCALE;
    if (u_old && u_new) {
      float s = MLP_TSCALE / mp.u_scale;
      float4 a = u_old[i], b = u_new[i];
      t1 = (b.x - a.x) * s; t2 = (b.y - a.y) * s; t3 = (b.z - a.z) * s;
    }
  }
  bool finite = t0 == t0 && t1 == t1 && t2 == t2 && t3 == t3 && fabsf(t0) < 1e30f;
  float m = ok && finite ? 1.f : 0.f;
  reinterpret_cast<float4*>(Ts)[t] = m > 0.f ? make_float4(t0, t1, t2, t3)
                                             : make_float4(0.f, 0.f, 0.f, 0.f);
  Ms[t] = m;
}

// hist.w = p - max(0, p(t-1)) after a solve (0 off the fluid) for every
// slot and, with `count` > 0, the sample of every slot i = phase + t stride
// (t < count) in the same coalesced sweep.  Fluid slots are those the
// solve's neighbour list marks valid (cnt: boundary and out-of-grid slots
// hold NBR_INVALID).
__global__ void k_mlp_hist(KP k, MLPP mp, const int* __restrict__ cnt,
                           const float4* __restrict__ phys,
                           const float* __restrict__ pprev, float4* __restrict__ hist,
                           int count, int stride, int phase, const float* __restrict__ X,
                           const float4* __restrict__ u_old, const float4* __restrict__ u_new,
                           float* __restrict__ Xs, float* __restrict__ Ts,
                           float* __restrict__ Ms) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.n) return;
  const bool ok = cnt[i] != NBR_INVALID;
  const float inc = phys[i].z - fmaxf(0.f, pprev[i]);
  float d = ok ? inc : 0.f;
  if (!(d == d)) d = 0.f;
  hist[i].w = d;
  const int r = i - phase;
  if (count == 0 || r < 0 || r % stride != 0) return;
  const int t = r / stride;
  if (t < count) record_sample(mp, i, t, ok, inc, X, u_old, u_new, Xs, Ts, Ms);
}

// p(t-1) when the features are not computed.
__global__ void k_mlp_pprev(int n, const float4* __restrict__ phys, float* __restrict__ pprev) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pprev[i] = phys[i].z;
}

__global__ void k_mlp_reorder(const uint32_t* __restrict__ skey,
                              const uint32_t* __restrict__ idsort, int n, uint32_t sentinel,
                              const float4* __restrict__ a, float4* __restrict__ b) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  b[s] = a[skey[s] != sentinel ? (int)idsort[s] : s];
}

// Uniform sample indices in [0, total) for `count` draws (SplitMix64).
__global__ void k_mlp_draw(int* __restrict__ order, int count, uint64_t total, uint64_t seed) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(t + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  order[t] = (int)(z % total);
}

MLPP make_mlpp(const sphx_context* c) {
  return MLPP{c->mlp.p_scale, c->mlp.u_scale, c->mlp.drift_blend};
}

int nbr_list(sphx_context* c, NbrList* out);
int mf_drift_into(sphx_context* c, const float4* dcorr, float4* out0, float4* out1,
                  float4* out2);

static int mlp_quiesce(sphx_context* c);

static int mlp_need(sphx_context* c, const char* who) {
  SPHX_CHECK(c, who, "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(c->mlp.ready, who, "network not initialised (call sphx_mlp_init)");
  if (c->mlp.train_inflight || c->mlp.train_launch) SPHX_TRY(mlp_quiesce(c));
  return 0;
}

// Gradient of the loss over samples order[base .. base+nb) into m.g; loss
// sum added to m.loss[0].
static int mlp_grad_batch(sphx_context* c, const float* X, const float* T, const float* M,
                          const int* order, int base, int nb, int nvalid, float4 lw,
                          SamplePerm sp = SamplePerm{0, 0, 0}) {
  auto& m = c->mlp;
  // tiles of 16 samples, at most TR_ROWS_MAX waves (gpart rows)
  const int tiles = (nb + TR_TS - 1) / TR_TS;
  const int tpw = (tiles + TR_ROWS_MAX - 1) / TR_ROWS_MAX;
  const int waves = (tiles + tpw - 1) / tpw;
  const int blocks = (waves + TR_WAVES - 1) / TR_WAVES;
  SPHX_TRY(m.gpart.alloc((size_t)TR_ROWS_MAX * MLP_NP));
  k_mlp_train_mfma<<<blocks, TR_WAVES * 64, 0, c->stream>>>(
      X, T, M, order, sp, base, nb, nvalid, m.w.p, lw, tpw, m.gpart.p, m.loss.p);
  SPHX_HIP(hipGetLastError());
  k_mlp_gsum<<<blocks_for(MLP_NP, GS_ENTRIES), 256, 0, c->stream>>>(m.gpart.p, blocks * TR_WAVES,
                                                                      m.g.p);
  SPHX_HIP(hipGetLastError());
  return 0;
}

static int mlp_refresh16(sphx_context* c) {
  auto& m = c->mlp;
  k_f32_to_f16<<<blocks_for(MLP_NP, 256), 256, 0, c->stream>>>(
      m.w.p, reinterpret_cast<__half*>(m.w16.p), MLP_NP);
  SPHX_HIP(hipGetLastError());
  return mlp_pack_mfma(c);
}

int iisph_solve_plain(sphx_context* c, int* iters, double* err);

// Online refresh: online_iters Adam steps, each on online_batch samples
// drawn uniformly (on the device) from the recorded ring.  The Adam steps
// write the FP16 weights the inference reads, or, asynchronous, run on the
// training stream after this step's samples are recorded and write staged
// FP16 weights (FP32 master and moments are touched by training only).
static int mlp_refresh_launch(sphx_context* c, __half* w16) {
  auto& m = c->mlp;
  const int nb = m.online_batch, steps = m.online_iters;
  const size_t draws = (size_t)nb * steps;
  SPHX_TRY(m.order.alloc(draws));
  m.online_seed += 0x632BE59BD9B4E019ull;
  k_mlp_draw<<<blocks_for((int)draws, 256), 256, 0, c->stream>>>(m.order.p, (int)draws,
                                                                  m.nsamples, m.online_seed);
  SPHX_HIP(hipGetLastError());
  float4 l4 = make_float4(m.online_lw[0], m.online_lw[1], m.online_lw[2], m.online_lw[3]);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int t = 0; t < steps; ++t) {
    SPHX_TRY(mlp_grad_batch(c, m.sX.p, m.sT.p, m.sM.p, m.order.p, t * nb, nb, nb, l4));
    ++m.adam_t;
    float bc1 = 1.f - std::pow(b1, (float)m.adam_t), bc2 = 1.f - std::pow(b2, (float)m.adam_t);
    k_mlp_adam<<<blocks_for(MLP_NP, 256), 256, 0, c->stream>>>(
        m.w.p, m.m.p, m.v.p, m.g.p, w16, MLP_NP, m.online_lr, b1, b2, eps, bc1, bc2);
    SPHX_HIP(hipGetLastError());
  }
  // the staged copy is packed when it is adopted
  if (w16 == reinterpret_cast<__half*>(m.w16.p)) SPHX_TRY(mlp_pack_mfma(c));
  return 0;
}

// Asynchronous refresh, at the end of the solve that recorded the samples:
// mark the main stream's position; the Adam steps are launched on the
// training stream by the next solve's start (launching them takes the host
// longer than the device needs to reach that point of its queue).
static int mlp_online_refresh(sphx_context* c) {
  auto& m = c->mlp;
  if (!m.online_async) return mlp_refresh_launch(c, reinterpret_cast<__half*>(m.w16.p));
  if (!m.tstream) {
    SPHX_HIP(hipStreamCreate(&m.tstream));
    SPHX_HIP(hipEventCreateWithFlags(&m.ev_main, hipEventDisableTiming));
    SPHX_HIP(hipEventCreateWithFlags(&m.ev_train, hipEventDisableTiming));
  }
  SPHX_TRY(m.w16_stage.alloc(MLP_NP + 1));
  SPHX_HIP(hipEventRecord(m.ev_main, c->stream));
  m.train_launch = true;
  return 0;
}

static int mlp_async_launch(sphx_context* c) {
  auto& m = c->mlp;
  if (!m.train_launch) return 0;
  m.train_launch = false;
  SPHX_HIP(hipStreamWaitEvent(m.tstream, m.ev_main, 0));
  hipStream_t main = c->stream;
  c->stream = m.tstream;
  int rc = mlp_refresh_launch(c, reinterpret_cast<__half*>(m.w16_stage.p));
  c->stream = main;
  SPHX_TRY(rc);
  SPHX_HIP(hipEventRecord(m.ev_train, m.tstream));
  m.train_pending = 1;
  m.train_inflight = true;
  return 0;
}

__global__ void k_u16_copy(const uint16_t* __restrict__ a, uint16_t* __restrict__ b, int n) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) b[e] = a[e];
}

// Adopt the staged weights now: the main stream waits for the training
// stream, then copies them in.
int mlp_async_settle(sphx_context* c) {
  auto& m = c->mlp;
  SPHX_TRY(mlp_async_launch(c));
  if (!m.train_inflight) return 0;
  SPHX_HIP(hipStreamWaitEvent(c->stream, m.ev_train, 0));
  k_u16_copy<<<blocks_for(MLP_NP, 256), 256, 0, c->stream>>>(m.w16_stage.p, m.w16.p, MLP_NP);
  SPHX_HIP(hipGetLastError());
  SPHX_TRY(mlp_pack_mfma(c));
  m.train_pending = 0;
  m.train_inflight = false;
  return 0;
}

// Host-side calls that read or replace the weights, moments or samples
// first finish an asynchronous refresh.
static int mlp_quiesce(sphx_context* c) {
  SPHX_TRY(mlp_async_settle(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// Before the a_ii pass evaluates the network: adopt the step's due
// refresh (which mlp_begin then skips) and pack the weights.
int mlp_prepare_start(sphx_context* c) {
  auto& m = c->mlp;
  if (!m.settled) {
    if (m.train_pending > 0 && --m.train_pending == 0) SPHX_TRY(mlp_async_settle(c));
    m.settled = true;
  }
  if (!m.wmf_ok) SPHX_TRY(mlp_pack_mfma(c));
  return 0;
}

// The pressure solve of a step with the learned start, the recorder and
// the fallback.  Called after the a_ii pass.
// Before a solve: features (or just p(t-1)) and the learned start.
int mlp_begin(sphx_context* c) {
  auto& m = c->mlp;
  if (!m.ready || c->num == 0) return 0;
  if (!m.settled && m.train_pending > 0 && --m.train_pending == 0) SPHX_TRY(mlp_async_settle(c));
  m.settled = false;
  SPHX_TRY(mlp_async_launch(c));
  KP k = make_kp(c);
  MLPP mp = make_mlpp(c);
  const bool mf = c->nphase > 0;
  const bool record = m.record || m.online_every > 0;
  SPHX_TRY(m.X.alloc((size_t)c->cap * MLP_NF));
  SPHX_TRY(m.pprev.alloc((size_t)c->cap));
  SPHX_TRY(m.pkeep.alloc((size_t)c->cap));
  float4* hist = m.hist[c->cur].p;
  if ((m.warm || record) && !m.feat_ok) {
    NbrList L;
    SPHX_TRY(nbr_list(c, &L));
    timer_begin(c, "mlp_features");
    k_mlp_features<<<blocks_for(k.n, c->tune[T_MLP_FEAT]), c->tune[T_MLP_FEAT], 0, c->stream>>>(
        k, L, mp, c->key.p, c->cell_start.p, c->POS(), c->VELEVAL(), c->PHYS(), c->dii.p,
        c->cst.p, c->rho_adv.p, mf ? c->mf_alpha[c->cur].p : nullptr, hist, m.X.p, m.pprev.p);
    timer_end(c);
  } else if (!(m.warm || record)) {
    k_mlp_pprev<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(k.n, c->PHYS(), m.pprev.p);
  }
  m.feat_ok = false;   // used once
  SPHX_HIP(hipGetLastError());
  float4* dcorr = m.drift && mf ? hist : nullptr;
  if ((m.warm || dcorr) && !m.start_ok) {
    if (!m.wmf_ok) SPHX_TRY(mlp_pack_mfma(c));
    timer_begin(c, "mlp_infer");
    k_mlp_infer<<<mlp_infer_blocks(k.n, c->tune[T_MLP_INFER]), c->tune[T_MLP_INFER], 0, c->stream>>>(
        k, mp, m.warm, m.X.p, m.wmf.p, c->POS(), c->key.p, c->PHYS(), c->press_l.p, c->ps.p,
        dcorr);
    timer_end(c);
    SPHX_HIP(hipGetLastError());
  }
  m.start_ok = false;   // used once
  // the solve's first iteration is followed by the guard (single scene)
  c->mlp.guard_armed = m.warm && m.guard && c->nscenes == 0;
  return 0;
}

// After a solve from the learned start: the plain solver judges it; a solve
// that missed the tolerance is redone from the plain start (per scene in an
// ensemble), and its iterations are added.
int mlp_fallback(sphx_context* c, int* iters, double* err) {
  auto& m = c->mlp;
  if (!m.ready || !m.warm || c->num == 0) return 0;
  ++m.learned_solves;
  m.last_fallback = 0;
  if (c->coop_fallbacks) {   // the persistent solve already fell back on the device
    m.fallbacks += c->coop_fallbacks;
    m.last_fallback = 1;
    c->coop_fallbacks = 0;
  }
  // the done flag is also set at max_iters: judge by the error itself
  double e = *err;
  if (e == e && e <= (double)c->O.tol) return 0;
  ++m.fallbacks;
  m.last_fallback = 1;
  KP k = make_kp(c);
  k_mlp_plain_start<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, c->POS(), c->key.p, c->PHYS(), m.pprev.p, c->press_l.p, c->ps.p, c->scenes(), c->O.tol);
  SPHX_HIP(hipGetLastError());
  int it2 = 0;
  SPHX_TRY(iisph_solve_plain(c, &it2, err));
  *iters += it2;
  return 0;
}

// After the final solve: the per-particle pressure increment (feature 20),
// then recording and the online refresh when enabled (record_ok = 0 inside
// a captured graph).
int mlp_after(sphx_context* c, int record_ok) {
  auto& m = c->mlp;
  if (!m.ready || c->num == 0) return 0;
  KP k = make_kp(c);
  MLPP mp = make_mlpp(c);
  size_t n = (size_t)c->num;
  const bool mf = c->nphase > 0;
  const bool online = m.online_every > 0;
  const bool record = record_ok && (m.record || online);
  float4* hist = m.hist[c->cur].p;
  // Recording with a stride keeps every stride-th slot of a solve (the
  // first slot rotating with the solve count); online, in a ring of as many
  // solves as the store holds full solves.
  const int stride = m.record_stride;
  const int phase = stride > 1 ? (int)(m.solves % stride) : 0;
  const size_t nrec = stride > 1 ? (n - (size_t)phase + stride - 1) / stride : n;
  const bool ring = online && !m.record;
  bool recorded = false;
  const size_t cap = (ring && stride > 1)
                         ? std::min(m.max_samples, (m.max_samples / n) * ((n + stride - 1) / stride))
                         : m.max_samples;
  if (record && (ring ? m.max_samples >= n : m.max_samples >= nrec)) {
    size_t base = m.wpos;
    if (base + nrec > cap) base = online ? 0 : cap;   // ring or full
    if (base + nrec <= cap) {
      const float4* u_old = nullptr;
      const float4* u_new = nullptr;
      // drift targets cost one more drift pass: only when they are used
      const bool drift_w = m.online_lw[1] != 0.f || m.online_lw[2] != 0.f || m.online_lw[3] != 0.f;
      if (mf && c->nphase >= 2 && !m.drift && (m.record || drift_w)) {
        SPHX_TRY(m.unew.alloc((size_t)c->cap * 3));
        float4* u = m.unew.p;
        SPHX_TRY(mf_drift_into(c, nullptr, u, u + c->cap, u + 2 * (size_t)c->cap));
        u_old = c->mf_drift[1].p;
        u_new = u + c->cap;
      }
      if (m.train_inflight) SPHX_HIP(hipStreamWaitEvent(c->stream, m.ev_train, 0));
      // the increments and the samples in one sweep over the slots
      timer_begin(c, "mlp_record");
      k_mlp_hist<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
          k, mp, c->nbr_cnt.p, c->PHYS(), m.pprev.p, hist, (int)nrec, stride, phase, m.X.p,
          u_old, u_new, m.sX.p + base * MLP_NF, m.sT.p + base * MLP_NO, m.sM.p + base);
      timer_end(c);
      SPHX_HIP(hipGetLastError());
      m.wpos = base + nrec;
      m.nsamples = std::min(std::max(m.nsamples, m.wpos), cap);
      recorded = true;
    }
  }
  if (!recorded) {
    k_mlp_hist<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
        k, mp, c->nbr_cnt.p, c->PHYS(), m.pprev.p, hist, 0, 1, 0, nullptr, nullptr, nullptr,
        nullptr, nullptr, nullptr);
    SPHX_HIP(hipGetLastError());
  }
  if (record_ok) {
    ++m.solves;
    if (online && m.solves % m.online_every == 0 && m.nsamples > 0) {
      timer_begin(c, "mlp_train");
      SPHX_TRY(mlp_online_refresh(c));
      timer_end(c);
    }
  }
  return 0;
}

// The pressure solve of a step with the learned start, the recorder and
// the fallback.  Called after the a_ii pass.
int mlp_solve(sphx_context* c, int* iters, double* err) {
  auto& m = c->mlp;
  if (!m.ready || c->num == 0) return iisph_solve_plain(c, iters, err);
  SPHX_TRY(mlp_begin(c));
  int it = 0;
  double e = 0.0;
  SPHX_TRY(iisph_solve_plain(c, &it, &e));
  SPHX_TRY(mlp_fallback(c, &it, &e));
  SPHX_TRY(mlp_after(c, 1));
  if (iters) *iters = it;
  if (err) *err = e;
  return 0;
}

int mlp_reorder(sphx_context* c, int a, int b) {
  if (!c->mlp.ready || c->num == 0) return 0;
  k_mlp_reorder<<<blocks_for(c->num, 256), 256, 0, c->stream>>>(
      c->key.p, c->idsort.p, c->num, (uint32_t)c->P.gridtotal, c->mlp.hist[a].p,
      c->mlp.hist[b].p);
  SPHX_HIP(hipGetLastError());
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

int sphx_mlp_nparam(void) { return MLP_NP; }

int sphx_mlp_guard(sphx_context* c, float factor) {
  SPHX_CHECK(c, "sphx_mlp_guard", "null context");
  SPHX_CHECK(factor >= 0.f, "sphx_mlp_guard", "factor must be >= 0 (0: off)");
  c->mlp.guard = factor > 0.f;
  c->mlp.guard_factor = factor > 0.f ? factor : 1.f;
  return 0;
}

int sphx_mlp_guard_stats(sphx_context* c, int64_t* probes, int64_t* plain_picks) {
  SPHX_CHECK(c && probes && plain_picks, "sphx_mlp_guard_stats", "null argument");
  *probes = *plain_picks = 0;
  if (!c->mlp.gstate.p) return 0;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  int g[4];
  SPHX_HIP(hipMemcpy(g, c->mlp.gstate.p, sizeof(g), hipMemcpyDeviceToHost));
  *probes = g[2];
  *plain_picks = g[3];
  return 0;
}

int sphx_mlp_init(sphx_context* c, uint64_t seed, float p_scale, float u_scale) {
  SPHX_CHECK(c, "sphx_mlp_init", "null context");
  SPHX_CHECK(p_scale > 0.f && u_scale > 0.f, "sphx_mlp_init", "scales must be positive");
  SPHX_TRY(ensure_device(c));
  if (c->mlp.train_inflight || c->mlp.train_launch) SPHX_TRY(mlp_quiesce(c));
  auto& m = c->mlp;
  SPHX_TRY(m.w.alloc(MLP_NP));
  SPHX_TRY(m.m.alloc(MLP_NP));
  SPHX_TRY(m.v.alloc(MLP_NP));
  SPHX_TRY(m.g.alloc(MLP_NP));
  SPHX_TRY(m.w16.alloc(MLP_NP + 1));
  SPHX_TRY(m.loss.alloc(4));
  SPHX_TRY(m.gstate.alloc(4));
  SPHX_HIP(hipMemset(m.gstate.p, 0, 4 * sizeof(int)));
  for (int b = 0; b < 2; ++b) {
    SPHX_TRY(m.hist[b].alloc((size_t)std::max(c->cap, 1)));
    SPHX_HIP(hipMemset(m.hist[b].p, 0, (size_t)std::max(c->cap, 1) * sizeof(float4)));
  }
  // He initialisation of the weights, zero biases; the output layer starts
  // at zero so an untrained network predicts "no change".
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<float> w(MLP_NP, 0.f);
  for (int e = MLP_W1; e < MLP_B1; ++e) w[e] = (float)(nd(rng) * std::sqrt(2.0 / MLP_NF));
  for (int e = MLP_W2; e < MLP_B2; ++e) w[e] = (float)(nd(rng) * std::sqrt(2.0 / MLP_H));
  SPHX_HIP(hipMemcpy(m.w.p, w.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemset(m.m.p, 0, MLP_NP * sizeof(float)));
  SPHX_HIP(hipMemset(m.v.p, 0, MLP_NP * sizeof(float)));
  m.adam_t = 0;
  m.p_scale = p_scale;
  m.u_scale = u_scale;
  m.ready = true;
  SPHX_TRY(mlp_refresh16(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int sphx_mlp_set_weights(sphx_context* c, const float* w, int n) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_set_weights"));
  SPHX_CHECK(w && n == MLP_NP, "sphx_mlp_set_weights", "expects 6020 weights");
  SPHX_HIP(hipStreamSynchronize(c->stream));
  SPHX_HIP(hipMemcpy(c->mlp.w.p, w, MLP_NP * sizeof(float), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemset(c->mlp.m.p, 0, MLP_NP * sizeof(float)));
  SPHX_HIP(hipMemset(c->mlp.v.p, 0, MLP_NP * sizeof(float)));
  c->mlp.adam_t = 0;
  SPHX_TRY(mlp_refresh16(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int sphx_mlp_get_weights(sphx_context* c, float* w, int n) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_get_weights"));
  SPHX_CHECK(w && n == MLP_NP, "sphx_mlp_get_weights", "expects 6020 weights");
  SPHX_HIP(hipStreamSynchronize(c->stream));
  SPHX_HIP(hipMemcpy(w, c->mlp.w.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// warm: learned pressure warm start; drift: learned drift correction;
// record: store (features, targets) of every solve.  Drift targets are
// recorded only while the drift correction is off (they are the change of
// the uncorrected drift); with it on they are stored as zero.
int sphx_mlp_mode(sphx_context* c, int warm, int drift, int record) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_mode"));
  auto& m = c->mlp;
  m.warm = warm != 0;
  m.drift = drift != 0;
  m.record = record != 0;
  return 0;
}

// Online refresh: every `every` solves (0 = off), `iters` Adam steps with
// learning rate lr on batches of `batch` samples drawn uniformly from the
// sample store, which then records every solve as a ring.
int sphx_mlp_online(sphx_context* c, int every, int iters, int batch, float lr, const float* lw) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_online"));
  SPHX_CHECK(every >= 0 && iters >= 0 && lr >= 0.f, "sphx_mlp_online", "bad argument");
  SPHX_CHECK(every == 0 || (batch > 0 && lw), "sphx_mlp_online", "bad batch or weights");
  SPHX_CHECK(every == 0 || (size_t)batch * (size_t)iters < (size_t)INT32_MAX, "sphx_mlp_online",
             "batch * iters too large");
  SPHX_CHECK(every == 0 || c->mlp.max_samples >= (size_t)c->num, "sphx_mlp_online",
             "reserve room for at least one step of samples first");
  SPHX_CHECK(every == 0 || !c->coop, "sphx_mlp_online",
             "not available together with the persistent solve");
  auto& m = c->mlp;
  m.online_every = every;
  m.online_iters = iters;
  m.online_batch = batch;
  m.online_lr = lr;
  if (lw)
    for (int o = 0; o < 4; ++o) m.online_lw[o] = lw[o];
  return 0;
}

// Asynchronous online refresh (on != 0): a refresh's Adam steps run on a
// second stream while the next solve proceeds with the current weights;
// the solve after that adopts the new ones.  Off: the refresh runs in the
// step's own stream and the next solve uses its result.
int sphx_mlp_online_async(sphx_context* c, int on) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_online_async"));
  c->mlp.online_async = on != 0;
  return 0;
}

// Online recording keeps every `stride`-th particle slot of a solve
// (stride 1: all).  The ring then holds as many solves as the sample store
// holds full solves, so its window in solves does not change.
int sphx_mlp_record_stride(sphx_context* c, int stride) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_record_stride"));
  SPHX_CHECK(stride >= 1, "sphx_mlp_record_stride", "stride must be >= 1");
  c->mlp.record_stride = stride;
  return 0;
}

int sphx_mlp_set_drift_blend(sphx_context* c, float blend) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_set_drift_blend"));
  c->mlp.drift_blend = blend;
  return 0;
}

// Reserve room for `max_samples` recorded samples (one per particle slot
// and step; boundary and out-of-grid slots are stored masked).
int sphx_mlp_reserve(sphx_context* c, int64_t max_samples) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_reserve"));
  SPHX_CHECK(max_samples >= 0, "sphx_mlp_reserve", "negative size");
  auto& m = c->mlp;
  size_t s = (size_t)max_samples;
  SPHX_TRY(m.sX.alloc(std::max<size_t>(s, 1) * MLP_NF));
  SPHX_TRY(m.sT.alloc(std::max<size_t>(s, 1) * MLP_NO));
  SPHX_TRY(m.sM.alloc(std::max<size_t>(s, 1)));
  m.max_samples = s;
  m.nsamples = 0;
  m.wpos = 0;
  return 0;
}

int sphx_mlp_clear_samples(sphx_context* c) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_clear_samples"));
  c->mlp.nsamples = 0;
  c->mlp.wpos = 0;
  return 0;
}

// Recorded sample count and the number of them that are valid (unmasked).
int sphx_mlp_samples(sphx_context* c, int64_t* total, int64_t* valid) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_samples"));
  auto& m = c->mlp;
  if (total) *total = (int64_t)m.nsamples;
  if (valid) {
    SPHX_HIP(hipStreamSynchronize(c->stream));
    std::vector<float> msk(m.nsamples);
    if (m.nsamples)
      SPHX_HIP(hipMemcpy(msk.data(), m.sM.p, m.nsamples * sizeof(float), hipMemcpyDeviceToHost));
    int64_t v = 0;
    for (float x : msk) v += x > 0.f;
    *valid = v;
  }
  return 0;
}

// Copy recorded samples [start, start+count) to the host (any pointer may
// be NULL): X [count][24], T [count][4], M [count].
int sphx_mlp_download_samples(sphx_context* c, int64_t start, int64_t count, float* X, float* T,
                              float* M) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_download_samples"));
  auto& m = c->mlp;
  SPHX_CHECK(start >= 0 && count >= 0 && (size_t)(start + count) <= m.nsamples,
             "sphx_mlp_download_samples", "range exceeds the recorded samples");
  SPHX_HIP(hipStreamSynchronize(c->stream));
  if (count == 0) return 0;
  if (X) SPHX_HIP(hipMemcpy(X, m.sX.p + start * MLP_NF, count * MLP_NF * sizeof(float), hipMemcpyDeviceToHost));
  if (T) SPHX_HIP(hipMemcpy(T, m.sT.p + start * MLP_NO, count * MLP_NO * sizeof(float), hipMemcpyDeviceToHost));
  if (M) SPHX_HIP(hipMemcpy(M, m.sM.p + start, count * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// Features of the current state (after the a_ii pass) into X [num][24],
// and the previous pressures into pprev [num] (either may be NULL).
int sphx_mlp_features(sphx_context* c, float* X, float* pprev, float* dp_last) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_features"));
  SPHX_CHECK(c->sorted_valid, "sphx_mlp_features", "grid not built");
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  auto& m = c->mlp;
  SPHX_TRY(m.X.alloc((size_t)c->cap * MLP_NF));
  SPHX_TRY(m.pprev.alloc((size_t)c->cap));
  k_mlp_features<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, L, make_mlpp(c), c->key.p, c->cell_start.p, c->POS(), c->VELEVAL(), c->PHYS(), c->dii.p,
      c->cst.p, c->rho_adv.p, c->nphase > 0 ? c->mf_alpha[c->cur].p : nullptr, m.hist[c->cur].p, m.X.p,
      m.pprev.p);
  SPHX_HIP(hipGetLastError());
  SPHX_HIP(hipStreamSynchronize(c->stream));
  if (dp_last) {
    std::vector<float4> h(k.n);
    SPHX_HIP(hipMemcpy(h.data(), m.hist[c->cur].p, (size_t)k.n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < k.n; ++i) dp_last[i] = h[i].w;
  }
  if (X) SPHX_HIP(hipMemcpy(X, m.X.p, (size_t)k.n * MLP_NF * sizeof(float), hipMemcpyDeviceToHost));
  if (pprev) SPHX_HIP(hipMemcpy(pprev, m.pprev.p, (size_t)k.n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// The feature buffer as the last pass left it (the a_ii pass computes the
// features when the network is in use; sphx_mlp_features recomputes them).
int sphx_mlp_download_features(sphx_context* c, float* X, float* pprev) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_download_features"));
  SPHX_CHECK(c->mlp.X.p && c->mlp.pprev.p, "sphx_mlp_download_features", "no features yet");
  SPHX_HIP(hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->num;
  if (X) SPHX_HIP(hipMemcpy(X, c->mlp.X.p, n * MLP_NF * sizeof(float), hipMemcpyDeviceToHost));
  if (pprev) SPHX_HIP(hipMemcpy(pprev, c->mlp.pprev.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// The network (FP16 weights, FP32 accumulation) on host features.
int sphx_mlp_eval(sphx_context* c, const float* X, int n, float* Y) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_eval"));
  SPHX_CHECK(X && Y && n >= 0, "sphx_mlp_eval", "null argument");
  if (n == 0) return 0;
  DevBuf<float> dx, dy;
  SPHX_TRY(dx.alloc((size_t)n * MLP_NF));
  SPHX_TRY(dy.alloc((size_t)n * MLP_NO));
  SPHX_HIP(hipMemcpy(dx.p, X, (size_t)n * MLP_NF * sizeof(float), hipMemcpyHostToDevice));
  SPHX_TRY(mlp_pack_mfma(c));
  k_mlp_eval<<<blocks_for(n, 256), 256, 0, c->stream>>>(dx.p, n, c->mlp.wmf.p, dy.p);
  SPHX_HIP(hipGetLastError());
  SPHX_HIP(hipStreamSynchronize(c->stream));
  SPHX_HIP(hipMemcpy(Y, dy.p, (size_t)n * MLP_NO * sizeof(float), hipMemcpyDeviceToHost));
  dx.release();
  dy.release();
  return 0;
}

// Loss (1/n_valid) sum_s M_s sum_o lw_o (y_so - T_so)^2 and its gradient
// with respect to the FP32 master weights, by the training kernels, for
// host samples X [n][24], T [n][4], M [n] (M may be NULL: all valid).
int sphx_mlp_grad(sphx_context* c, const float* X, const float* T, const float* M, int n,
                  const float* lw, float* grad, double* loss) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_grad"));
  SPHX_CHECK(X && T && lw && n > 0, "sphx_mlp_grad", "bad argument");
  auto& m = c->mlp;
  std::vector<float> msk(n, 1.f);
  if (M) std::memcpy(msk.data(), M, n * sizeof(float));
  int nvalid = 0;
  for (float v : msk) nvalid += v > 0.f;
  std::vector<int> ord(n);
  for (int s = 0; s < n; ++s) ord[s] = s;
  DevBuf<float> dx, dt, dm;
  DevBuf<int> dord;
  SPHX_TRY(dx.alloc((size_t)n * MLP_NF));
  SPHX_TRY(dt.alloc((size_t)n * MLP_NO));
  SPHX_TRY(dm.alloc(n));
  SPHX_TRY(dord.alloc(n));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  SPHX_HIP(hipMemcpy(dx.p, X, (size_t)n * MLP_NF * sizeof(float), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(dt.p, T, (size_t)n * MLP_NO * sizeof(float), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(dm.p, msk.data(), n * sizeof(float), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(dord.p, ord.data(), n * sizeof(int), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemsetAsync(m.loss.p, 0, sizeof(double), c->stream));
  float4 l4 = make_float4(lw[0], lw[1], lw[2], lw[3]);
  SPHX_TRY(mlp_grad_batch(c, dx.p, dt.p, dm.p, dord.p, 0, n, std::max(nvalid, 1), l4));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  if (grad) SPHX_HIP(hipMemcpy(grad, m.g.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) {
    double l;
    SPHX_HIP(hipMemcpy(&l, m.loss.p, sizeof(double), hipMemcpyDeviceToHost));
    *loss = l / std::max(nvalid, 1);
  }
  return 0;
}

// Adam on the recorded samples: `epochs` passes in shuffled mini-batches
// of `batch` valid samples.  loss_out[e] (may be NULL) receives the mean
// loss of epoch e.  The FP16 inference copy is refreshed after each step.
int sphx_mlp_train(sphx_context* c, int epochs, int batch, float lr, const float* lw,
                   uint64_t seed, double* loss_out) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_train"));
  SPHX_CHECK(epochs >= 0 && batch > 0 && lr > 0.f && lw, "sphx_mlp_train", "bad argument");
  auto& m = c->mlp;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  std::vector<float> msk(m.nsamples);
  if (m.nsamples)
    SPHX_HIP(hipMemcpy(msk.data(), m.sM.p, m.nsamples * sizeof(float), hipMemcpyDeviceToHost));
  std::vector<int> valid;
  valid.reserve(m.nsamples);
  SPHX_CHECK(m.nsamples < (size_t)INT32_MAX, "sphx_mlp_train", "too many samples");
  for (size_t s = 0; s < m.nsamples; ++s)
    if (msk[s] > 0.f) valid.push_back((int)s);
  SPHX_CHECK(!valid.empty(), "sphx_mlp_train", "no recorded samples");
  const int nv = (int)valid.size();
  SPHX_TRY(m.order.alloc(nv));
  SPHX_HIP(hipMemcpyAsync(m.order.p, valid.data(), (size_t)nv * sizeof(int),
                          hipMemcpyHostToDevice, c->stream));
  // each epoch visits the valid samples in the order of a keyed bijection
  // of [0, nv), evaluated on the device
  int h = 1;
  while ((uint64_t)1 << (2 * h) < (uint64_t)nv) ++h;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float4 l4 = make_float4(lw[0], lw[1], lw[2], lw[3]);
  for (int ep = 0; ep < epochs; ++ep) {
    const SamplePerm sp{rng() | 1ull, (uint32_t)nv, h};
    SPHX_HIP(hipMemsetAsync(m.loss.p, 0, sizeof(double), c->stream));
    for (int base = 0; base < nv; base += batch) {
      int nb = std::min(batch, nv - base);
      SPHX_TRY(mlp_grad_batch(c, m.sX.p, m.sT.p, m.sM.p, m.order.p, base, nb, nb, l4, sp));
      ++m.adam_t;
      float bc1 = 1.f - std::pow(b1, (float)m.adam_t), bc2 = 1.f - std::pow(b2, (float)m.adam_t);
      k_mlp_adam<<<blocks_for(MLP_NP, 256), 256, 0, c->stream>>>(
          m.w.p, m.m.p, m.v.p, m.g.p, reinterpret_cast<__half*>(m.w16.p), MLP_NP, lr, b1, b2,
          eps, bc1, bc2);
      SPHX_HIP(hipGetLastError());
    }
    if (loss_out) {
      double l;
      SPHX_HIP(hipMemcpy(&l, m.loss.p, sizeof(double), hipMemcpyDeviceToHost));
      loss_out[ep] = l / nv;
    }
  }
  return mlp_pack_mfma(c);
}

// Learned solves so far, how many fell back to the plain start, and
// whether the last one did.
int sphx_mlp_stats(sphx_context* c, int64_t* learned, int64_t* fallbacks, int* last_fallback) {
  SPHX_TRY(mlp_need(c, "sphx_mlp_stats"));
  if (learned) *learned = c->mlp.learned_solves;
  if (fallbacks) *fallbacks = c->mlp.fallbacks;
  if (last_fallback) *last_fallback = c->mlp.last_fallback;
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------ checkpoint --
namespace sphx {

// The network's state: settings, the FP32 weights with their Adam moments
// and step, the solve counter, and the per-particle history (drift
// correction, last pressure increment) in slot order.  The online sample
// ring is not part of it.
struct MlpHead {
  int32_t warm, drift, record, adam_t;
  int32_t online_every, online_iters, online_batch, record_stride;
  int32_t online_async, guard, np, n;
  int64_t solves, seed;
  float p_scale, u_scale, drift_blend, online_lr, online_lw[4];
};

int mlp_state_get(sphx_context* c, std::vector<char>& out) {
  out.clear();
  if (!c || !c->mlp.ready) return 0;
  auto& m = c->mlp;
  SPHX_TRY(mlp_quiesce(c));
  MlpHead h{};
  h.warm = m.warm; h.drift = m.drift; h.record = m.record; h.adam_t = m.adam_t;
  h.online_every = m.online_every; h.online_iters = m.online_iters;
  h.online_batch = m.online_batch; h.record_stride = m.record_stride;
  h.online_async = m.online_async; h.guard = m.guard; h.np = MLP_NP; h.n = c->num;
  h.solves = m.solves; h.seed = (int64_t)m.online_seed;
  h.p_scale = m.p_scale; h.u_scale = m.u_scale; h.drift_blend = m.drift_blend;
  h.online_lr = m.online_lr;
  for (int o = 0; o < 4; ++o) h.online_lw[o] = m.online_lw[o];
  const size_t wb = (size_t)MLP_NP * sizeof(float), hb = (size_t)c->num * sizeof(float4);
  out.resize(sizeof(h) + 3 * wb + hb);
  char* p = out.data();
  std::memcpy(p, &h, sizeof(h));
  SPHX_HIP(hipMemcpy(p + sizeof(h), m.w.p, wb, hipMemcpyDeviceToHost));
  SPHX_HIP(hipMemcpy(p + sizeof(h) + wb, m.m.p, wb, hipMemcpyDeviceToHost));
  SPHX_HIP(hipMemcpy(p + sizeof(h) + 2 * wb, m.v.p, wb, hipMemcpyDeviceToHost));
  if (c->num > 0)
    SPHX_HIP(hipMemcpy(p + sizeof(h) + 3 * wb, m.hist[c->cur].p, hb, hipMemcpyDeviceToHost));
  return 0;
}

int mlp_state_set(sphx_context* c, const char* p, size_t bytes) {
  SPHX_CHECK(c && p && bytes >= sizeof(MlpHead), "sphx_checkpoint_restore",
             "truncated network section");
  MlpHead h;
  std::memcpy(&h, p, sizeof(h));
  const size_t wb = (size_t)MLP_NP * sizeof(float), hb = (size_t)c->num * sizeof(float4);
  SPHX_CHECK(h.np == MLP_NP && h.n == c->num && bytes == sizeof(h) + 3 * wb + hb,
             "sphx_checkpoint_restore", "network section does not match this build or the particles");
  SPHX_CHECK(h.online_every == 0 || !c->coop, "sphx_checkpoint_restore",
             "the network's online refresh is not available together with the persistent solve");
  SPHX_TRY(sphx_mlp_init(c, 1, h.p_scale, h.u_scale));
  auto& m = c->mlp;
  SPHX_HIP(hipMemcpy(m.w.p, p + sizeof(h), wb, hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(m.m.p, p + sizeof(h) + wb, wb, hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(m.v.p, p + sizeof(h) + 2 * wb, wb, hipMemcpyHostToDevice));
  if (c->num > 0)
    SPHX_HIP(hipMemcpy(m.hist[c->cur].p, p + sizeof(h) + 3 * wb, hb, hipMemcpyHostToDevice));
  m.warm = h.warm; m.drift = h.drift; m.record = h.record; m.adam_t = h.adam_t;
  m.online_every = h.online_every; m.online_iters = h.online_iters;
  m.online_batch = h.online_batch; m.record_stride = h.record_stride;
  m.online_async = h.online_async; m.guard = h.guard;
  m.solves = (int)h.solves; m.online_seed = (uint64_t)h.seed;
  m.drift_blend = h.drift_blend; m.online_lr = h.online_lr;
  for (int o = 0; o < 4; ++o) m.online_lw[o] = h.online_lw[o];
  SPHX_TRY(mlp_refresh16(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace sphx
