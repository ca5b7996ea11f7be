// Stokesian mobility of spheres (SPH_Solver/stokesian/Mobility.cu).
//
// The polydisperse Rotne-Prager-Yamakawa mobility is applied to K
// right-hand sides, U = M F (M is 3N x 3N, row 3p + a), without storing M:
// each workgroup generates 64 x 32-sphere tiles of the 3x3 pair blocks on
// the VALU, splits them into bf16 pieces in LDS and multiplies them against
// the bf16 pieces of F on MFMA (v_mfma_f32_16x16x32_bf16, FP32
// accumulation).  The dominant self term M_ii = I/(6 pi mu a_i) is applied
// in FP32 outside the MFMA.  Precision modes:
//   1  hi*hi                                  (plain bf16)
//   3  hi*hi + hi*lo + lo*hi                  (bf16x3)
//   6  bf16x3 + lo*lo + hi*F3 + M3*hi         (second-order split terms)
// where x = hi + lo + x3 with each piece bf16 (round to nearest even).
// The reduction over spheres is split over `split` workgroups per row tile
// (partials in FP32, summed by a second kernel) so N = 4096 fills the chip.
//
// Also here: the reference's equal-sphere walkthrough (Kern_getForce
// :257-300, Kern_interpolateVelocity :322-366, Kern_getmobu_walkthrough
// :207-255) and the dense product getMobU (:85-109).
#include "common.hpp"

#include <mutex>
#include <utility>
#include <vector>

namespace sphx {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int RPY_TP = 64;     // spheres per row tile (4 waves x 16)
constexpr int RPY_TQ = 32;     // spheres per reduction chunk (MFMA K)
constexpr int RPY_THREADS = 256;
constexpr float RPY_PI = 3.14159265358979f;

// round to nearest even to bf16 (finite inputs)
__host__ __device__ __forceinline__ uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf16_float(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// x = hi + lo + r3 (three bf16 pieces)
__device__ __forceinline__ void split3(float x, uint16_t& h, uint16_t& l, uint16_t& r) {
  h = bf16_bits(x);
  float d = x - bf16_float(h);
  l = bf16_bits(d);
  r = bf16_bits(d - bf16_float(l));
}

// F [3n][K] FP32 row-major -> pieces FT[b][c][q] (sphere q contiguous),
// zero-padded to np spheres.
__global__ void k_rpy_split_f(const float* __restrict__ F, int n, int np, int K,
                              uint16_t* __restrict__ Fh, uint16_t* __restrict__ Fl,
                              uint16_t* __restrict__ F3) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;   // over 3 * K * np
  if (t >= 3 * K * np) return;
  int q = t % np, c = (t / np) % K, b = t / (np * K);
  float x = q < n ? F[(3 * (size_t)q + b) * K + c] : 0.f;
  uint16_t h, l, r;
  split3(x, h, l, r);
  Fh[t] = h;
  Fl[t] = l;
  F3[t] = r;
}

typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;

// Two floats to packed bf16 (element 0 in the low half), round to nearest
// even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pk_bf16(float a, float b) {
  const bf16x2 v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, v);
}
// x0 = hi + lo + r3 and x1 likewise, the pieces of both packed per word.
__device__ __forceinline__ void split3_pk(float x0, float x1, uint32_t& h, uint32_t& l,
                                          uint32_t& r) {
  h = pk_bf16(x0, x1);
  const float d0 = x0 - __uint_as_float(h << 16), d1 = x1 - __uint_as_float(h & 0xFFFF0000u);
  l = pk_bf16(d0, d1);
  const float e0 = d0 - __uint_as_float(l << 16), e1 = d1 - __uint_as_float(l & 0xFFFF0000u);
  r = pk_bf16(e0, e1);
}

// The six distinct entries (00 11 22 01 02 12) of the polydisperse RPY
// pair block for spheres p = {x, y, z, a} and q, zero for q == p or
// outside the set.
__device__ __forceinline__ void rpy_pair(float4 sp, float4 sq, bool valid, float inv8pimu,
                                         float m[6]) {
#pragma unroll
  for (int k = 0; k < 6; ++k) m[k] = 0.f;
  if (!valid) return;
  const float dx = sq.x - sp.x, dy = sq.y - sp.y, dz = sq.z - sp.z;
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float ir2 = 1.f / r2;
  const float c0 = inv8pimu * sqrtf(ir2);
  const float s2 = sp.w * sp.w + sq.w * sq.w;
  const float ci = c0 * (1.f + s2 * ir2 * (1.f / 3.f));
  const float cr = c0 * (1.f - s2 * ir2) * ir2;
  m[0] = ci + cr * dx * dx;
  m[1] = ci + cr * dy * dy;
  m[2] = ci + cr * dz * dz;
  m[3] = cr * dx * dy;
  m[4] = cr * dx * dz;
  m[5] = cr * dy * dz;
}

// Symmetric 3x3 block index: 00 11 22 01 02 12
__device__ __forceinline__ int sym_index(int a, int b) {
  if (a == b) return a;
  int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo == 0 ? (hi == 1 ? 3 : 4) : 5;
}

__device__ __forceinline__ bf16x8 lds8(const uint16_t* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}
__device__ __forceinline__ bf16x8 glb8(const uint16_t* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}

// Partial products of the off-diagonal mobility for row tile blockIdx.x and
// reduction segment blockIdx.y: part[s][a][p][c].
// The pieces planes a precision mode multiplies: hi (mode 1), hi and lo
// (mode 3), all three (mode 6).
template <int MODE>
struct RpyPieces {
  static constexpr int N = MODE == 1 ? 1 : (MODE == 3 ? 2 : 3);
  uint16_t M[N][6][RPY_TP][RPY_TQ];
};

template <int MODE>
__global__ __launch_bounds__(RPY_THREADS, MODE == 1 ? 3 : 2) void k_rpy_tile(
    const float4* __restrict__ sph, int n, int np, float mu, int K, int seg,
    const uint16_t* __restrict__ Fh, const uint16_t* __restrict__ Fl,
    const uint16_t* __restrict__ F3, float* __restrict__ part) {
  constexpr int mode = MODE;
  __shared__ __align__(16) RpyPieces<MODE> P;
  auto& Mh = P.M[0];
  auto& Ml = P.M[MODE >= 3 ? 1 : 0];
  auto& M3 = P.M[MODE == 6 ? 2 : 0];
  const int p0 = blockIdx.x * RPY_TP;
  const int s = blockIdx.y;
  const int q_begin = s * seg, q_end = min(np, q_begin + seg);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // pair generation: thread -> spheres q0 + 2 (tid & 15) + {0, 1} and rows
  // p0 + (tid >> 4) + 16 t, t = 0..3; a wave writes 4 consecutive rows of
  // packed words (256 contiguous bytes per store)
  const int qp = 2 * (tid & 15), pr = tid >> 4;
  const float inv8pimu = 1.f / (8.f * RPY_PI * mu);
  // MFMA work of a wave: rows p0 + 32 wr .. +32 (two 16-row tiles) and
  // right-hand sides 32 wc .. +32 (two 16-column tiles), so each F piece is
  // loaded by two waves of the workgroup instead of four
  const int wr = wave & 1, wc = wave >> 1;
  f32x4 acc[3][2][2];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) acc[a][rt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int ncb = K / 16;
  // the thread's four row spheres are the same for every chunk; the two
  // column spheres of the next chunk are fetched during this chunk's MFMAs
  float4 sp4[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int gp = p0 + pr + 16 * t;
    sp4[t] = gp < n ? sph[gp] : make_float4(0.f, 0.f, 0.f, 1.f);
  }
  auto col = [&](int gq) { return gq < n ? sph[gq] : make_float4(0.f, 0.f, 0.f, 1.f); };
  float4 nq0 = col(q_begin + qp), nq1 = col(q_begin + qp + 1);
  for (int q0 = q_begin; q0 < q_end; q0 += RPY_TQ) {
    __syncthreads();
    {
      const int gq0 = q0 + qp, gq1 = gq0 + 1;
      const float4 sq0 = nq0, sq1 = nq1;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int pl = pr + 16 * t, gp = p0 + pl;
        const float4 sp = sp4[t];
        float m0[6], m1[6];
        rpy_pair(sp, sq0, gp < n && gq0 < n && gq0 != gp, inv8pimu, m0);
        rpy_pair(sp, sq1, gp < n && gq1 < n && gq1 != gp, inv8pimu, m1);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          uint32_t h, l, r;
          split3_pk(m0[k], m1[k], h, l, r);
          *reinterpret_cast<uint32_t*>(&Mh[k][pl][qp]) = h;
          if (mode >= 3) *reinterpret_cast<uint32_t*>(&Ml[k][pl][qp]) = l;
          if (mode == 6) *reinterpret_cast<uint32_t*>(&M3[k][pl][qp]) = r;
        }
      }
    }
    __syncthreads();
    if (q0 + RPY_TQ < q_end) {
      nq0 = col(q0 + RPY_TQ + qp);
      nq1 = col(q0 + RPY_TQ + qp + 1);
    }
    // MFMA lane layout of v_mfma_f32_16x16x32_bf16: A row = lane % 16,
    // k = 8 (lane / 16) .. +8; B column = lane % 16, same k; D rows
    // 4 (lane / 16) + i, column lane % 16.
    const int kk = 8 * (lane >> 4);
    // all three b unrolled: their F loads issue together
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      bf16x8 fh[2], fl[2], f3[2];
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        const int cb = 2 * wc + ct;
        if (cb >= ncb) break;
        size_t off = ((size_t)b * K + 16 * cb + (lane & 15)) * np + q0 + kk;
        fh[ct] = glb8(Fh + off);
        if (mode >= 3) fl[ct] = glb8(Fl + off);
        if (mode == 6) f3[ct] = glb8(F3 + off);
      }
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int k6 = sym_index(a, b);
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) {
          const int arow = 32 * wr + 16 * rt + (lane & 15);
          bf16x8 ah = lds8(&Mh[k6][arow][kk]);
          bf16x8 al = lds8(&Ml[k6][arow][kk]);
          bf16x8 a3;
          if (mode == 6) a3 = lds8(&M3[k6][arow][kk]);
#pragma unroll
          for (int ct = 0; ct < 2; ++ct) {
            if (2 * wc + ct >= ncb) break;
            f32x4 c = acc[a][rt][ct];
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, fh[ct], c, 0, 0, 0);
            if (mode >= 3) {
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, fl[ct], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, fh[ct], c, 0, 0, 0);
            }
            if (mode == 6) {
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, fl[ct], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, f3[ct], c, 0, 0, 0);
              c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, fh[ct], c, 0, 0, 0);
            }
            acc[a][rt][ct] = c;
          }
        }
      }
    }
  }
  // partials: part[s][a][p][c]
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        const int cb = 2 * wc + ct;
        if (cb >= ncb) break;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          int p = p0 + 32 * wr + 16 * rt + 4 * (lane >> 4) + i;
          int c = 16 * cb + (lane & 15);
          part[(((size_t)s * 3 + a) * np + p) * K + c] = acc[a][rt][ct][i];
        }
      }
}

// U[3p + a][c] = F/(6 pi mu a_p) + sum_s part[s][a][p][c]
__global__ void k_rpy_reduce(const float4* __restrict__ sph, int n, int np, float mu, int K,
                             int nseg, const float* __restrict__ F,
                             const float* __restrict__ part, float* __restrict__ U) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;   // over 3 n K
  if (t >= 3 * n * K) return;
  int c = t % K, row = t / K, a = row % 3, p = row / 3;
  float v = F[t] / (6.f * RPY_PI * mu * sph[p].w);
  for (int s = 0; s < nseg; ++s) v += part[(((size_t)s * 3 + a) * np + p) * K + c];
  U[t] = v;
}

// ------------------------------------------------------------- reference --
// PairwiseMobMatrix (:37-82): A1N and B1N for equal spheres a = 1.
__device__ void pair_ref(float dx, float dy, float dz, float A[9], float B[9]) {
  float dr = sqrtf(dx * dx + dy * dy + dz * dz);
  float eu[3] = {dx / dr, dy / dr, dz / dr};
  float dr3 = dr * dr * dr;
  float X1NA = 1.5f / dr - 1.f / dr3;
  float Y1NA = 0.75f / dr + 0.5f / dr3;
  float Y1NB = -3.f / (4.f * dr * dr);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      A[i * 3 + j] = X1NA * eu[i] * eu[j] + Y1NA * ((i == j ? 1.f : 0.f) - eu[i] * eu[j]);
  // B1N[i][j] = -Y1NB sum_k eps(j, i, k) e_k
  B[0] = 0.f;
  B[1] = -Y1NB * (-eu[2]);   // eps(1,0,2) = -1
  B[2] = -Y1NB * (eu[1]);    // eps(2,0,1) = +1
  B[3] = -Y1NB * (eu[2]);    // eps(0,1,2) = +1
  B[4] = 0.f;
  B[5] = -Y1NB * (-eu[0]);   // eps(2,1,0) = -1
  B[6] = -Y1NB * (-eu[1]);   // eps(0,2,1) = -1
  B[7] = -Y1NB * (eu[0]);    // eps(1,2,0) = +1
  B[8] = 0.f;
}

// Kern_getForce (:257-300): near-field repulsion added to force.
__global__ void k_walk_force(const sphx_displayPack* __restrict__ disp, int n,
                             sphx_float3* __restrict__ force) {
  __shared__ float4 tile[256];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float3 pi = i < n ? f3(disp[i].pos.x, disp[i].pos.y, disp[i].pos.z) : f3(0.f, 0.f, 0.f);
  float3 f = f3(0.f, 0.f, 0.f);
  for (int k0 = 0; k0 < n; k0 += 256) {
    __syncthreads();
    int k = k0 + threadIdx.x;
    tile[threadIdx.x] = k < n ? make_float4(disp[k].pos.x, disp[k].pos.y, disp[k].pos.z, 0.f)
                              : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    int cnt = min(256, n - k0);
    for (int t = 0; t < cnt; ++t) {
      if (k0 + t == i) continue;
      float4 q = tile[t];
      float dx = q.x - pi.x, dy = q.y - pi.y, dz = q.z - pi.z;
      float dr = sqrtf(dx * dx + dy * dy + dz * dz);
      const float actdist = 1.5f;
      if (dr < actdist) {
        float fac = 5.f * (expf(actdist / dr) - 2.7128f * dr / actdist);
        f.x += -dx * fac; f.y += -dy * fac; f.z += -dz * fac;
      }
    }
  }
  if (i < n) { force[i].x += f.x; force[i].y += f.y; force[i].z += f.z; }
}

// Kern_interpolateVelocity (:322-366): trilinear background velocity from
// the MPM nodes ({v, m} per node).
__global__ void k_walk_background(const sphx_displayPack* __restrict__ disp, int n,
                                  const float4* __restrict__ nodev, float3 xmin, float cs,
                                  int3 res, float ss, sphx_float3* __restrict__ unew) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float px = disp[i].pos.x - xmin.x, py = disp[i].pos.y - xmin.y, pz = disp[i].pos.z - xmin.z;
  int ix = (int)(px / cs), iy = (int)(py / cs), iz = (int)(pz / cs);
  float3 v = f3(0.f, 0.f, 0.f);
  for (int c = 0; c < 8; ++c) {
    int x = ix + (c & 1), y = iy + ((c >> 1) & 1), z = iz + (c >> 2);
    if (x < 0 || y < 0 || z < 0 || x >= res.x || y >= res.y || z >= res.z) continue;
    float w = (1.f - fabsf(px - x * cs) / cs) * (1.f - fabsf(py - y * cs) / cs) *
              (1.f - fabsf(pz - z * cs) / cs);
    float4 nv = nodev[((size_t)y * res.z + z) * res.x + x];
    v.x += nv.x * w; v.y += nv.y * w; v.z += nv.z * w;
  }
  unew[i] = {v.x / ss, v.y / ss, v.z / ss};
}

// Kern_getmobu_walkthrough (:207-255), positions double-buffered (the
// reference moves them in place while other threads read them, A10).
__global__ void k_walk_mob(const sphx_displayPack* __restrict__ disp, int n,
                           const sphx_float3* __restrict__ force, float dt,
                           sphx_float3* __restrict__ unew, sphx_float3* __restrict__ omega,
                           float4* __restrict__ newpos) {
  __shared__ float4 tp[256], tf[256];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float3 pi = i < n ? f3(disp[i].pos.x, disp[i].pos.y, disp[i].pos.z) : f3(0.f, 0.f, 0.f);
  float3 u = i < n ? f3(unew[i].x, unew[i].y, unew[i].z) : f3(0.f, 0.f, 0.f);
  float3 w = f3(0.f, 0.f, 0.f);
  for (int k0 = 0; k0 < n; k0 += 256) {
    __syncthreads();
    int k = k0 + threadIdx.x;
    tp[threadIdx.x] = k < n ? make_float4(disp[k].pos.x, disp[k].pos.y, disp[k].pos.z, 0.f)
                            : make_float4(0.f, 0.f, 0.f, 0.f);
    tf[threadIdx.x] = k < n ? make_float4(force[k].x, force[k].y, force[k].z, 0.f)
                            : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    int cnt = min(256, n - k0);
    for (int t = 0; t < cnt; ++t) {
      float4 q = tp[t], f = tf[t];
      if (k0 + t != i) {
        float A[9], B[9];
        pair_ref(q.x - pi.x, q.y - pi.y, q.z - pi.z, A, B);
        u.x += A[0] * f.x + A[1] * f.y + A[2] * f.z;
        u.y += A[3] * f.x + A[4] * f.y + A[5] * f.z;
        u.z += A[6] * f.x + A[7] * f.y + A[8] * f.z;
        w.x += B[0] * f.x + B[1] * f.y + B[2] * f.z;
        w.y += B[3] * f.x + B[4] * f.y + B[5] * f.z;
        w.z += B[6] * f.x + B[7] * f.y + B[8] * f.z;
      } else {
        u.x += f.x; u.y += f.y; u.z += f.z;
      }
    }
  }
  if (i >= n) return;
  unew[i] = {u.x, u.y, u.z};
  omega[i] = {w.x, w.y, w.z};
  newpos[i] = make_float4(pi.x + u.x * dt, pi.y + u.y * dt, pi.z + u.z * dt, 0.f);
}

__global__ void k_walk_commit(sphx_displayPack* __restrict__ disp, int n,
                              const float4* __restrict__ newpos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = newpos[i];
  disp[i].pos = {p.x, p.y, p.z};
}

__global__ void k_walk_zero(sphx_float3* __restrict__ u, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] = {0.f, 0.f, 0.f};
}

// Position double buffer of sphx_rpy_walkthrough_noalloc: one copy per GPU
// the library runs on, held until the process ends.  Host code reaches it
// only through the address hipGetSymbolAddress returns for that GPU.
static __device__ __align__(256) float4 scratch[SPHX_WALK_NOALLOC_MAX];

// MVmultiply (:85-98): one wave per row of the dense (6n x 6n) matrix.
__global__ void k_gemv(const float* __restrict__ mat, const float* __restrict__ f,
                       float* __restrict__ u, int dim) {
  int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (row >= dim) return;
  const float* m = mat + (size_t)row * dim;
  double s = 0.0;
  for (int j = lane; j < dim; j += 64) s += (double)m[j] * (double)f[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) u[row] = (float)s;
}

}  // namespace sphx

using namespace sphx;

struct sphx_rpy {
  int n = 0, np = 0, K = 0, split = 16, device = 0;
  float mu = 1.f;
  hipStream_t stream = nullptr;
  sphx::DevBuf<float4> sph;
  sphx::DevBuf<uint16_t> Fh, Fl, F3;
  sphx::DevBuf<float> part, F, U;
  // tile-kernel timing: an event pair around each k_rpy_tile launch
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

extern "C" {

int sphx_rpy_create(sphx_rpy** out, int n, int nrhs) {
  SPHX_CHECK(out, "sphx_rpy_create", "null argument");
  *out = nullptr;
  SPHX_CHECK(n > 0, "sphx_rpy_create", "n must be positive");
  SPHX_CHECK(nrhs > 0 && nrhs <= 64 && nrhs % 16 == 0, "sphx_rpy_create",
             "nrhs must be 16, 32, 48 or 64");
  int dev = 0;
  SPHX_HIP(hipGetDevice(&dev));
  sphx_rpy* h = new sphx_rpy;
  h->n = n;
  h->np = (n + RPY_TP - 1) / RPY_TP * RPY_TP;
  h->K = nrhs;
  h->device = dev;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete h; return hip_fail("hipStreamCreate", e); }
  size_t fsz = 3 * (size_t)nrhs * h->np;
  int rc = 0;
  if ((rc = h->sph.alloc(h->np)) || (rc = h->Fh.alloc(fsz)) || (rc = h->Fl.alloc(fsz)) ||
      (rc = h->F3.alloc(fsz)) || (rc = h->F.alloc(3 * (size_t)n * nrhs)) ||
      (rc = h->U.alloc(3 * (size_t)n * nrhs))) {
    (void)hipStreamDestroy(h->stream);
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

static void rpy_drop_events(sphx_rpy* h) {
  for (auto& e : h->ev) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  h->ev.clear();
}

int sphx_rpy_destroy(sphx_rpy* h) {
  if (!h) return 0;
  (void)hipStreamSynchronize(h->stream);
  rpy_drop_events(h);
  h->sph.release(); h->Fh.release(); h->Fl.release(); h->F3.release();
  h->part.release(); h->F.release(); h->U.release();
  (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

// Sphere centres [n][3] and radii [n] (host), viscosity mu.
int sphx_rpy_set_spheres(sphx_rpy* h, const float* pos, const float* radius, float mu) {
  SPHX_CHECK(h && pos && radius, "sphx_rpy_set_spheres", "null argument");
  SPHX_CHECK(mu > 0.f, "sphx_rpy_set_spheres", "mu must be positive");
  std::vector<float4> s(h->np, make_float4(0.f, 0.f, 0.f, 1.f));
  for (int i = 0; i < h->n; ++i) {
    SPHX_CHECK(radius[i] > 0.f, "sphx_rpy_set_spheres", "radii must be positive");
    s[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], radius[i]);
  }
  SPHX_HIP(hipSetDevice(h->device));
  SPHX_HIP(hipMemcpy(h->sph.p, s.data(), h->np * sizeof(float4), hipMemcpyHostToDevice));
  h->mu = mu;
  return 0;
}

// Reduction segments per row tile (work split over the spheres).
int sphx_rpy_set_split(sphx_rpy* h, int split) {
  SPHX_CHECK(h && split >= 1, "sphx_rpy_set_split", "bad argument");
  h->split = split;
  return 0;
}

// U = M F with device arrays dF, dU of [3n][nrhs] FP32 (row 3p + a),
// asynchronous on the handle's stream.  mode 1, 3 or 6 (see above).
int sphx_rpy_apply_device(sphx_rpy* h, const float* dF, float* dU, int mode) {
  SPHX_CHECK(h && dF && dU, "sphx_rpy_apply_device", "null argument");
  SPHX_CHECK(mode == 1 || mode == 3 || mode == 6, "sphx_rpy_apply_device", "mode must be 1, 3 or 6");
  SPHX_HIP(hipSetDevice(h->device));
  const int np = h->np, K = h->K;
  int chunks = np / RPY_TQ;
  int split = std::max(1, std::min(h->split, chunks));
  int seg = (chunks + split - 1) / split * RPY_TQ;
  int nseg = (np + seg - 1) / seg;
  SPHX_TRY(h->part.alloc((size_t)nseg * 3 * np * K));
  int tot = 3 * K * np;
  k_rpy_split_f<<<blocks_for(tot, 256), 256, 0, h->stream>>>(dF, h->n, np, K, h->Fh.p, h->Fl.p,
                                                             h->F3.p);
  dim3 grid(np / RPY_TP, nseg);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->timing) {
    // timing only: no system-scope fence at the records
    SPHX_HIP(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
    SPHX_HIP(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
    h->ev.emplace_back(e0, e1);
    SPHX_HIP(hipEventRecord(e0, h->stream));
  }
  if (mode == 1)
    k_rpy_tile<1><<<grid, RPY_THREADS, 0, h->stream>>>(h->sph.p, h->n, np, h->mu, K, seg, h->Fh.p,
                                                       h->Fl.p, h->F3.p, h->part.p);
  else if (mode == 3)
    k_rpy_tile<3><<<grid, RPY_THREADS, 0, h->stream>>>(h->sph.p, h->n, np, h->mu, K, seg, h->Fh.p,
                                                       h->Fl.p, h->F3.p, h->part.p);
  else
    k_rpy_tile<6><<<grid, RPY_THREADS, 0, h->stream>>>(h->sph.p, h->n, np, h->mu, K, seg, h->Fh.p,
                                                       h->Fl.p, h->F3.p, h->part.p);
  if (h->timing) SPHX_HIP(hipEventRecord(e1, h->stream));
  int outs = 3 * h->n * K;
  k_rpy_reduce<<<blocks_for(outs, 256), 256, 0, h->stream>>>(h->sph.p, h->n, np, h->mu, K, nseg,
                                                             dF, h->part.p, dU);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// Host arrays F, U of [3n][nrhs]; synchronous.
int sphx_rpy_apply(sphx_rpy* h, const float* F, float* U, int mode) {
  SPHX_CHECK(h && F && U, "sphx_rpy_apply", "null argument");
  SPHX_HIP(hipSetDevice(h->device));
  size_t sz = 3 * (size_t)h->n * h->K * sizeof(float);
  SPHX_HIP(hipMemcpyAsync(h->F.p, F, sz, hipMemcpyHostToDevice, h->stream));
  SPHX_TRY(sphx_rpy_apply_device(h, h->F.p, h->U.p, mode));
  SPHX_HIP(hipMemcpyAsync(U, h->U.p, sz, hipMemcpyDeviceToHost, h->stream));
  SPHX_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

// Time each k_rpy_tile launch with HIP events on the handle's stream (on)
// or stop and forget the recorded launches (off).
int sphx_rpy_timing(sphx_rpy* h, int on) {
  SPHX_CHECK(h, "sphx_rpy_timing", "null handle");
  SPHX_HIP(hipStreamSynchronize(h->stream));
  rpy_drop_events(h);
  h->timing = on != 0;
  return 0;
}

// Total milliseconds and count of the tile launches timed since timing was
// switched on; forgets them.
int sphx_rpy_timing_get(sphx_rpy* h, double* ms, int* count) {
  SPHX_CHECK(h && ms && count, "sphx_rpy_timing_get", "null argument");
  SPHX_HIP(hipStreamSynchronize(h->stream));
  double t = 0.0;
  for (auto& e : h->ev) {
    float f = 0.f;
    SPHX_HIP(hipEventElapsedTime(&f, e.first, e.second));
    t += f;
  }
  *ms = t;
  *count = (int)h->ev.size();
  rpy_drop_events(h);
  return 0;
}

int sphx_rpy_sync(sphx_rpy* h) {
  SPHX_CHECK(h, "sphx_rpy_sync", "null argument");
  SPHX_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

hipStream_t sphx_rpy_stream(sphx_rpy* h) { return h ? h->stream : nullptr; }

}  // extern "C"

namespace {

// The walkthrough's launches on the current device's null stream, with
// `newpos` (n float4) as the position double buffer; synchronous.
int walk_run(sphx_context* nodes, const sphx_stokesianBufList* b, int n, int with_force,
             float4* newpos) {
  int bl = blocks_for(n, 256);
  if (with_force) k_walk_force<<<bl, 256>>>(b->dispBuffer, n, b->cuForce);
  if (nodes && nodes->mpm.on) {
    const sphx_ParamCarrier& P = nodes->P;
    k_walk_background<<<bl, 256>>>(b->dispBuffer, n, nodes->mpm.nodev.p,
                                   make_float3(P.mpmXmin.x, P.mpmXmin.y, P.mpmXmin.z),
                                   P.mpmcellsize, make_int3(P.mpmRes.x, P.mpmRes.y, P.mpmRes.z),
                                   b->simscale, b->cuUnew);
  } else {
    k_walk_zero<<<bl, 256>>>(b->cuUnew, n);
  }
  k_walk_mob<<<bl, 256>>>(b->dispBuffer, n, b->cuForce, b->dt, b->cuUnew, b->cuOmega, newpos);
  k_walk_commit<<<bl, 256>>>(b->dispBuffer, n, newpos);
  SPHX_HIP(hipGetLastError());
  SPHX_HIP(hipDeviceSynchronize());
  return 0;
}

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (each GPU holds its own copy of the array).
constexpr int WALK_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float4* g_scratch[WALK_MAX_DEVICES] = {};

int walk_scratch(float4** out) {
  int dev = 0;
  SPHX_HIP(hipGetDevice(&dev));
  SPHX_CHECK(dev >= 0 && dev < WALK_MAX_DEVICES, "sphx_rpy_walkthrough_noalloc",
             "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    SPHX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SPHX_CHECK(p != nullptr, "sphx_rpy_walkthrough_noalloc", "no device address for the scratch");
    g_scratch[dev] = static_cast<float4*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

extern "C" {

// The reference's walkthrough step on device buffers: near-field force
// (added to cuForce), background velocity from the node grid of `nodes`
// (may be NULL: zero background), mobility velocity and rotation,
// positions advanced by U dt.
int sphx_rpy_walkthrough(sphx_context* nodes, const sphx_stokesianBufList* b, int n,
                         int with_force) {
  SPHX_CHECK(b && n >= 0, "sphx_rpy_walkthrough", "bad argument");
  if (n == 0) return 0;
  DevBuf<float4> np;
  SPHX_TRY(np.alloc(n));
  return walk_run(nodes, b, n, with_force, np.p);
}

// Same step with the position double buffer in the library's static device
// scratch: no allocation.  Every element the step reads of the scratch is
// written by k_walk_mob of the same call first, so nothing is cleared.
// Calls that overlap on one GPU share the scratch and race.
int sphx_rpy_walkthrough_noalloc(sphx_context* nodes, const sphx_stokesianBufList* b, int n,
                                 int with_force) {
  SPHX_CHECK(b && n >= 0, "sphx_rpy_walkthrough_noalloc", "bad argument");
  SPHX_CHECK(n <= SPHX_WALK_NOALLOC_MAX, "sphx_rpy_walkthrough_noalloc",
             "more spheres than the static scratch holds (SPHX_WALK_NOALLOC_MAX)");
  if (n == 0) return 0;
  float4* np = nullptr;
  SPHX_TRY(walk_scratch(&np));
  return walk_run(nodes, b, n, with_force, np);
}

// MVmultiply / getMobU: u = mat f with mat (6n x 6n) row-major, device.
int sphx_mob_gemv(const float* mat, const float* f, float* u, int nsize) {
  SPHX_CHECK(mat && f && u && nsize >= 0, "sphx_mob_gemv", "bad argument");
  int dim = 6 * nsize;
  if (dim == 0) return 0;
  k_gemv<<<blocks_for(dim, 4), 256>>>(mat, f, u, dim);
  SPHX_HIP(hipGetLastError());
  SPHX_HIP(hipDeviceSynchronize());
  return 0;
}

}  // extern "C"
