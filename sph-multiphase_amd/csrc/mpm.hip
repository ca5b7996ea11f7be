// MPM transfer with Drucker-Prager plasticity on a dense node grid.
//
// Reference (SPH_Solver/fluids/MpmSolver.cu): node-centric gathers over the
// 27 SPH cells around each node (P2G :823-889, node force :658-704) and
// particle-centric gathers over 4^3 nodes (stress :477-592, G2P :891-951),
// one thread per node or particle.  Here every transfer is particle-centric
// over the cell-sorted particles:
//   * P2G and the node force are scatters.  A workgroup of 256 consecutive
//     sorted particles accumulates into an LDS tile covering the bounding
//     box of its stencils (ds_add_f32), then adds each touched tile node to
//     global memory once: ~14 global atomics per particle at one particle
//     per cell instead of 4 x 27 (quadratic) or 4 x 64 (cubic).  A
//     workgroup whose box exceeds the tile adds directly.
//   * The node grid is dense (y-major ids as GetMpmIdx :35-38); only the
//     4^3-node blocks a transfer touched are listed, cleared and updated,
//     so a mostly empty grid costs nothing.
//   * Node velocity and mass live in one float4 so a stencil gather is one
//     16-byte load per node.
// Kernels: cubic B-spline of :155-195 (gradients in cell units, as the
// reference) or the quadratic spline; `physical` scales gradients to metres
// and uses the momentum-conserving node force m_p / (rho_p m_I) sigma grad N
// instead of the reference's sigma grad N / rho_p.
#include "common.hpp"

#include <climits>

namespace sphx {

struct MP {
  int n;
  uint32_t sentinel;
  float3 xmin;
  float dx;
  int3 res;
  int3 nblk;
  float ss, dt;
  float3 g;
  float3 smin, smax;
  float a_phi, a_psi, k_c;
  int kernel, physical;
  float iDp;
  float rc;
};

constexpr int MPM_THREADS = 256;
constexpr int MPM_TILE = 2048;   // nodes (float4) per workgroup tile: 32 KB
constexpr int MPM_BLK = 4;       // node block edge

// MpmSolver.cu:155-165, constants as written there
__device__ __forceinline__ float sp_cubic(float x) {
  x = fabsf(x);
  if (x > 2.f) return 0.f;
  if (x > 1.f) return -0.166666f * x * x * x + x * x - 2.f * x + 1.33333f;
  return 0.5f * x * x * x - x * x + 0.666666f;
}
// :173-186
__device__ __forceinline__ float sp_dcubic(float x) {
  if (x <= -2.f) return 0.f;
  if (x <= -1.f) return 0.5f * x * x + 2.f * x + 2.f;
  if (x <= 0.f) return -1.5f * x * x - 2.f * x;
  if (x < 1.f) return 1.5f * x * x - 2.f * x;
  if (x < 2.f) return -0.5f * x * x + 2.f * x - 2.f;
  return 0.f;
}
__device__ __forceinline__ float sp_quad(float x) {
  float a = fabsf(x);
  if (a < 0.5f) return 0.75f - a * a;
  if (a < 1.5f) { float t = 1.5f - a; return 0.5f * t * t; }
  return 0.f;
}
__device__ __forceinline__ float sp_dquad(float x) {
  float a = fabsf(x);
  if (a < 0.5f) return -2.f * x;
  if (a < 1.5f) return (x > 0.f ? -1.f : 1.f) * (1.5f - a);
  return 0.f;
}

// Stencil of a particle along the three axes: first node, weights and
// derivatives (argument (x_p - x_I) / dx, cell units) for W nodes.
template <int W>
struct Stencil {
  int base[3];
  float w[3][W], dw[3][W];
};

template <int W>
__device__ __forceinline__ void make_stencil(const MP& m, float4 p, Stencil<W>& s) {
  const float pc[3] = {p.x, p.y, p.z};
  const float xm[3] = {m.xmin.x, m.xmin.y, m.xmin.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float rel = (pc[a] - xm[a]) / m.dx;
    s.base[a] = W == 4 ? (int)rel - 1 : (int)floorf(rel - 0.5f);
#pragma unroll
    for (int k = 0; k < W; ++k) {
      float xi = xm[a] + (float)(s.base[a] + k) * m.dx;   // initMpm :51-53
      float r = (pc[a] - xi) / m.dx;
      s.w[a][k] = W == 4 ? sp_cubic(r) : sp_quad(r);
      s.dw[a][k] = W == 4 ? sp_dcubic(r) : sp_dquad(r);
    }
  }
}

__device__ __forceinline__ bool in_res(const MP& m, int x, int y, int z) {
  return x >= 0 && x < m.res.x && y >= 0 && y < m.res.y && z >= 0 && z < m.res.z;
}
__device__ __forceinline__ size_t node_id(const MP& m, int x, int y, int z) {
  return ((size_t)y * m.res.z + z) * m.res.x + x;
}
__device__ __forceinline__ int block_id(const MP& m, int bx, int by, int bz) {
  return (by * m.nblk.z + bz) * m.nblk.x + bx;
}

__device__ __forceinline__ void mark_block(int* flags, int* list, int* count, int b) {
  if (atomicCAS(&flags[b], 0, 1) == 0) list[atomicAdd(count, 1)] = b;
}

// Adds contrib(s, a, b, c) (float4) of every valid lane to the nodes of its
// stencil: through an LDS tile when the workgroup's box fits, else
// directly.  Marks every node block the stencils overlap.  Must be called
// by the whole workgroup (it synchronises).
template <int W, class F>
__device__ void scatter(const MP& m, bool valid, const Stencil<W>& s, float4* __restrict__ out,
                        int* __restrict__ flags, int* __restrict__ list, int* __restrict__ count,
                        F&& contrib) {
  __shared__ float4 tile[MPM_TILE];
  __shared__ int box[7];
  if (threadIdx.x < 3) { box[threadIdx.x] = INT_MAX; box[3 + threadIdx.x] = INT_MIN; }
  if (threadIdx.x == 0) box[6] = 0;
  __syncthreads();
  if (valid) {
    for (int a = 0; a < 3; ++a) { atomicMin(&box[a], s.base[a]); atomicMax(&box[3 + a], s.base[a]); }
    box[6] = 1;
  }
  __syncthreads();
  if (box[6] == 0) return;
  const int lo[3] = {box[0], box[1], box[2]};
  const int hi[3] = {box[3] + W - 1, box[4] + W - 1, box[5] + W - 1};
  const long long ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
  const long long V = ex * ey * ez;
  if (V <= MPM_TILE) {
    for (int t = threadIdx.x; t < (int)V; t += blockDim.x) tile[t] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int b = 0; b < W; ++b)
#pragma unroll
        for (int c = 0; c < W; ++c)
#pragma unroll
          for (int a = 0; a < W; ++a) {
            int x = s.base[0] + a, y = s.base[1] + b, z = s.base[2] + c;
            if (!in_res(m, x, y, z)) continue;
            float4 v = contrib(a, b, c);
            int li = (int)(((y - lo[1]) * ez + (z - lo[2])) * ex + (x - lo[0]));
            atomicAdd(&tile[li].x, v.x);
            atomicAdd(&tile[li].y, v.y);
            atomicAdd(&tile[li].z, v.z);
            atomicAdd(&tile[li].w, v.w);
          }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < (int)V; t += blockDim.x) {
      float4 v = tile[t];
      if (v.x == 0.f && v.y == 0.f && v.z == 0.f && v.w == 0.f) continue;
      int x = lo[0] + (int)(t % ex), z = lo[2] + (int)((t / ex) % ez), y = lo[1] + (int)(t / (ex * ez));
      if (!in_res(m, x, y, z)) continue;
      float4* o = out + node_id(m, x, y, z);
      atomicAdd(&o->x, v.x);
      atomicAdd(&o->y, v.y);
      atomicAdd(&o->z, v.z);
      atomicAdd(&o->w, v.w);
    }
    // node blocks of the box (clamped to the grid)
    int b0[3], nb[3];
    for (int a = 0; a < 3; ++a) {
      int r = a == 0 ? m.res.x : (a == 1 ? m.res.y : m.res.z);
      int l = max(lo[a], 0), h = min(hi[a], r - 1);
      b0[a] = l / MPM_BLK;
      nb[a] = h >= l ? h / MPM_BLK - b0[a] + 1 : 0;
    }
    int tot = nb[0] * nb[1] * nb[2];
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
      int bx = b0[0] + t % nb[0], bz = b0[2] + (t / nb[0]) % nb[2], by = b0[1] + t / (nb[0] * nb[2]);
      mark_block(flags, list, count, block_id(m, bx, by, bz));
    }
  } else if (valid) {
#pragma unroll
    for (int b = 0; b < W; ++b)
#pragma unroll
      for (int c = 0; c < W; ++c)
#pragma unroll
        for (int a = 0; a < W; ++a) {
          int x = s.base[0] + a, y = s.base[1] + b, z = s.base[2] + c;
          if (!in_res(m, x, y, z)) continue;
          float4 v = contrib(a, b, c);
          float4* o = out + node_id(m, x, y, z);
          atomicAdd(&o->x, v.x);
          atomicAdd(&o->y, v.y);
          atomicAdd(&o->z, v.z);
          atomicAdd(&o->w, v.w);
        }
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c)
        for (int a = 0; a < 2; ++a) {
          int x = min(max(s.base[0] + a * (W - 1), 0), m.res.x - 1);
          int y = min(max(s.base[1] + b * (W - 1), 0), m.res.y - 1);
          int z = min(max(s.base[2] + c * (W - 1), 0), m.res.z - 1);
          mark_block(flags, list, count, block_id(m, x / MPM_BLK, y / MPM_BLK, z / MPM_BLK));
        }
  }
}

// ------------------------------------------------------------------ P2G --
// MpmParticleToGrid_APIC (:823-889) / MpmParticleToGrid (:197-236):
// node {sum w m_p (v_p + B_p iDp (x_I - x_p) ss), sum w m_p}.
template <int W>
__global__ __launch_bounds__(MPM_THREADS) void k_mpm_p2g(
    MP m, int apic, const uint32_t* __restrict__ key, const float4* __restrict__ pos,
    const float4* __restrict__ vel, const float4* __restrict__ phys, const float* __restrict__ B,
    int cap, float4* __restrict__ acc, int* __restrict__ flags, int* __restrict__ list,
    int* __restrict__ count) {
  const int i = xcd_index();
  bool valid = false;
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f), v = p;
  float mp = 0.f, Bi[9];
  if (i < m.n) {
    p = pos[i];
    valid = type_of(p) != SPHX_TYPE_BOUNDARY && key[i] != m.sentinel;
    v = vel[i];
    mp = phys[i].x;
#pragma unroll
    for (int q = 0; q < 9; ++q) Bi[q] = apic ? B[(size_t)q * cap + i] * m.iDp : 0.f;
  }
  Stencil<W> s;
  make_stencil<W>(m, p, s);
  scatter<W>(m, valid, s, acc, flags, list, count, [&](int a, int b, int c) {
    float wt = s.w[0][a] * s.w[1][b] * s.w[2][c];
    float vb[3] = {v.x, v.y, v.z};
    if (apic) {
      float xi = m.xmin.x + (float)(s.base[0] + a) * m.dx;
      float yi = m.xmin.y + (float)(s.base[1] + b) * m.dx;
      float zi = m.xmin.z + (float)(s.base[2] + c) * m.dx;
      float d0 = (xi - p.x) * m.ss, d1 = (yi - p.y) * m.ss, d2 = (zi - p.z) * m.ss;
#pragma unroll
      for (int q = 0; q < 3; ++q) vb[q] += Bi[3 * q] * d0 + Bi[3 * q + 1] * d1 + Bi[3 * q + 2] * d2;
    }
    return make_float4(vb[0] * mp * wt, vb[1] * mp * wt, vb[2] * mp * wt, wt * mp);
  });
}

// Persistent loop over the listed node blocks: 64 threads, one node each.
#define FOR_LISTED_NODES(m, list, count, body)                                      \
  for (int bi = blockIdx.x; bi < *(count); bi += gridDim.x) {                      \
    const int blk = (list)[bi];                                                     \
    const int bx = blk % (m).nblk.x, bz = (blk / (m).nblk.x) % (m).nblk.z,          \
              by = blk / ((m).nblk.x * (m).nblk.z);                                 \
    const int t = threadIdx.x;                                                      \
    const int x = bx * MPM_BLK + (t & 3), z = bz * MPM_BLK + ((t >> 2) & 3),        \
              y = by * MPM_BLK + (t >> 4);                                          \
    if (in_res((m), x, y, z)) {                                                     \
      const size_t id = node_id((m), x, y, z);                                      \
      body                                                                          \
    }                                                                               \
  }

// node velocity = momentum / mass (or zero below 1e-9, :881-888)
__global__ __launch_bounds__(64) void k_mpm_normalize(MP m, const int* __restrict__ list,
                                                      const int* __restrict__ count,
                                                      const float4* __restrict__ acc,
                                                      float4* __restrict__ nodev) {
  FOR_LISTED_NODES(m, list, count, {
    float4 a = acc[id];
    nodev[id] = a.w > 0.000000001f ? make_float4(a.x / a.w, a.y / a.w, a.z / a.w, a.w)
                                   : make_float4(0.f, 0.f, 0.f, 0.f);
  })
}

// zero the accumulators of the listed blocks and unlist them
__global__ __launch_bounds__(64) void k_mpm_clear(MP m, const int* __restrict__ list,
                                                  const int* __restrict__ count,
                                                  float4* __restrict__ acc,
                                                  float4* __restrict__ force, int* __restrict__ flags) {
  for (int bi = blockIdx.x; bi < *count; bi += gridDim.x) {
    const int blk = list[bi];
    if (threadIdx.x == 0) flags[blk] = 0;
    const int bx = blk % m.nblk.x, bz = (blk / m.nblk.x) % m.nblk.z, by = blk / (m.nblk.x * m.nblk.z);
    const int t = threadIdx.x;
    const int x = bx * MPM_BLK + (t & 3), z = bz * MPM_BLK + ((t >> 2) & 3), y = by * MPM_BLK + (t >> 4);
    if (in_res(m, x, y, z)) {
      const size_t id = node_id(m, x, y, z);
      acc[id] = make_float4(0.f, 0.f, 0.f, 0.f);
      force[id] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

__global__ void k_mpm_set_count(int* count, int v) { *count = v; }

__global__ __launch_bounds__(64) void k_mpm_zero_force(MP m, const int* __restrict__ list,
                                                       const int* __restrict__ count,
                                                       float4* __restrict__ force) {
  FOR_LISTED_NODES(m, list, count, { force[id] = make_float4(0.f, 0.f, 0.f, 0.f); })
}

// ------------------------------------------------------ Drucker-Prager --
// :310-371 with the yield path completed as the formula of :361-368.
__device__ void drucker_prager(const MP& m, float G, float K, const float* eps, const float* sig,
                               float* plastic) {
#pragma unroll
  for (int i = 0; i < 9; ++i) plastic[i] = 0.f;
  float s[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) s[i] = sig[i];
  float trs = (s[0] + s[4] + s[8]) / 3.0f;
  s[0] -= trs; s[4] -= trs; s[8] -= trs;
  float j2 = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) j2 += 0.5f * s[i] * s[i];
  if (j2 <= 0.0000001f) return;
  j2 = sqrtf(j2);
  float i1 = sig[0] + sig[4] + sig[8];
  float f = j2 + m.a_phi * i1 - m.k_c;
  if (f <= 0.f) return;
  float tr_e = eps[0] + eps[4] + eps[8];
  float se = 0.f;
#pragma unroll
  for (int i = 0; i < 9; ++i) se += s[i] * eps[i];
  float lambda = 3.f * m.a_phi * K * tr_e + G / j2 * se;
  lambda /= (9.f * m.a_phi * m.a_psi * K + G);
#pragma unroll
  for (int ii = 0; ii < 3; ++ii)
#pragma unroll
    for (int ij = 0; ij < 3; ++ij) {
      plastic[ii * 3 + ij] = G / j2 * s[ii * 3 + ij] * lambda;
      if (ii == ij) plastic[ii * 3 + ij] += 3.f * m.a_psi * K * lambda;
    }
}

// ------------------------------------------------------- particle stress --
// MpmParticleStress (:477-592): velocity gradient and density from the
// nodes, rate-form stress with the Jaumann terms, Drucker-Prager return for
// granular particles, ProjectStress (:373-378).
template <int W>
__global__ __launch_bounds__(MPM_THREADS) void k_mpm_stress(
    MP m, const float4* __restrict__ pos, float4* __restrict__ phys,
    const float2* __restrict__ gk, const int* __restrict__ perm, const float4* __restrict__ nodev,
    float* __restrict__ S, int cap) {
  const int i = xcd_index();
  if (i >= m.n) return;
  const float4 p = pos[i];
  const int type = type_of(p);
  if (type == SPHX_TYPE_BOUNDARY) return;
  Stencil<W> s;
  make_stencil<W>(m, p, s);
  float vg[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float rho = 0.f;
#pragma unroll
  for (int b = 0; b < W; ++b)
#pragma unroll
    for (int c = 0; c < W; ++c)
#pragma unroll
      for (int a = 0; a < W; ++a) {
        int x = s.base[0] + a, y = s.base[1] + b, z = s.base[2] + c;
        if (!in_res(m, x, y, z)) continue;
        float4 nv = nodev[node_id(m, x, y, z)];
        float n0 = s.w[0][a], n1 = s.w[1][b], n2 = s.w[2][c];
        float wt = n0 * n1 * n2;
        float nw0 = s.dw[0][a] * n1 * n2, nw1 = s.dw[1][b] * n0 * n2, nw2 = s.dw[2][c] * n0 * n1;
        vg[0] += nv.x * nw0; vg[1] += nv.x * nw1; vg[2] += nv.x * nw2;
        vg[3] += nv.y * nw0; vg[4] += nv.y * nw1; vg[5] += nv.y * nw2;
        vg[6] += nv.z * nw0; vg[7] += nv.z * nw1; vg[8] += nv.z * nw2;
        rho += nv.w * wt / m.rc / m.rc / m.rc;
      }
  if (m.physical)
#pragma unroll
    for (int q = 0; q < 9; ++q) vg[q] /= m.rc;
  phys[i].y = rho;
  float de[9], om[9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      de[3 * a + b] = (vg[3 * a + b] + vg[3 * b + a]) * 0.5f;
      om[3 * a + b] = (vg[3 * a + b] - vg[3 * b + a]) * 0.5f;
    }
  float trE = (de[0] + de[4] + de[8]) / 3.0f;
  float sig[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) sig[q] = S[(size_t)q * cap + i];
  const float2 g2 = gk[perm[i]];
  const float G = g2.x, K = g2.y;
  float ds[9];
#pragma unroll
  for (int ii = 0; ii < 3; ++ii)
#pragma unroll
    for (int ij = 0; ij < 3; ++ij) {
      float v = 2.f * G * de[3 * ii + ij];
#pragma unroll
      for (int ik = 0; ik < 3; ++ik)
        v += om[3 * ii + ik] * sig[3 * ik + ij] - sig[3 * ii + ik] * om[3 * ik + ij];
      if (ii == ij) v += K * trE * 3.0f - 2.f * G * trE;
      ds[3 * ii + ij] = v;
    }
  float pl[9];
  if (type == SPHX_TYPE_GRANULAR) drucker_prager(m, G, K, de, sig, pl);
  else
#pragma unroll
    for (int q = 0; q < 9; ++q) pl[q] = 0.f;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    sig[q] += ds[q] * m.dt;
    sig[q] -= pl[q] * m.dt;
  }
  if (type == SPHX_TYPE_GRANULAR && sig[0] + sig[4] + sig[8] > 0.f)
#pragma unroll
    for (int q = 0; q < 9; ++q) sig[q] = 0.f;
#pragma unroll
  for (int q = 0; q < 9; ++q) S[(size_t)q * cap + i] = sig[q];
}

// ------------------------------------------------------------ node force --
// MpmNodeUpdate's sum (:676-697) as a scatter: sigma_p grad N((x_I - x_p)/dx)
// / rho_p (reference) or m_p / (rho_p rc) sigma_p grad N (physical; divided
// by m_I in the node pass).
template <int W>
__global__ __launch_bounds__(MPM_THREADS) void k_mpm_force(
    MP m, const uint32_t* __restrict__ key, const float4* __restrict__ pos,
    const float4* __restrict__ phys, const float* __restrict__ S, int cap,
    float4* __restrict__ force, int* __restrict__ flags, int* __restrict__ list,
    int* __restrict__ count) {
  const int i = xcd_index();
  bool valid = false;
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
  float sig[9], scale = 0.f;
  if (i < m.n) {
    p = pos[i];
    valid = type_of(p) != SPHX_TYPE_BOUNDARY && key[i] != m.sentinel;
    float4 q = phys[i];
    scale = m.physical ? q.x / (q.y * m.rc) : 1.f / q.y;
#pragma unroll
    for (int k = 0; k < 9; ++k) sig[k] = S[(size_t)k * cap + i];
  }
  Stencil<W> s;
  make_stencil<W>(m, p, s);
  scatter<W>(m, valid, s, force, flags, list, count, [&](int a, int b, int c) {
    float n0 = s.w[0][a], n1 = s.w[1][b], n2 = s.w[2][c];
    // the reference differentiates at (x_I - x_p)/dx: the stencil's
    // derivative (at (x_p - x_I)/dx) with the sign flipped
    float nw0 = -s.dw[0][a] * n1 * n2, nw1 = -s.dw[1][b] * n0 * n2, nw2 = -s.dw[2][c] * n0 * n1;
    float f0 = sig[0] * nw0 + sig[1] * nw1 + sig[2] * nw2;
    float f1 = sig[3] * nw0 + sig[4] * nw1 + sig[5] * nw2;
    float f2 = sig[6] * nw0 + sig[7] * nw1 + sig[8] * nw2;
    return make_float4(f0 * scale, f1 * scale, f2 * scale, 0.f);
  });
}

// GridCollisionDetection (:595-656), EPSILON 1e-6 at this use site
__device__ __forceinline__ void collide(const MP& m, float x, float y, float z, float4& v) {
  const float eps = 1e-6f;
  if (-(y - m.smin.y) * m.ss > eps && v.y < 0.f) { v.x = 0.f; v.y = 0.f; v.z = 0.f; }
  if (-(x - m.smin.x) * m.ss > eps && v.x < 0.f) v.x = 0.f;
  if (-(z - m.smin.z) * m.ss > eps && v.z < 0.f) v.z = 0.f;
  if ((x - m.smax.x) * m.ss > eps && -v.x < 0.f) v.x = 0.f;
  if ((z - m.smax.z) * m.ss > eps && -v.z < 0.f) v.z = 0.f;
}

// MpmNodeUpdate (:658-704): v += (f + g) dt for nodes with mass, collision.
__global__ __launch_bounds__(64) void k_mpm_node(MP m, const int* __restrict__ list,
                                                 const int* __restrict__ count,
                                                 const float4* __restrict__ force,
                                                 float4* __restrict__ nodev) {
  FOR_LISTED_NODES(m, list, count, {
    float4 v = nodev[id];
    if (v.w >= 0.000000001f) {
      float4 f = force[id];
      float sc = m.physical ? 1.f / v.w : 1.f;
      v.x += (f.x * sc + m.g.x) * m.dt;
      v.y += (f.y * sc + m.g.y) * m.dt;
      v.z += (f.z * sc + m.g.z) * m.dt;
      collide(m, m.xmin.x + (float)x * m.dx, m.xmin.y + (float)y * m.dx,
              m.xmin.z + (float)z * m.dx, v);
      nodev[id] = v;
    }
  })
}

// ------------------------------------------------------------------ G2P --
// MpmParticleUpdate_APIC (:891-951) / MpmParticleUpdate (:706-756).
template <int W>
__global__ __launch_bounds__(MPM_THREADS) void k_mpm_g2p(MP m, int apic, float4* __restrict__ pos,
                                                         float4* __restrict__ vel,
                                                         const float4* __restrict__ nodev,
                                                         float* __restrict__ B, int cap) {
  const int i = xcd_index();
  if (i >= m.n) return;
  float4 p = pos[i];
  if (type_of(p) == SPHX_TYPE_BOUNDARY) return;
  Stencil<W> s;
  make_stencil<W>(m, p, s);
  float v[3] = {0.f, 0.f, 0.f}, Bn[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int b = 0; b < W; ++b)
#pragma unroll
    for (int c = 0; c < W; ++c)
#pragma unroll
      for (int a = 0; a < W; ++a) {
        int x = s.base[0] + a, y = s.base[1] + b, z = s.base[2] + c;
        if (!in_res(m, x, y, z)) continue;
        float4 nv = nodev[node_id(m, x, y, z)];
        float wt = s.w[0][a] * s.w[1][b] * s.w[2][c];
        float d0 = (m.xmin.x + (float)x * m.dx - p.x) * m.ss;
        float d1 = (m.xmin.y + (float)y * m.dx - p.y) * m.ss;
        float d2 = (m.xmin.z + (float)z * m.dx - p.z) * m.ss;
        v[0] += nv.x * wt; v[1] += nv.y * wt; v[2] += nv.z * wt;
        Bn[0] += wt * (nv.x * d0); Bn[1] += wt * (nv.x * d1); Bn[2] += wt * (nv.x * d2);
        Bn[3] += wt * (nv.y * d0); Bn[4] += wt * (nv.y * d1); Bn[5] += wt * (nv.y * d2);
        Bn[6] += wt * (nv.z * d0); Bn[7] += wt * (nv.z * d1); Bn[8] += wt * (nv.z * d2);
      }
  float4 vo = vel[i];
  vo.x = v[0]; vo.y = v[1]; vo.z = v[2];
  vel[i] = vo;
  if (apic)
#pragma unroll
    for (int q = 0; q < 9; ++q) B[(size_t)q * cap + i] = Bn[q];
  p.x += vo.x * m.dt / m.ss;
  p.y += vo.y * m.dt / m.ss;
  p.z += vo.z * m.dt / m.ss;
  pos[i] = p;
}

// --------------------------------------------- hot B / sigma <-> cold mats --
// cold_mat holds per upload slot 36 floats: deformGrad, stress, B, invA.
__global__ void k_mpm_load(const int* __restrict__ perm, int n, const float* __restrict__ mat,
                           float* __restrict__ S, float* __restrict__ B, int cap) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float* src = mat + (size_t)perm[s] * 36;
  for (int q = 0; q < 9; ++q) {
    S[(size_t)q * cap + s] = src[9 + q];
    B[(size_t)q * cap + s] = src[18 + q];
  }
}
__global__ void k_mpm_store(const int* __restrict__ perm, int n, float* __restrict__ mat,
                            const float* __restrict__ S, const float* __restrict__ B, int cap) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float* dst = mat + (size_t)perm[s] * 36;
  for (int q = 0; q < 9; ++q) {
    dst[9 + q] = S[(size_t)q * cap + s];
    dst[18 + q] = B[(size_t)q * cap + s];
  }
}
__global__ void k_mpm_reorder(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ idsort,
                              int n, uint32_t sentinel, const float* __restrict__ Sa,
                              float* __restrict__ Sb, const float* __restrict__ Ba,
                              float* __restrict__ Bb, int cap) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int src = skey[s] != sentinel ? (int)idsort[s] : s;
  for (int q = 0; q < 9; ++q) {
    Sb[(size_t)q * cap + s] = Sa[(size_t)q * cap + src];
    Bb[(size_t)q * cap + s] = Ba[(size_t)q * cap + src];
  }
}

// node {v, m} from host arrays; every block listed
__global__ void k_mpm_list_all(int* __restrict__ flags, int* __restrict__ list, int nb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  flags[b] = 1;
  list[b] = b;
}

// ----------------------------------------------------------------- host --
MP make_mp(const sphx_context* c) {
  MP m;
  const sphx_ParamCarrier& P = c->P;
  m.n = c->num;
  m.sentinel = (uint32_t)P.gridtotal;
  m.xmin = make_float3(P.mpmXmin.x, P.mpmXmin.y, P.mpmXmin.z);
  m.dx = P.mpmcellsize;
  m.res = make_int3(P.mpmRes.x, P.mpmRes.y, P.mpmRes.z);
  m.nblk = make_int3(c->mpm.nblk[0], c->mpm.nblk[1], c->mpm.nblk[2]);
  m.ss = P.simscale;
  m.dt = P.dt;
  m.g = make_float3(P.gravity.x, P.gravity.y, P.gravity.z);
  m.smin = make_float3(P.softminx.x, P.softminx.y, P.softminx.z);
  m.smax = make_float3(P.softmaxx.x, P.softmaxx.y, P.softmaxx.z);
  m.a_phi = P.a_phi;
  m.a_psi = P.a_psi;
  m.k_c = P.k_c;
  m.kernel = c->mpm.kernel;
  m.physical = c->mpm.physical;
  float iDp = P.mpmcellsize * P.mpmcellsize * P.simscale * P.simscale /
              (c->mpm.kernel == 0 ? 3.0f : 4.0f);   // :844-845 (D_p = dx^2/3 cubic, /4 quadratic)
  m.iDp = 1.0f / iDp;
  m.rc = P.mpmcellsize * P.simscale;
  return m;
}

int mpm_reorder(sphx_context* c, int a, int b) {
  auto& M = c->mpm;
  if (!M.on || !M.loaded || c->num == 0) return 0;
  k_mpm_reorder<<<blocks_for(c->num, 256), 256, 0, c->stream>>>(
      c->key.p, c->idsort.p, c->num, (uint32_t)c->P.gridtotal, M.S[a].p, M.S[b].p, M.B[a].p,
      M.B[b].p, c->cap);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// Before a download: write the particles' B and sigma back to the cold mats.
int mpm_sync_cold(sphx_context* c) {
  auto& M = c->mpm;
  if (!M.on || !M.loaded || c->num == 0) return 0;
  k_mpm_store<<<blocks_for(c->num, 256), 256, 0, c->stream>>>(c->PERM(), c->num, c->cold_mat.p,
                                                               M.S[c->cur].p, M.B[c->cur].p, c->cap);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// After an upload: the cold mats are the truth until the next MPM call.
void mpm_invalidate(sphx_context* c) { c->mpm.loaded = false; }

static int mpm_need(sphx_context* c, const char* who) {
  SPHX_CHECK(c, who, "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(c->mpm.on, who, "MPM not set up (call sphx_mpm_setup)");
  auto& M = c->mpm;
  if (!M.loaded && c->num > 0) {
    k_mpm_load<<<blocks_for(c->num, 256), 256, 0, c->stream>>>(c->PERM(), c->num, c->cold_mat.p,
                                                                M.S[c->cur].p, M.B[c->cur].p, c->cap);
    SPHX_HIP(hipGetLastError());
  }
  M.loaded = true;
  return 0;
}

static int persistent_blocks(const sphx_context* c) {
  return (int)std::min<size_t>(c->mpm.blocks, 4096);
}

int grid_hash(sphx_context* c);
int grid_sort(sphx_context* c);
int grid_reorder(sphx_context* c);

}  // namespace sphx

using namespace sphx;

extern "C" {

int sphx_mpm_setup(sphx_context* c, int kernel, int physical) {
  SPHX_CHECK(c, "sphx_mpm_setup", "null context");
  SPHX_CHECK(kernel == 0 || kernel == 1, "sphx_mpm_setup", "kernel must be 0 (cubic) or 1 (quadratic)");
  SPHX_TRY(ensure_device(c));
  const sphx_ParamCarrier& P = c->P;
  SPHX_CHECK(P.mpmcellsize > 0.f && P.simscale > 0.f, "sphx_mpm_setup", "mpmcellsize must be > 0");
  SPHX_CHECK(P.mpmRes.x > 0 && P.mpmRes.y > 0 && P.mpmRes.z > 0, "sphx_mpm_setup",
             "mpmRes must be positive");
  size_t nodes = (size_t)P.mpmRes.x * P.mpmRes.y * P.mpmRes.z;
  SPHX_CHECK(nodes < ((size_t)1 << 31), "sphx_mpm_setup", "node grid too large");
  // the reference gathers node sums over the 27 SPH cells around a node:
  // the same particles as this scatter only if a cell spans the support
  float support = (kernel == 0 ? 2.0f : 1.5f) * P.mpmcellsize;
  float cell = 1.0f / std::max(std::max(P.gridIdfac.x, P.gridIdfac.y), P.gridIdfac.z);
  SPHX_CHECK(cell >= support * 0.999f, "sphx_mpm_setup",
             "SPH grid cell narrower than the B-spline support (node sums would miss particles)");
  auto& M = c->mpm;
  M.kernel = kernel;
  M.physical = physical != 0;
  M.nodes = nodes;
  for (int a = 0; a < 3; ++a) {
    int r = a == 0 ? P.mpmRes.x : (a == 1 ? P.mpmRes.y : P.mpmRes.z);
    M.nblk[a] = (r + MPM_BLK - 1) / MPM_BLK;
  }
  M.blocks = (size_t)M.nblk[0] * M.nblk[1] * M.nblk[2];
  size_t cap = (size_t)std::max(c->cap, 1);
  for (int b = 0; b < 2; ++b) {
    SPHX_TRY(M.B[b].alloc(9 * cap));
    SPHX_TRY(M.S[b].alloc(9 * cap));
  }
  SPHX_TRY(M.acc.alloc(nodes));
  SPHX_TRY(M.force.alloc(nodes));
  SPHX_TRY(M.nodev.alloc(nodes));
  SPHX_TRY(M.flags.alloc(M.blocks));
  SPHX_TRY(M.list.alloc(M.blocks));
  SPHX_TRY(M.count.alloc(1));
  SPHX_HIP(hipMemsetAsync(M.acc.p, 0, nodes * sizeof(float4), c->stream));
  SPHX_HIP(hipMemsetAsync(M.force.p, 0, nodes * sizeof(float4), c->stream));
  SPHX_HIP(hipMemsetAsync(M.nodev.p, 0, nodes * sizeof(float4), c->stream));
  SPHX_HIP(hipMemsetAsync(M.flags.p, 0, M.blocks * sizeof(int), c->stream));
  SPHX_HIP(hipMemsetAsync(M.count.p, 0, sizeof(int), c->stream));
  M.on = true;
  M.loaded = false;
  SPHX_TRY(mpm_need(c, "sphx_mpm_setup"));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// MpmParticleToGrid_APIC (apic = 1) / MpmParticleToGrid (apic = 0)
int sphx_mpm_p2g(sphx_context* c, int apic) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_p2g"));
  SPHX_CHECK(c->sorted_valid, "sphx_mpm_p2g", "grid not built (call sphx_grid_build first)");
  auto& M = c->mpm;
  MP m = make_mp(c);
  int pb = persistent_blocks(c);
  timer_begin(c, "mpm_clear");
  k_mpm_clear<<<pb, 64, 0, c->stream>>>(m, M.list.p, M.count.p, M.acc.p, M.force.p, M.flags.p);
  k_mpm_set_count<<<1, 1, 0, c->stream>>>(M.count.p, 0);
  timer_end(c);
  if (m.n > 0) {
    timer_begin(c, "mpm_p2g");
    if (M.kernel == 0)
      k_mpm_p2g<4><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
          m, apic, c->key.p, c->POS(), c->VEL(), c->PHYS(), M.B[c->cur].p, c->cap, M.acc.p,
          M.flags.p, M.list.p, M.count.p);
    else
      k_mpm_p2g<3><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
          m, apic, c->key.p, c->POS(), c->VEL(), c->PHYS(), M.B[c->cur].p, c->cap, M.acc.p,
          M.flags.p, M.list.p, M.count.p);
    timer_end(c);
  }
  timer_begin(c, "mpm_normalize");
  k_mpm_normalize<<<pb, 64, 0, c->stream>>>(m, M.list.p, M.count.p, M.acc.p, M.nodev.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// MpmParticleStress
int sphx_mpm_stress(sphx_context* c) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_stress"));
  auto& M = c->mpm;
  MP m = make_mp(c);
  if (m.n == 0) return 0;
  timer_begin(c, "mpm_stress");
  if (M.kernel == 0)
    k_mpm_stress<4><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
        m, c->POS(), c->PHYS(), c->cold_GK.p, c->PERM(), M.nodev.p, M.S[c->cur].p, c->cap);
  else
    k_mpm_stress<3><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
        m, c->POS(), c->PHYS(), c->cold_GK.p, c->PERM(), M.nodev.p, M.S[c->cur].p, c->cap);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// MpmNodeUpdate: node force from the particles' stress, then v += (f+g) dt.
int sphx_mpm_node_update(sphx_context* c) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_node_update"));
  SPHX_CHECK(c->sorted_valid, "sphx_mpm_node_update", "grid not built");
  auto& M = c->mpm;
  MP m = make_mp(c);
  k_mpm_zero_force<<<persistent_blocks(c), 64, 0, c->stream>>>(m, M.list.p, M.count.p, M.force.p);
  if (m.n > 0) {
    timer_begin(c, "mpm_force");
    if (M.kernel == 0)
      k_mpm_force<4><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
          m, c->key.p, c->POS(), c->PHYS(), M.S[c->cur].p, c->cap, M.force.p, M.flags.p,
          M.list.p, M.count.p);
    else
      k_mpm_force<3><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
          m, c->key.p, c->POS(), c->PHYS(), M.S[c->cur].p, c->cap, M.force.p, M.flags.p,
          M.list.p, M.count.p);
    timer_end(c);
  }
  timer_begin(c, "mpm_node");
  k_mpm_node<<<persistent_blocks(c), 64, 0, c->stream>>>(m, M.list.p, M.count.p, M.force.p,
                                                          M.nodev.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// MpmParticleUpdate_APIC (apic = 1) / MpmParticleUpdate (apic = 0)
int sphx_mpm_g2p(sphx_context* c, int apic) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_g2p"));
  auto& M = c->mpm;
  MP m = make_mp(c);
  if (m.n == 0) return 0;
  c->nbr_valid = false;
  timer_begin(c, "mpm_g2p");
  if (M.kernel == 0)
    k_mpm_g2p<4><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
        m, apic, c->POS(), c->VEL(), M.nodev.p, M.B[c->cur].p, c->cap);
  else
    k_mpm_g2p<3><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
        m, apic, c->POS(), c->VEL(), M.nodev.p, M.B[c->cur].p, c->cap);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// RunMPM (fluid_system.cpp:541-564): sort, APIC P2G, stress, node update,
// APIC G2P.
int sphx_mpm_substep(sphx_context* c) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_substep"));
  SPHX_TRY(grid_hash(c));
  SPHX_TRY(grid_sort(c));
  SPHX_TRY(grid_reorder(c));
  SPHX_TRY(sphx_mpm_p2g(c, 1));
  SPHX_TRY(sphx_mpm_stress(c));
  SPHX_TRY(sphx_mpm_node_update(c));
  SPHX_TRY(sphx_mpm_g2p(c, 1));
  return timer_flush(c);
}

// Node velocities [N][3] and masses [N] in GetMpmIdx order (either NULL).
int sphx_mpm_download_nodes(sphx_context* c, float* v, float* mass) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_download_nodes"));
  auto& M = c->mpm;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  std::vector<float4> t(M.nodes);
  SPHX_HIP(hipMemcpy(t.data(), M.nodev.p, M.nodes * sizeof(float4), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < M.nodes; ++i) {
    if (v) { v[3 * i] = t[i].x; v[3 * i + 1] = t[i].y; v[3 * i + 2] = t[i].z; }
    if (mass) mass[i] = t[i].w;
  }
  return 0;
}

// Set node velocities and masses (tests and restarts); lists every block.
int sphx_mpm_upload_nodes(sphx_context* c, const float* v, const float* mass) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_upload_nodes"));
  SPHX_CHECK(v && mass, "sphx_mpm_upload_nodes", "null argument");
  auto& M = c->mpm;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  std::vector<float4> t(M.nodes);
  for (size_t i = 0; i < M.nodes; ++i) t[i] = make_float4(v[3 * i], v[3 * i + 1], v[3 * i + 2], mass[i]);
  SPHX_HIP(hipMemcpy(M.nodev.p, t.data(), M.nodes * sizeof(float4), hipMemcpyHostToDevice));
  MP m = make_mp(c);
  k_mpm_clear<<<persistent_blocks(c), 64, 0, c->stream>>>(m, M.list.p, M.count.p, M.acc.p,
                                                           M.force.p, M.flags.p);
  k_mpm_list_all<<<blocks_for((int)M.blocks, 256), 256, 0, c->stream>>>(M.flags.p, M.list.p,
                                                                        (int)M.blocks);
  k_mpm_set_count<<<1, 1, 0, c->stream>>>(M.count.p, (int)M.blocks);
  SPHX_HIP(hipGetLastError());
  SPHX_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int sphx_mpm_active_blocks(sphx_context* c, int* count) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_active_blocks"));
  SPHX_CHECK(count, "sphx_mpm_active_blocks", "null argument");
  SPHX_HIP(hipStreamSynchronize(c->stream));
  SPHX_HIP(hipMemcpy(count, c->mpm.count.p, sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
