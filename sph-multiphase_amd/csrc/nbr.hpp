// Neighbour traversal over the dense cell table.
//
// The reference visits 27 cells c = gc + neighborid[k] for k = 0..26
// (fluid_system.cpp:140-143: k = i*9 + j*3 + l with i->y, j->z, l->x) and
// the members [gridoff, gridoff+gridcnt) of each.  The three cells of one
// (dy, dz) row have consecutive ids, so their members form one contiguous
// range of the sorted arrays; visiting nine ranges in row order visits the
// same particles in the same order as the reference's 27-cell loop.
#pragma once

#include "common.hpp"

namespace sphx {

template <class F>
__device__ __forceinline__ void for_each_neighbor(const KP& k, const int* __restrict__ cell_start,
                                                  uint32_t cell, F&& f) {
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    int c0 = (int)cell + k.rowoff[r];
    int jb = cell_start[c0];
    int je = cell_start[c0 + 3];
#pragma unroll 1
    for (int j = jb; j < je; ++j) f(j);
  }
}

// Per-step neighbour list, sliced and packed.  Entries are 16 bits:
// e = d << 14 | (j - S_d), where d = dy + 1 is the neighbour's cell layer
// relative to particle i and S_d = cell_start of the first cell of that
// layer's first row (dz = -1, dx = -1), i.e. rows 3d .. 3d+2 of the 27-cell
// neighbourhood are one index window starting at S_d.  Eight entries form
// one 16-byte group; particles are grouped in slices of 256, and group g of
// particle i is at idx[(slice * capg + g) * 256 + (i & 255)], so a wave
// reading group g of its 64 particles touches one 1 KiB run.  Entries follow
// the reference's visiting order and hold every j != i with |x_ij| < 2h (the
// widest support any kernel of the step uses).  Positions do not move
// between the density pass and the last pressure iteration, so one list
// serves every neighbour loop of an IISPH step.  A particle with more than
// `cap` neighbours, one whose offset j - S_d does not fit 14 bits, or one
// whose layer starts are 2^24 or more particles apart has count cap + 1
// and its loops fall back to the cell scan, which visits a superset in the
// same order.
constexpr int NBR_SLICE = 256;
constexpr int NBR_GROUP = 8;                 // entries per 16-byte group
constexpr int NBR_TBITS = 14;                // offset bits of an entry
constexpr uint32_t NBR_TMAX = (1u << NBR_TBITS) - 1;

// Count of a boundary particle or one outside the grid: it has no list
// (the loops that skip such particles read only the count).
constexpr int NBR_INVALID = -2;

__host__ __device__ __forceinline__ size_t nbr_gslot(int i, int g, int capg) {
  return ((size_t)(i / NBR_SLICE) * capg + g) * NBR_SLICE + (i % NBR_SLICE);
}

struct NbrList {
  const uint4* __restrict__ idx;   // packed groups
  const int* __restrict__ cnt;     // entries; > cap: use the cell scan
  int cap;                         // entries per particle, a multiple of NBR_GROUP
};

// Window starts of the three cell layers around `cell`: S_0 and the
// offsets S_1 - S_0, S_2 - S_0 (the list builder falls back to the cell
// scan for a particle whose offsets do not fit 24 bits).
struct NbrBase {
  int s0;
  uint32_t d1, d2;
};
__device__ __forceinline__ NbrBase nbr_bases(const KP& k, const int* __restrict__ cs,
                                             uint32_t cell) {
  const int s0 = cs[(int)cell + k.rowoff[0]];
  return NbrBase{s0, (uint32_t)(cs[(int)cell + k.rowoff[3]] - s0),
                 (uint32_t)(cs[(int)cell + k.rowoff[6]] - s0)};
}

// j = S_d + t for the 16-bit entry e = d << 14 | t, d in {0, 1, 2}: the
// layer bits select the offsets by 24-bit multiply-adds instead of compares.
__device__ __forceinline__ int nbr_decode(uint32_t e, const NbrBase& b) {
  const uint32_t t = (uint32_t)b.s0 + (e & NBR_TMAX);
  return (int)(t + __umul24((e >> NBR_TBITS) & 1u, b.d1) + __umul24(e >> (NBR_TBITS + 1), b.d2));
}

// Batched traversal of the list: B entries at a time, and the records of
// all of them fetched (load(j)) before any is used (use(j, rec)), so a
// batch's gathers are in flight together instead of one dependent round
// trip per neighbour; with PREFETCH the next batch's entry groups are
// fetched before the current batch's records are used (loops whose live
// state leaves room for the extra registers).  Entries past the count decode to S_0
// (a sorted index <= i, inside the arrays) and are not used; groups past
// the last written one are not read.  The visiting order, and so every
// sum's order, is that of the list.
constexpr int NBR_BATCH = 8;          // default batch (wide records)
constexpr int NBR_BATCH_SMALL = 16;   // loops with few live values (density)

template <int B = NBR_BATCH, bool PREFETCH = true, class Load, class Use>
__device__ __forceinline__ void for_each_listed_b(const KP& k, const NbrList& L,
                                                  const int* __restrict__ cell_start,
                                                  uint32_t cell, int i, Load&& load, Use&& use) {
  static_assert(B % NBR_GROUP == 0, "batch must be whole groups");
  constexpr int G = B / NBR_GROUP;
  const int c = L.cnt[i];
  if (c <= L.cap) {
    if (c <= 0) return;
    const NbrBase nb = nbr_bases(k, cell_start, cell);
    const uint4* __restrict__ p = L.idx + nbr_gslot(i, 0, L.cap / NBR_GROUP);
    const int glast = (c - 1) / NBR_GROUP;
    uint4 g[G];
    if (PREFETCH) {
#pragma unroll
      for (int u = 0; u < G; ++u) g[u] = p[(size_t)min(u, glast) * NBR_SLICE];
    }
    using Rec = decltype(load(0));
    for (int q0 = 0; q0 < c; q0 += B) {
      if (!PREFETCH) {
#pragma unroll
        for (int u = 0; u < G; ++u) g[u] = p[(size_t)min(q0 / NBR_GROUP + u, glast) * NBR_SLICE];
      }
      int j[B];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const uint32_t w[4] = {g[u].x, g[u].y, g[u].z, g[u].w};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          j[u * NBR_GROUP + 2 * h] = nbr_decode(w[h] & 0xFFFFu, nb);
          j[u * NBR_GROUP + 2 * h + 1] = nbr_decode(w[h] >> 16, nb);
        }
      }
      if (PREFETCH) {
        const int gn = q0 / NBR_GROUP + G;
#pragma unroll
        for (int u = 0; u < G; ++u) g[u] = p[(size_t)min(gn + u, glast) * NBR_SLICE];
      }
      Rec v[B];
#pragma unroll
      for (int b = 0; b < B; ++b) v[b] = load(j[b]);
#pragma unroll
      for (int b = 0; b < B; ++b)
        if (q0 + b < c) use(j[b], v[b]);
    }
  } else {
    for_each_neighbor(k, cell_start, cell, [&](int j) {
      if (j != i) use(j, load(j));
    });
  }
}

// A record of `base` at a 32-bit byte offset: the address is the base (a
// kernel argument, in SGPRs) plus the zero-extended offset, which the
// global_load saddr form adds in the address unit, so a gather costs one
// shift per lane instead of 64-bit address arithmetic.  Every array the
// solve passes gather is below 4 GiB.
template <class T>
__device__ __forceinline__ T load_at(const T* __restrict__ base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Branch-free traversal for the pressure-solve passes: as
// for_each_listed_b, but use(j, rec, valid) is called for every slot of
// every batch, with valid false for the slots past the particle's count,
// and the loop body has no per-neighbour branch, so the compiler
// interleaves a batch's arithmetic.  use must add nothing for an invalid
// slot (its record is a real particle's, at S_0).  j is the sorted index.
template <int B, bool PREFETCH, class Load, class Use>
__device__ __forceinline__ void for_each_listed_m(const KP& k, const NbrList& L,
                                                  const int* __restrict__ cell_start,
                                                  uint32_t cell, int i, Load&& load, Use&& use) {
  static_assert(B % NBR_GROUP == 0, "batch must be whole groups");
  constexpr int G = B / NBR_GROUP;
  const int c = L.cnt[i];
  if (c <= L.cap) {
    if (c <= 0) return;
    const NbrBase nb = nbr_bases(k, cell_start, cell);
    const uint4* __restrict__ p = L.idx + nbr_gslot(i, 0, L.cap / NBR_GROUP);
    const int glast = (c - 1) / NBR_GROUP;
    using Rec = decltype(load(0u));
    uint4 g[G];
    if (PREFETCH) {
#pragma unroll
      for (int u = 0; u < G; ++u) g[u] = p[(size_t)min(u, glast) * NBR_SLICE];
    }
    for (int q0 = 0; q0 < c; q0 += B) {
      if (!PREFETCH) {
#pragma unroll
        for (int u = 0; u < G; ++u) g[u] = p[(size_t)min(q0 / NBR_GROUP + u, glast) * NBR_SLICE];
      }
      uint32_t j[B];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const uint32_t w[4] = {g[u].x, g[u].y, g[u].z, g[u].w};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          j[u * NBR_GROUP + 2 * h] = (uint32_t)nbr_decode(w[h] & 0xFFFFu, nb);
          j[u * NBR_GROUP + 2 * h + 1] = (uint32_t)nbr_decode(w[h] >> 16, nb);
        }
      }
      if (PREFETCH) {
        const int gn = q0 / NBR_GROUP + G;
#pragma unroll
        for (int u = 0; u < G; ++u) g[u] = p[(size_t)min(gn + u, glast) * NBR_SLICE];
      }
      Rec v[B];
#pragma unroll
      for (int b = 0; b < B; ++b) v[b] = load(j[b]);
#pragma unroll
      for (int b = 0; b < B; ++b) use(j[b], v[b], q0 + b < c);
    }
  } else {
    for_each_neighbor(k, cell_start, cell, [&](int j) {
      if (j != i) use((uint32_t)j, load((uint32_t)j), true);
    });
  }
}

// Cubic-spline kernel gradient for branch-free neighbour loops, in the
// units of the stored positions: with x_u = x_i - x_j before the simscale
// s, the reference's grad W_ij = x_u s g(q) kspline / (h r), r = s |x_u|,
// q = r / h (fluid_system_kern.cu:637-671), is x_u g(q) kspline / (h |x_u|):
// s cancels except in q.  g is -3q + 2.25q^2 (q <= 1) or -0.75(2 - q)^2,
// with kspline / h folded into the coefficients; the support test
// 0 < |x_u|^2 < lim / s^2 is one unsigned compare of the bit patterns
// (|x_u|^2 >= 0; NaN and +inf fail it).  Agrees with cubic_gradf to rounding.
struct SolveGrad {
  float q_of_r;       // s / h
  float co, ca, cb;   // -0.75, -3, 2.25 times kspline / h
  uint32_t lim_m1;    // bits of lim / s^2, minus 1
};

// support 2h when `wide`, else h
__device__ __forceinline__ SolveGrad solve_grad(const KP& k, bool wide) {
  const float ks = k.kspline * k.inv_h;
  const float lim = (wide ? 4.f * k.h2 : k.h2) / (k.ss * k.ss);
  return SolveGrad{k.ss * k.inv_h, -0.75f * ks, -3.f * ks, 2.25f * ks, __float_as_uint(lim) - 1u};
}
// the support of the IISPH loops (SolverOptions::full_support)
__device__ __forceinline__ SolveGrad solve_grad(const KP& k) {
  return solve_grad(k, k.full_support != 0);
}

// |x_u|^2 with the FMAs written out (the same bits in every kernel).
__device__ __forceinline__ float pair_dsq(float dx, float dy, float dz) {
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// f with grad W = x_u f; 0 outside the support (dsq_u = +inf: an unused slot).
// No contraction inside (FMAs only where written), so every kernel that
// inlines it computes the same bits.
__device__ __forceinline__ float solve_gradf(const SolveGrad& s, float dsq_u) {
#pragma clang fp contract(off)
  const float ir = __builtin_amdgcn_rsqf(dsq_u);
  const float q = dsq_u * ir * s.q_of_r;
  const float t = 2.f - q;
  const float outer = s.co * (t * t);
  const float inner = q * fmaf(s.cb, q, s.ca);
  const float g = (q > 1.f ? outer : inner) * ir;
  return (__float_as_uint(dsq_u) - 1u) < s.lim_m1 ? g : 0.f;
}

// Writer side of the packed list: entries are shifted in at the top of a
// 128-bit register, so after eight pushes the first is the lowest 16 bits,
// and each full group is stored as one 16-byte write.
struct NbrPacker {
  uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
  __device__ __forceinline__ void push(uint32_t e) {
    w0 = __builtin_amdgcn_alignbit(w1, w0, 16);
    w1 = __builtin_amdgcn_alignbit(w2, w1, 16);
    w2 = __builtin_amdgcn_alignbit(w3, w2, 16);
    w3 = __builtin_amdgcn_alignbit(e, w3, 16);
  }
  __device__ __forceinline__ uint4 get() const { return make_uint4(w0, w1, w2, w3); }
};

// r = |x| and 1/r from r^2 (v_rsq_f32, about 1 ulp; r = 0 for r^2 = 0).
__device__ __forceinline__ void dist_rinv(float dsq, float& r, float& ir) {
  ir = __builtin_amdgcn_rsqf(dsq);
  r = dsq > 0.f ? dsq * ir : 0.f;
}

}  // namespace sphx
