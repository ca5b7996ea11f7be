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

// Per-step neighbour list in sliced-ELL layout: particles are grouped in
// slices of 256; entry q of particle i is idx[(slice * cap + q) * 256 +
// (i & 255)], so a wave reading entry q of its 64 particles touches one
// 256-byte run and a slice's whole list is one contiguous cap KiB block.
// Entries follow the reference's visiting order and hold every j != i with
// |x_ij| < 2h (the widest support any kernel of the step uses).  Positions
// do not move between the density pass and the last pressure iteration, so
// one list serves every neighbour loop of an IISPH step.  A particle with
// more than `cap` neighbours keeps its true count and its loops fall back
// to the cell scan, which visits a superset in the same order.
constexpr int NBR_SLICE = 256;

__host__ __device__ __forceinline__ size_t nbr_slot(int i, int q, int cap) {
  return ((size_t)(i / NBR_SLICE) * cap + q) * NBR_SLICE + (i % NBR_SLICE);
}

struct NbrList {
  const int* __restrict__ idx;
  const int* __restrict__ cnt;
  int cap;
};

template <class F>
__device__ __forceinline__ void for_each_listed(const KP& k, const NbrList& L,
                                                const int* __restrict__ cell_start, uint32_t cell,
                                                int i, F&& f) {
  const int c = L.cnt[i];
  if (c <= L.cap) {
    const int* __restrict__ p = L.idx + nbr_slot(i, 0, L.cap);
#pragma unroll 4
    for (int q = 0; q < c; ++q) f(p[(size_t)q * NBR_SLICE]);
  } else {
    for_each_neighbor(k, cell_start, cell, [&](int j) {
      if (j != i) f(j);
    });
  }
}

// Batched traversal of the list: B entries at a time, and the
// records of all of them fetched (load(j)) before any is used (use(j, rec)),
// so a batch's gathers are in flight together instead of one dependent
// round trip per neighbour.  Entries past the count reload the last entry
// (inside the particle's own list) and are not used; the visiting order,
// and so every sum's order, is that of for_each_listed.
constexpr int NBR_BATCH = 8;    // default batch (wide records)
constexpr int NBR_BATCH_SMALL = 16;   // loops with few live values (density, DP)

template <int B = NBR_BATCH, class Load, class Use>
__device__ __forceinline__ void for_each_listed_b(const KP& k, const NbrList& L,
                                                  const int* __restrict__ cell_start,
                                                  uint32_t cell, int i, Load&& load, Use&& use) {
  const int c = L.cnt[i];
  if (c <= L.cap) {
    const int* __restrict__ p = L.idx + nbr_slot(i, 0, L.cap);
    using Rec = decltype(load(0));
    for (int q0 = 0; q0 < c; q0 += B) {
      int j[B];
      Rec v[B];
#pragma unroll
      for (int b = 0; b < B; ++b) j[b] = p[(size_t)min(q0 + b, c - 1) * NBR_SLICE];
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

// r = |x| and 1/r from r^2 (v_rsq_f32, about 1 ulp; r = 0 for r^2 = 0).
__device__ __forceinline__ void dist_rinv(float dsq, float& r, float& ir) {
  ir = __builtin_amdgcn_rsqf(dsq);
  r = dsq > 0.f ? dsq * ir : 0.f;
}

}  // namespace sphx
