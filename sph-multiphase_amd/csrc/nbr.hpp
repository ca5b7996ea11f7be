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

}  // namespace sphx
