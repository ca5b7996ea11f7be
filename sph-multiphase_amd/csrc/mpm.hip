// MPM transfer with Drucker-Prager plasticity on a dense node grid.
//
// Reference (SPH_Solver/fluids/MpmSolver.cu): node-centric gathers over the
// 27 SPH cells around each node (P2G :823-889, node force :658-704) and
// particle-centric gathers over 4^3 nodes (stress :477-592, G2P :891-951),
// one thread per node or particle.  Here every transfer is particle-centric
// over the cell-sorted particles:
//   * P2G and the node force run in two passes.  A workgroup of 256
//     consecutive sorted particles bins them by stencil base cell in LDS
//     (one int atomic each) and accumulates into an LDS tile covering the
//     bounding box of their stencils: the W^3 stencil offsets are phases,
//     in which distinct cells write distinct nodes and the particles of a
//     cell add in rank order (no float atomics).  The tile is stored whole.  A per-node-block pass sums
//     the tiles of the workgroups whose sorted particle range can reach the
//     block (found through the sort grid's cell_start).  Particles of a
//     box too large for a tile add to the nodes directly with atomics.
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
  if (flags[b] == 0 && atomicCAS(&flags[b], 0, 1) == 0) list[atomicAdd(&count[0], 1)] = b;
}

// Where a scatter leaves its sums.  Workgroup w (sorted particles
// [256 w, 256 w + 256)) writes its node sums, box by box, to
// tiles[w * MPM_TILE ...] and the boxes to hdr[3 w ..]:
//   hdr[3w]   = {lo0.x, lo0.y, lo0.z, ex0.x}
//   hdr[3w+1] = {ex0.y, ex0.z, lo1.x, lo1.y}
//   hdr[3w+2] = {lo1.z, ex1.x, ex1.y, ex1.z}
// (extent 0: no tile for that box).  Tile entries are y-major within a box,
// ((y - lo.y) ex.z + (z - lo.z)) ex.x + (x - lo.x); box 1 follows box 0.
// Particles of a box that does not fit add to `direct` with atomics and
// set *direct_flag.
struct TileOut {
  float4* tiles;
  int4* hdr;
  float4* direct;
  int* direct_flag;
  int* flags;
  int* list;
  int* count;
};

template <int W>
__device__ __forceinline__ void stencil_base(const MP& m, float4 p, int base[3]) {
  const float pc[3] = {p.x, p.y, p.z};
  const float xm[3] = {m.xmin.x, m.xmin.y, m.xmin.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float rel = (pc[a] - xm[a]) / m.dx;
    base[a] = W == 4 ? (int)rel - 1 : (int)floorf(rel - 0.5f);
  }
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

constexpr int MPM_CELLS = 1536;   // base cells binned per workgroup

// LDS of a workgroup's binning.
struct BinShared {
  int sb[13];                   // [box][lo x y z, hi x y z] of bases, first row
  int cnt[MPM_CELLS];
  int start[MPM_CELLS + 1];
  int part[MPM_THREADS / 64];
};

// The workgroup's two boxes (first sort-grid row, the rest): base cells
// lo .. lo + nb - 1 per axis, tile nodes lo .. lo + ex - 1 (ex = nb + W - 1;
// nb = ex = 0 for a box not tiled), and where each box's tile entries and
// cell bins start.
struct BinBox {
  int lo[2][3], nb[2][3], ex[2][3];
  int off[2], coff[2];
  int tot, nbin;   // tile entries of the tiled boxes; binned particles
  // element k of box g (g varies by lane: a select, not an indexed load)
  __device__ __forceinline__ static int get(const int (&a)[2][3], int g, int k) {
    return g ? a[1][k] : a[0][k];
  }
  __device__ __forceinline__ bool tiled(int g) const { return get(nb, g, 0) > 0; }
};

// Bins the valid particles of the workgroup by base cell.  The particles
// split into the first sort-grid row (y, z) they occupy and the rest, so a
// workgroup that wraps from the end of one row to the start of the next
// keeps two compact boxes.  A box is tiled when its nodes fit MPM_TILE and
// its cells MPM_CELLS; seg is the lane's box, slot its place in the binned
// order and cell its bin (-1: not binned).  S.start[c] .. S.start[c + 1] are the slots of
// cell c; cells of one x-row are consecutive.  Must be called by the whole
// workgroup (it synchronises).
template <int W>
__device__ __forceinline__ void block_bins(BinShared& S, bool valid, int row, const int base[3], BinBox& B,
                           int& seg, int& slot, int& cell) {
  const int t0 = threadIdx.x, lane = t0 & 63, wv = t0 >> 6;
  if (t0 < 13) S.sb[t0] = (t0 % 6 < 3 || t0 == 12) ? INT_MAX : INT_MIN;
  __syncthreads();
  const int r0 = wave_min(valid ? row : INT_MAX);
  if (lane == 0 && r0 != INT_MAX) atomicMin(&S.sb[12], r0);
  __syncthreads();
  seg = valid && row != S.sb[12] ? 1 : 0;
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const bool in = valid && seg == g;
      const int mn = wave_min(in ? base[a] : INT_MAX);
      const int mx = wave_max(in ? base[a] : INT_MIN);
      if (lane == 0 && mn != INT_MAX) {
        atomicMin(&S.sb[6 * g + a], mn);
        atomicMax(&S.sb[6 * g + 3 + a], mx);
      }
    }
  __syncthreads();
  long long V[2], C[2];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const bool has = S.sb[6 * g] != INT_MAX;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      B.lo[g][a] = S.sb[6 * g + a];
      B.nb[g][a] = has ? S.sb[6 * g + 3 + a] - B.lo[g][a] + 1 : 0;
      B.ex[g][a] = has ? B.nb[g][a] + W - 1 : 0;
    }
    V[g] = (long long)B.ex[g][0] * B.ex[g][1] * B.ex[g][2];
    C[g] = (long long)B.nb[g][0] * B.nb[g][1] * B.nb[g][2];
  }
  const bool tiled0 = V[0] > 0 && V[0] <= MPM_TILE && C[0] <= MPM_CELLS;
  B.off[0] = 0;
  B.coff[0] = 0;
  B.off[1] = tiled0 ? (int)V[0] : 0;
  B.coff[1] = tiled0 ? (int)C[0] : 0;
  const bool tiled1 = V[1] > 0 && B.off[1] + V[1] <= MPM_TILE && B.coff[1] + C[1] <= MPM_CELLS;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!tiled0) B.nb[0][a] = B.ex[0][a] = 0;
    if (!tiled1) B.nb[1][a] = B.ex[1][a] = 0;
  }
  const int ctot = B.coff[1] + (tiled1 ? (int)C[1] : 0);
  B.tot = B.off[1] + (tiled1 ? (int)V[1] : 0);
  for (int c = t0; c < ctot; c += blockDim.x) S.cnt[c] = 0;
  __syncthreads();
  int cb = -1, rank = 0;
  if (valid && B.tiled(seg)) {
    const int l0 = B.get(B.lo, seg, 0), l1 = B.get(B.lo, seg, 1), l2 = B.get(B.lo, seg, 2);
    const int n0 = B.get(B.nb, seg, 0), n2 = B.get(B.nb, seg, 2);
    cb = (seg ? B.coff[1] : 0) + ((base[1] - l1) * n2 + (base[2] - l2)) * n0 + (base[0] - l0);
    rank = atomicAdd(&S.cnt[cb], 1);
  }
  __syncthreads();
  // exclusive scan of cnt[0 .. ctot) into start[0 .. ctot]
  const int per = (ctot + (int)blockDim.x - 1) / (int)blockDim.x;
  const int c0 = min(t0 * per, ctot), c1 = min(c0 + per, ctot);
  int local = 0;
  for (int c = c0; c < c1; ++c) local += S.cnt[c];
  int incl = local;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int v = __shfl_up(incl, o);
    if (lane >= o) incl += v;
  }
  if (lane == 63) S.part[wv] = incl;
  __syncthreads();
  int run = incl - local;
  for (int k = 0; k < wv; ++k) run += S.part[k];
  for (int c = c0; c < c1; ++c) {
    S.start[c] = run;
    run += S.cnt[c];
  }
  if (t0 == (int)blockDim.x - 1) S.start[ctot] = run;
  __syncthreads();
  slot = cb >= 0 ? S.start[cb] + rank : -1;
  cell = cb;
  B.nbin = S.start[ctot];
}

// LDS of the phased tile accumulation: the tile of the workgroup's boxes,
// and the particle and bin of each binned slot.
struct PhaseShared {
  float4 tile[MPM_TILE];
  int pidx[MPM_THREADS];
  int pcell[MPM_THREADS];
};

// The binned slot a thread accumulates: its particle, the cell's place in
// its box (local coordinates), the tile entry of the cell's node (0, 0, 0)
// and the box extents, the particle's rank within its cell, and whether the
// cell's slots straddle two waves.
struct SlotRef {
  bool act, cross;
  int j, rank, entry0, ex0, exz;
};

__device__ __forceinline__ SlotRef slot_ref(const BinShared& S, const PhaseShared& T, const BinBox& B) {
  SlotRef r = {false, false, 0, 0, 0, 0, 0};
  const int t = threadIdx.x;
  if (t >= B.nbin) return r;
  r.act = true;
  r.j = T.pidx[t];
  const int c = T.pcell[t];
  const int s0 = S.start[c], s1 = S.start[c + 1];
  r.rank = t - s0;
  r.cross = (s0 >> 6) != ((s1 - 1) >> 6);
  // box of the cell: box 0 holds cells [0, coff1) when it is tiled
  const int g = (B.tiled(0) && c < B.coff[1]) ? 0 : 1;
  const int cc = c - (g ? B.coff[1] : 0);
  const int n0 = BinBox::get(B.nb, g, 0), n2 = BinBox::get(B.nb, g, 2);
  const int e0 = BinBox::get(B.ex, g, 0), e2 = BinBox::get(B.ex, g, 2);
  const int cx = cc % n0, cz = (cc / n0) % n2, cy = cc / (n0 * n2);
  r.ex0 = e0;
  r.exz = e0 * e2;
  r.entry0 = (g ? B.off[1] : 0) + (cy * e2 + cz) * e0 + cx;
  return r;
}

// Accumulates contrib(a, b, c) of every binned slot into the LDS tile
// without float atomics.  The stencil offsets (a, b, c) are phases
// separated by barriers; within a phase distinct cells write distinct
// nodes, and the particles of one cell (consecutive slots, one wave) add in
// rank order, each rank a separate read-modify-write of the wave.  Cells
// whose slots straddle two waves add with LDS atomics.  Called by the whole
// workgroup.
template <int W, class F>
__device__ __forceinline__ void phased_tile(PhaseShared& T, const SlotRef& r, F&& contrib) {
  const int maxr = wave_max(r.act && !r.cross ? r.rank : -1);
#pragma clang loop unroll(full)
  for (int b = 0; b < W; ++b)
#pragma clang loop unroll(full)
    for (int c = 0; c < W; ++c)
#pragma clang loop unroll(full)
      for (int a = 0; a < W; ++a) {
        const float4 v = r.act ? contrib(a, b, c) : make_float4(0.f, 0.f, 0.f, 0.f);
        float4* d = &T.tile[r.entry0 + b * r.exz + c * r.ex0 + a];
        if (r.act && r.cross) {
          atomicAdd(&d->x, v.x);
          atomicAdd(&d->y, v.y);
          atomicAdd(&d->z, v.z);
          atomicAdd(&d->w, v.w);
        }
        for (int k = 0; k <= maxr; ++k) {
          if (r.act && !r.cross && r.rank == k) {
            float4 o = *d;
            o.x += v.x; o.y += v.y; o.z += v.z; o.w += v.w;
            *d = o;
          }
          // each rank's update reaches LDS before the next rank reads it
          asm volatile("" ::: "memory");
        }
        __syncthreads();
      }
}

// Header of workgroup w and the node blocks of its tiled boxes (clamped to
// the grid) marked active.
template <int W>
__device__ __forceinline__ void tile_finish(const MP& m, const TileOut& o, int w, const BinBox& B) {
  const int t0 = threadIdx.x;
  if (t0 == 0) {
    o.hdr[3 * (size_t)w] = make_int4(B.lo[0][0], B.lo[0][1], B.lo[0][2], B.ex[0][0]);
    o.hdr[3 * (size_t)w + 1] = make_int4(B.ex[0][1], B.ex[0][2], B.lo[1][0], B.lo[1][1]);
    o.hdr[3 * (size_t)w + 2] = make_int4(B.lo[1][2], B.ex[1][0], B.ex[1][1], B.ex[1][2]);
  }
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    if (B.ex[g][0] == 0) continue;
    int b0[3], nb[3];
    for (int a = 0; a < 3; ++a) {
      int r = a == 0 ? m.res.x : (a == 1 ? m.res.y : m.res.z);
      int lw = max(B.lo[g][a], 0), hg = min(B.lo[g][a] + B.ex[g][a] - 1, r - 1);
      b0[a] = lw / MPM_BLK;
      nb[a] = hg >= lw ? hg / MPM_BLK - b0[a] + 1 : 0;
    }
    int nt = nb[0] * nb[1] * nb[2];
    for (int t = t0; t < nt; t += blockDim.x) {
      int bx = b0[0] + t % nb[0], bz = b0[2] + (t / nb[0]) % nb[2], by = b0[1] + t / (nb[0] * nb[2]);
      mark_block(o.flags, o.list, o.count, block_id(m, bx, by, bz));
    }
  }
}

// A particle of a box that is not tiled adds contrib(a, b, c) to node
// base + (a, b, c) of its stencil in global memory and marks the blocks its
// stencil spans.
template <int W, class F>
__device__ __forceinline__ void direct_add(const MP& m, const TileOut& o, const int base[3], F&& contrib) {
#pragma unroll
  for (int b = 0; b < W; ++b)
#pragma unroll
    for (int c = 0; c < W; ++c)
#pragma unroll
      for (int a = 0; a < W; ++a) {
        int x = base[0] + a, y = base[1] + b, z = base[2] + c;
        if (!in_res(m, x, y, z)) continue;
        float4 v = contrib(a, b, c);
        float4* d = o.direct + node_id(m, x, y, z);
        atomicAdd(&d->x, v.x);
        atomicAdd(&d->y, v.y);
        atomicAdd(&d->z, v.z);
        atomicAdd(&d->w, v.w);
      }
  o.direct_flag[0] = 1;
  for (int b = 0; b < 2; ++b)
    for (int c = 0; c < 2; ++c)
      for (int a = 0; a < 2; ++a) {
        int x = min(max(base[0] + a * (W - 1), 0), m.res.x - 1);
        int y = min(max(base[1] + b * (W - 1), 0), m.res.y - 1);
        int z = min(max(base[2] + c * (W - 1), 0), m.res.z - 1);
        mark_block(o.flags, o.list, o.count, block_id(m, x / MPM_BLK, y / MPM_BLK, z / MPM_BLK));
      }
}

// The sort grid as the gather needs it: which sorted particles lie in
// which cell (cell_start), and the map from a position to a cell (the
// hash's (p - gmin) * fac, truncated; valid cells 1..res-2 per axis).
struct SortGrid {
  const int* cell_start;
  int3 res;
  float3 gmin, fac;
};

constexpr int GATHER_CAND = 512;   // candidate workgroups listed at a time
constexpr int GATHER_BOX = 256;    // intersecting boxes held at a time

// LDS of a one-wave gather workgroup.
struct GatherShared {
  int cand[GATHER_CAND];
  int4 box[GATHER_BOX][2];   // {lo x, y, z, ex x}, {ex y, ex z, tile entry 0, -}
};

// Adds to `sum`, in list order, the tile entry of this lane's node (x, y,
// z) from each of the first nbox listed boxes; four loads in flight.
__device__ __forceinline__ void sum_boxes(const GatherShared& G, int nbox, const float4* __restrict__ tiles,
                                          int x, int y, int z, float4& sum) {
  for (int k = 0; k < nbox; k += 4) {
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (k + u < nbox) {
        const int4 a = G.box[k + u][0], b = G.box[k + u][1];
        const int dx = x - a.x, dy = y - a.y, dz = z - a.z;
        if (dx >= 0 && dx < a.w && dy >= 0 && dy < b.x && dz >= 0 && dz < b.y)
          v[u] = tiles[(size_t)b.z + (dy * b.y + dz) * a.w + dx];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sum.x += v[u].x; sum.y += v[u].y; sum.z += v[u].z; sum.w += v[u].w;
    }
  }
}

// Sum, for node (x, y, z) of node block (bx, by, bz), the tile entries of
// every workgroup whose particles can reach the block.  Those particles lie
// within R cells of the block (R = support + 1/2 cell of slack, far above
// the hash's rounding), so in the sort-grid cells that range maps to; their
// sorted ranges name the workgroups.  The wave works lane-parallel: one
// sort-grid row per lane gives a workgroup range (rows in key order, so a
// running maximum removes the workgroups already listed), the candidates'
// headers are loaded one per lane, and the boxes that meet the block are
// listed in workgroup order and summed.  Called by the whole (one-wave)
// workgroup.
template <int W>
__device__ float4 gather_tiles(const MP& m, const SortGrid& g, const int4* __restrict__ hdr,
                               const float4* __restrict__ tiles, GatherShared& G, int bx, int by,
                               int bz, int x, int y, int z) {
  const int R = W == 4 ? 3 : 2;
  const int lane = threadIdx.x & 63;
  const int bb[3] = {bx, by, bz};
  const float xm[3] = {m.xmin.x, m.xmin.y, m.xmin.z};
  const float gm[3] = {g.gmin.x, g.gmin.y, g.gmin.z};
  const float fc[3] = {g.fac.x, g.fac.y, g.fac.z};
  const int rs[3] = {g.res.x, g.res.y, g.res.z};
  int clo[3], chi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float lo = xm[a] + (float)(bb[a] * MPM_BLK - R) * m.dx;
    float hi = xm[a] + (float)(bb[a] * MPM_BLK + MPM_BLK - 1 + R) * m.dx;
    clo[a] = max((int)floorf((lo - gm[a]) * fc[a]), 1);
    chi[a] = min((int)floorf((hi - gm[a]) * fc[a]), rs[a] - 2);
  }
  float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
  if (clo[0] > chi[0] || clo[1] > chi[1] || clo[2] > chi[2]) return sum;
  const int nz = chi[2] - clo[2] + 1;
  const int nrows = (chi[1] - clo[1] + 1) * nz;
  const int nlo[3] = {bx * MPM_BLK, by * MPM_BLK, bz * MPM_BLK};
  int last = -1, nbox = 0;
  for (int r0 = 0; r0 < nrows; r0 += 64) {
    const int r = r0 + lane;
    int w0 = 0, w1 = -1;
    if (r < nrows) {
      const int cy = clo[1] + r / nz, cz = clo[2] + r % nz;
      const int rowbase = (cy * g.res.z + cz) * g.res.x;
      const int s0 = g.cell_start[rowbase + clo[0]], s1 = g.cell_start[rowbase + chi[0] + 1];
      if (s1 > s0) {
        w0 = s0 / MPM_THREADS;
        w1 = (s1 - 1) / MPM_THREADS;
      }
    }
    int pm = w1;   // running maximum of the last workgroup listed
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(pm, o);
      if (lane >= o) pm = max(pm, v);
    }
    int prev = __shfl_up(pm, 1);
    if (lane == 0) prev = -1;
    w0 = max(w0, max(prev, last) + 1);
    const int cnt = w1 >= w0 ? w1 - w0 + 1 : 0;
    int inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(inc, o);
      if (lane >= o) inc += v;
    }
    const int off = inc - cnt, total = __shfl(inc, 63);
    last = max(last, __shfl(pm, 63));
    for (int cb = 0; cb < total; cb += GATHER_CAND) {
      const int j0 = max(off, cb), j1 = min(off + cnt, cb + GATHER_CAND);
      for (int j = j0; j < j1; ++j) G.cand[j - cb] = w0 + (j - off);
      __syncthreads();
      const int nc = min(total - cb, GATHER_CAND);
      for (int c0 = 0; c0 < nc; c0 += 64) {
        const int c = c0 + lane;
        bool hit[2] = {false, false};
        int4 e[2][2];
        if (c < nc) {
          const int w = G.cand[c];
          const int4 h0 = hdr[3 * (size_t)w], h1 = hdr[3 * (size_t)w + 1], h2 = hdr[3 * (size_t)w + 2];
          const int lo0[3] = {h0.x, h0.y, h0.z}, ex0[3] = {h0.w, h1.x, h1.y};
          const int lo1[3] = {h1.z, h1.w, h2.x}, ex1[3] = {h2.y, h2.z, h2.w};
          hit[0] = ex0[0] > 0;
          hit[1] = ex1[0] > 0;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            hit[0] = hit[0] && lo0[a] <= nlo[a] + MPM_BLK - 1 && lo0[a] + ex0[a] > nlo[a];
            hit[1] = hit[1] && lo1[a] <= nlo[a] + MPM_BLK - 1 && lo1[a] + ex1[a] > nlo[a];
          }
          const int base = w * MPM_TILE;
          e[0][0] = make_int4(lo0[0], lo0[1], lo0[2], ex0[0]);
          e[0][1] = make_int4(ex0[1], ex0[2], base, 0);
          e[1][0] = make_int4(lo1[0], lo1[1], lo1[2], ex1[0]);
          e[1][1] = make_int4(ex1[1], ex1[2], base + ex0[0] * ex0[1] * ex0[2], 0);
        }
        const uint64_t b0 = __ballot(hit[0]), b1 = __ballot(hit[1]);
        const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int rel = __popcll(b0 & below) + __popcll(b1 & below);
        const int add = __popcll(b0) + __popcll(b1);   // at most 128
        if (nbox + add > GATHER_BOX) {   // full: sum what is listed first
          __syncthreads();
          sum_boxes(G, nbox, tiles, x, y, z, sum);
          __syncthreads();
          nbox = 0;
        }
        const int p0 = nbox + rel;
        if (hit[0]) { G.box[p0][0] = e[0][0]; G.box[p0][1] = e[0][1]; }
        if (hit[1]) { G.box[p0 + hit[0]][0] = e[1][0]; G.box[p0 + hit[0]][1] = e[1][1]; }
        nbox += add;
      }
      __syncthreads();
    }
  }
  __syncthreads();
  sum_boxes(G, nbox, tiles, x, y, z, sum);
  __syncthreads();
  return sum;
}

// ------------------------------------------------------------------ P2G --
// MpmParticleToGrid_APIC (:823-889) / MpmParticleToGrid (:197-236):
// node {sum w m_p (v_p + B_p iDp (x_I - x_p) ss), sum w m_p}.
template <int W>
__global__ __launch_bounds__(MPM_THREADS) void k_mpm_p2g(
    MP m, int apic, const uint32_t* __restrict__ key, const float4* __restrict__ pos,
    const float4* __restrict__ vel, const float4* __restrict__ phys, const float* __restrict__ B,
    int cap, int sort_resx, TileOut o) {
  __shared__ BinShared S;
  __shared__ PhaseShared T;
  const int i = xcd_index();
  bool valid = false;
  int row = 0, base[3] = {0, 0, 0};
  if (i < m.n) {
    const float4 p = pos[i];
    const uint32_t k = key[i];
    valid = type_of(p) != SPHX_TYPE_BOUNDARY && k != m.sentinel;
    row = (int)(k / (uint32_t)sort_resx);
    stencil_base<W>(m, p, base);
  }
  BinBox bx;
  int seg, slot, cell;
  block_bins<W>(S, valid, row, base, bx, seg, slot, cell);
  if (slot >= 0) {
    T.pidx[slot] = i;
    T.pcell[slot] = cell;
  }
  for (int t = threadIdx.x; t < bx.tot; t += blockDim.x) T.tile[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();
  // the particle of this thread's slot (or, for a box not tiled, its own)
  const SlotRef r = slot_ref(S, T, bx);
  const bool direct = valid && !bx.tiled(seg);
  const int j = r.act ? r.j : i;
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f), v = p;
  float mp = 0.f, Bi[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (r.act || direct) {
    p = pos[j];
    v = vel[j];
    mp = phys[j].x;
    if (apic)
#pragma unroll
      for (int q = 0; q < 9; ++q) Bi[q] = B[(size_t)q * cap + j] * m.iDp;
  }
  Stencil<W> s;
  make_stencil<W>(m, p, s);
  auto contrib = [&](int a, int b, int c) {
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
  };
  phased_tile<W>(T, r, contrib);
  const int w = xcd_tile();
  float4* dst = o.tiles + (size_t)w * MPM_TILE;
  for (int t = threadIdx.x; t < bx.tot; t += blockDim.x) dst[t] = T.tile[t];
  if (direct) {
    // this thread's own particle (slot threads above reloaded theirs)
    if (r.act) {
      p = pos[i];
      v = vel[i];
      mp = phys[i].x;
      if (apic)
#pragma unroll
        for (int q = 0; q < 9; ++q) Bi[q] = B[(size_t)q * cap + i] * m.iDp;
      make_stencil<W>(m, p, s);
    }
    direct_add<W>(m, o, s.base, contrib);
  }
  tile_finish<W>(m, o, w, bx);
}

// Persistent loop over the listed node blocks: 64 threads, one node each,
// nodes outside the grid included (the body checks in_res).
#define FOR_LISTED_NODES(m, list, count, body)                                  \
  for (int bi = blockIdx.x; bi < *(count); bi += gridDim.x) {                      \
    const int blk = (list)[bi];                                                     \
    const int bx = blk % (m).nblk.x, bz = (blk / (m).nblk.x) % (m).nblk.z,          \
              by = blk / ((m).nblk.x * (m).nblk.z);                                 \
    const int t = threadIdx.x;                                                      \
    const int x = bx * MPM_BLK + (t & 3), z = bz * MPM_BLK + ((t >> 2) & 3),        \
              y = by * MPM_BLK + (t >> 4);                                          \
    body                                                                            \
  }

// Per listed node block (64 threads, one node each): the node's sum over
// the workgroup tiles, plus what was added directly (then cleared), then
// velocity = momentum / mass (or zero below 1e-9, :881-888).
template <int W>
__global__ __launch_bounds__(64) void k_mpm_gather_p2g(MP m, SortGrid g, const int* __restrict__ list,
                                                       const int* __restrict__ count,
                                                       const int4* __restrict__ hdr,
                                                       const float4* __restrict__ tiles,
                                                       float4* __restrict__ acc,
                                                       const int* __restrict__ direct_flag,
                                                       float4* __restrict__ nodev) {
  __shared__ GatherShared G;
  const int any_direct = direct_flag[0];
  FOR_LISTED_NODES(m, list, count, {
    float4 a = gather_tiles<W>(m, g, hdr, tiles, G, bx, by, bz, x, y, z);
    if (in_res(m, x, y, z)) {
      const size_t id = node_id(m, x, y, z);
      if (any_direct) {
        float4 d = acc[id];
        a.x += d.x; a.y += d.y; a.z += d.z; a.w += d.w;
        acc[id] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      nodev[id] = a.w > 0.000000001f ? make_float4(a.x / a.w, a.y / a.w, a.z / a.w, a.w)
                                     : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  })
}

// unlist the listed blocks (their direct-add accumulators are already zero:
// each gather clears what the direct path added)
__global__ __launch_bounds__(64) void k_mpm_clear(const int* __restrict__ list,
                                                  const int* __restrict__ count,
                                                  int* __restrict__ flags) {
  for (int bi = blockIdx.x * 64 + threadIdx.x; bi < *count; bi += gridDim.x * 64) flags[list[bi]] = 0;
}

// count = v; the direct-add flags of P2G and the node force cleared
__global__ void k_mpm_set_count(int* count, int v, int* direct) {
  *count = v;
  direct[0] = 0;
  direct[1] = 0;
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
    const float4* __restrict__ phys, const float* __restrict__ S9, int cap, int sort_resx,
    TileOut o) {
  __shared__ BinShared S;
  __shared__ PhaseShared T;
  const int i = xcd_index();
  bool valid = false;
  int row = 0, base[3] = {0, 0, 0};
  if (i < m.n) {
    const float4 p = pos[i];
    const uint32_t k = key[i];
    valid = type_of(p) != SPHX_TYPE_BOUNDARY && k != m.sentinel;
    row = (int)(k / (uint32_t)sort_resx);
    stencil_base<W>(m, p, base);
  }
  BinBox bx;
  int seg, slot, cell;
  block_bins<W>(S, valid, row, base, bx, seg, slot, cell);
  if (slot >= 0) {
    T.pidx[slot] = i;
    T.pcell[slot] = cell;
  }
  for (int t = threadIdx.x; t < bx.tot; t += blockDim.x) T.tile[t] = make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();
  const SlotRef r = slot_ref(S, T, bx);
  const bool direct = valid && !bx.tiled(seg);
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
  float sig[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, scale = 0.f;
  auto load = [&](int j) {
    p = pos[j];
    const float4 q = phys[j];
    scale = m.physical ? q.x / (q.y * m.rc) : 1.f / q.y;
#pragma unroll
    for (int k9 = 0; k9 < 9; ++k9) sig[k9] = S9[(size_t)k9 * cap + j];
  };
  if (r.act) load(r.j);
  else if (direct) load(i);
  Stencil<W> s;
  make_stencil<W>(m, p, s);
  // the reference differentiates at (x_I - x_p)/dx: the stencil's
  // derivative (at (x_p - x_I)/dx) with the sign flipped
  auto contrib = [&](int a, int b, int c) {
    float n0 = s.w[0][a], n1 = s.w[1][b], n2 = s.w[2][c];
    float nw0 = -s.dw[0][a] * n1 * n2, nw1 = -s.dw[1][b] * n0 * n2, nw2 = -s.dw[2][c] * n0 * n1;
    float f0 = sig[0] * nw0 + sig[1] * nw1 + sig[2] * nw2;
    float f1 = sig[3] * nw0 + sig[4] * nw1 + sig[5] * nw2;
    float f2 = sig[6] * nw0 + sig[7] * nw1 + sig[8] * nw2;
    return make_float4(f0 * scale, f1 * scale, f2 * scale, 0.f);
  };
  phased_tile<W>(T, r, contrib);
  const int w = xcd_tile();
  float4* dst = o.tiles + (size_t)w * MPM_TILE;
  for (int t = threadIdx.x; t < bx.tot; t += blockDim.x) dst[t] = T.tile[t];
  if (direct) {
    if (r.act) {
      load(i);
      make_stencil<W>(m, p, s);
    }
    direct_add<W>(m, o, s.base, contrib);
  }
  tile_finish<W>(m, o, w, bx);
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

// MpmNodeUpdate (:658-704): the node force summed from the workgroup tiles
// (plus what was added directly, then cleared), v += (f + g) dt for nodes
// with mass, collision.
template <int W>
__global__ __launch_bounds__(64) void k_mpm_node(MP m, SortGrid g, const int* __restrict__ list,
                                                 const int* __restrict__ count,
                                                 const int4* __restrict__ hdr,
                                                 const float4* __restrict__ tiles,
                                                 float4* __restrict__ force,
                                                 const int* __restrict__ direct_flag,
                                                 float4* __restrict__ nodev) {
  __shared__ GatherShared G;
  const int any_direct = direct_flag[1];
  FOR_LISTED_NODES(m, list, count, {
    float4 f = gather_tiles<W>(m, g, hdr, tiles, G, bx, by, bz, x, y, z);
    if (in_res(m, x, y, z)) {
      const size_t id = node_id(m, x, y, z);
      if (any_direct) {
        float4 d = force[id];
        f.x += d.x; f.y += d.y; f.z += d.z;
        force[id] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      float4 v = nodev[id];
      if (v.w >= 0.000000001f) {
        float sc = m.physical ? 1.f / v.w : 1.f;
        v.x += (f.x * sc + m.g.x) * m.dt;
        v.y += (f.y * sc + m.g.y) * m.dt;
        v.z += (f.z * sc + m.g.z) * m.dt;
        collide(m, m.xmin.x + (float)x * m.dx, m.xmin.y + (float)y * m.dx,
                m.xmin.z + (float)z * m.dx, v);
        nodev[id] = v;
      }
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

static TileOut tile_out(sphx_context* c, float4* direct, int pass) {
  auto& M = c->mpm;
  return TileOut{M.tiles.p, M.hdr.p, direct, M.direct.p + pass, M.flags.p, M.list.p, M.count.p};
}

static SortGrid sort_grid(const sphx_context* c) {
  KP k = make_kp(c);
  return SortGrid{c->cell_start.p, k.gridres, k.gridmin, k.gridfac};
}

// The scatters index tiles by workgroup: room for every workgroup of a
// full context.
static int mpm_tiles_fit(sphx_context* c, const char* who) {
  SPHX_CHECK((size_t)blocks_for(c->num, MPM_THREADS) <= c->mpm.tiles_cap, who,
             "more particles than the MPM tiles were sized for");
  return 0;
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
  M.tiles_cap = (size_t)blocks_for((int)cap, MPM_THREADS);
  SPHX_CHECK(M.tiles_cap * MPM_TILE < (size_t)INT32_MAX, "sphx_mpm_setup",
             "too many particles for the MPM tile index");
  SPHX_TRY(M.tiles.alloc(M.tiles_cap * MPM_TILE));
  SPHX_TRY(M.hdr.alloc(M.tiles_cap * 3));
  SPHX_TRY(M.direct.alloc(2));
  SPHX_HIP(hipMemsetAsync(M.direct.p, 0, 2 * sizeof(int), c->stream));
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
  SPHX_TRY(mpm_tiles_fit(c, "sphx_mpm_p2g"));
  auto& M = c->mpm;
  MP m = make_mp(c);
  SortGrid g = sort_grid(c);
  TileOut o = tile_out(c, M.acc.p, 0);
  int pb = persistent_blocks(c);
  timer_begin(c, "mpm_clear");
  k_mpm_clear<<<pb, 64, 0, c->stream>>>(M.list.p, M.count.p, M.flags.p);
  k_mpm_set_count<<<1, 1, 0, c->stream>>>(M.count.p, 0, M.direct.p);
  timer_end(c);
  if (m.n > 0) {
    timer_begin(c, "mpm_p2g");
    if (M.kernel == 0)
      k_mpm_p2g<4><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
          m, apic, c->key.p, c->POS(), c->VEL(), c->PHYS(), M.B[c->cur].p, c->cap, g.res.x, o);
    else
      k_mpm_p2g<3><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
          m, apic, c->key.p, c->POS(), c->VEL(), c->PHYS(), M.B[c->cur].p, c->cap, g.res.x, o);
    timer_end(c);
  }
  timer_begin(c, "mpm_gather");
  if (M.kernel == 0)
    k_mpm_gather_p2g<4><<<pb, 64, 0, c->stream>>>(m, g, M.list.p, M.count.p, M.hdr.p, M.tiles.p,
                                                  M.acc.p, M.direct.p, M.nodev.p);
  else
    k_mpm_gather_p2g<3><<<pb, 64, 0, c->stream>>>(m, g, M.list.p, M.count.p, M.hdr.p, M.tiles.p,
                                                  M.acc.p, M.direct.p, M.nodev.p);
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
  SPHX_TRY(mpm_tiles_fit(c, "sphx_mpm_node_update"));
  auto& M = c->mpm;
  MP m = make_mp(c);
  SortGrid g = sort_grid(c);
  TileOut o = tile_out(c, M.force.p, 1);
  if (m.n > 0) {
    timer_begin(c, "mpm_force");
    if (M.kernel == 0)
      k_mpm_force<4><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
          m, c->key.p, c->POS(), c->PHYS(), M.S[c->cur].p, c->cap, g.res.x, o);
    else
      k_mpm_force<3><<<blocks_for(m.n, MPM_THREADS), MPM_THREADS, 0, c->stream>>>(
          m, c->key.p, c->POS(), c->PHYS(), M.S[c->cur].p, c->cap, g.res.x, o);
    timer_end(c);
  }
  timer_begin(c, "mpm_node");
  if (M.kernel == 0)
    k_mpm_node<4><<<persistent_blocks(c), 64, 0, c->stream>>>(m, g, M.list.p, M.count.p, M.hdr.p,
                                                               M.tiles.p, M.force.p, M.direct.p,
                                                               M.nodev.p);
  else
    k_mpm_node<3><<<persistent_blocks(c), 64, 0, c->stream>>>(m, g, M.list.p, M.count.p, M.hdr.p,
                                                               M.tiles.p, M.force.p, M.direct.p,
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
  positions_changed(c);
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
  return timer_collect(c);
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
  k_mpm_clear<<<persistent_blocks(c), 64, 0, c->stream>>>(M.list.p, M.count.p, M.flags.p);
  k_mpm_list_all<<<blocks_for((int)M.blocks, 256), 256, 0, c->stream>>>(M.flags.p, M.list.p,
                                                                        (int)M.blocks);
  k_mpm_set_count<<<1, 1, 0, c->stream>>>(M.count.p, (int)M.blocks, M.direct.p);
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

// ------------------------------------------------ node velocity sampling --
namespace sphx {

// SPH cell of a node as initMpm computes it (MpmSolver.cu:49-61): node
// position x_I = mpmXmin + (x, y, z) mpmcellsize, cell g = (x_I - gridmin)
// gridIdfac truncated, valid when 1 <= g <= res - 1 on every axis (FP32
// operations in the reference's order, no contraction).
__device__ __forceinline__ bool node_cell(const KP& k, const MP& m, int i, float3& xn, int3& g) {
  const int x = i % m.res.x, z = (i / m.res.x) % m.res.z, y = i / (m.res.x * m.res.z);
  xn = make_float3(__fadd_rn(m.xmin.x, __fmul_rn((float)x, m.dx)),
                   __fadd_rn(m.xmin.y, __fmul_rn((float)y, m.dx)),
                   __fadd_rn(m.xmin.z, __fmul_rn((float)z, m.dx)));
  g = make_int3((int)__fmul_rn(__fsub_rn(xn.x, k.gridmin.x), k.gridfac.x),
                (int)__fmul_rn(__fsub_rn(xn.y, k.gridmin.y), k.gridfac.y),
                (int)__fmul_rn(__fsub_rn(xn.z, k.gridmin.z), k.gridfac.z));
  return g.x >= 1 && g.x <= k.gridres.x - 1 && g.y >= 1 && g.y <= k.gridres.y - 1 &&
         g.z >= 1 && g.z <= k.gridres.z - 1;
}

// Visit the members of the 27 cells around cell g in the reference's order
// (rows dy, dz; cells dx ascending; members ascending).  Cells past the
// grid's last layer (g = res - 1 is a valid node cell, A11) are skipped.
template <class F>
__device__ __forceinline__ void node_cells(const KP& k, const int* __restrict__ cs, int3 g, F&& f) {
  const int x0 = g.x - 1, x1 = min(g.x + 1, k.gridres.x - 1);
#pragma unroll 1
  for (int dy = -1; dy <= 1; ++dy) {
    const int y = g.y + dy;
    if (y >= k.gridres.y) continue;
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
      const int z = g.z + dz;
      if (z >= k.gridres.z) continue;
      const int row = (y * k.gridres.z + z) * k.gridres.x;
      const int jb = cs[row + x0], je = cs[row + x1 + 1];
#pragma unroll 1
      for (int j = jb; j < je; ++j) f(j);
    }
  }
}

// MpmGetMomentum (MpmSolver.cu:107-153): node velocity = trilinear-weighted
// mean of the velocities of the particles closer than one node spacing on
// every axis; 0 when the weights sum to at most EPSILON = 1e-6 (A6) or the
// node lies outside the valid cells.  The node mass (.w) is kept.
__global__ void k_mpm_momentum(KP k, MP m, const int* __restrict__ cs,
                               const float4* __restrict__ pos, const float4* __restrict__ vel,
                               float4* __restrict__ nodev, int nodes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  float3 xn;
  int3 g;
  float sx = 0.f, sy = 0.f, sz = 0.f, sw = 0.f;
  if (node_cell(k, m, i, xn, g)) {
    const float c = m.dx;
    node_cells(k, cs, g, [&](int j) {
      const float4 pj = pos[j];
      const float dx = fabsf(xn.x - pj.x), dy = fabsf(xn.y - pj.y), dz = fabsf(xn.z - pj.z);
      if (dx >= c || dy >= c || dz >= c) return;
      const float w = __fmul_rn(__fmul_rn(1.f - dx / c, 1.f - dy / c), 1.f - dz / c);
      const float4 v = vel[j];
      sx = __fadd_rn(sx, __fmul_rn(v.x, w));
      sy = __fadd_rn(sy, __fmul_rn(v.y, w));
      sz = __fadd_rn(sz, __fmul_rn(v.z, w));
      sw = __fadd_rn(sw, w);
    });
  }
  float4 o = nodev[i];
  if (sw > 1e-6f) {
    o.x = sx / sw; o.y = sy / sw; o.z = sz / sw;
  } else {
    o.x = o.y = o.z = 0.f;
  }
  nodev[i] = o;
}

// MpmColorTest (MpmSolver.cu:75-105): particles in the 27 cells around a
// node left of x = 0 are coloured white (a debugging aid of the reference).
__global__ void k_mpm_color_test(KP k, MP m, const int* __restrict__ cs,
                                 const int* __restrict__ perm, float4* __restrict__ color,
                                 int nodes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes) return;
  float3 xn;
  int3 g;
  const bool ok = node_cell(k, m, i, xn, g);
  if (!(xn.x < 0.f) || !ok) return;
  node_cells(k, cs, g, [&](int j) { color[perm[j]] = make_float4(1.f, 1.f, 1.f, 1.f); });
}

}  // namespace sphx

extern "C" {

int sphx_mpm_get_momentum(sphx_context* c) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_get_momentum"));
  SPHX_CHECK(c->sorted_valid, "sphx_mpm_get_momentum", "grid not built (call sphx_grid_build first)");
  const int nodes = (int)c->mpm.nodes;
  k_mpm_momentum<<<blocks_for(nodes, 256), 256, 0, c->stream>>>(
      make_kp(c), make_mp(c), c->cell_start.p, c->POS(), c->VEL(), c->mpm.nodev.p, nodes);
  SPHX_HIP(hipGetLastError());
  return timer_collect(c);
}

int sphx_mpm_color_test(sphx_context* c) {
  SPHX_TRY(mpm_need(c, "sphx_mpm_color_test"));
  SPHX_CHECK(c->sorted_valid, "sphx_mpm_color_test", "grid not built (call sphx_grid_build first)");
  const int nodes = (int)c->mpm.nodes;
  k_mpm_color_test<<<blocks_for(nodes, 256), 256, 0, c->stream>>>(
      make_kp(c), make_mp(c), c->cell_start.p, c->PERM(), c->cold_color.p, nodes);
  SPHX_HIP(hipGetLastError());
  return timer_collect(c);
}

}  // extern "C"
