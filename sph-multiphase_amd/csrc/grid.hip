// Spatial-hash grid build: cell hash, stable LSD radix sort, cell ranges,
// reorder gather.
//
// Reference: InitialSort / CalcFirstCnt / GetCnt / RearrangeData
// (fluid_system_kern.cu:36-121) and thrust::sort_by_key
// (fluid_system_host.cu:147-210).  Outputs are bit-identical to a stable
// sort of (cell, index) with GRID_UNDEF last.
//
// Design:
//  * keys are cell ids in [0, gridtotal); out-of-grid particles get the
//    sentinel `gridtotal`, so a key needs only bit_width(gridtotal) bits
//    (18 for a 52^3 grid) and the sort runs ceil(bits/9) passes of <= 9 bits.
//  * each pass is upsweep (per-tile digit histogram, one LDS atomic per run
//    of equal digits in lane order) -> one-workgroup scan of the digit-major
//    histogram -> downsweep (stable in-tile ranking: when a round's digits
//    do not decrease in lane order, as for nearly sorted keys, a lane's
//    rank is its offset in its run of equal digits; otherwise lanes holding
//    equal digits find each other with one wave64 ballot per digit bit; a
//    per-wave running count in LDS orders the rounds; a block prefix over
//    the four waves orders the waves).
//  * cell ranges are stored dense: cell_start[c] = number of valid particles
//    in cells < c, for every c in [0, gridtotal].  The reference's
//    mgridoff/mgridcnt are derived from it on download.  Because cell ids
//    are x-fastest (kern.cu:55), the three cells dx=-1..1 of a neighbour
//    row are contiguous and a row is the single range
//    [cell_start[c-1], cell_start[c+2]).
#include "common.hpp"

namespace sphx {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;   // 2048 keys per workgroup
constexpr int RS_MAX_BITS = 9;
constexpr int RS_MAX_BINS = 1 << RS_MAX_BITS;
constexpr int RS_WAVES = RS_THREADS / 64;

// --------------------------------------------------------------- hashing --
// InitialSort (kern.cu:36-64): gc = int((pos - gridmin) * gridIdfac), valid
// when 1 <= gc < res-1 on every axis; key = (gy*resz + gz)*resx + gx.
__global__ void k_hash(const float4* __restrict__ pos, uint32_t* __restrict__ key,
                       uint32_t* __restrict__ val, int n, float3 gmin, float3 fac, int3 res,
                       uint32_t sentinel) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  // plain FP32 subtract then multiply (no contraction possible), truncation
  float fx = (p.x - gmin.x) * fac.x;
  float fy = (p.y - gmin.y) * fac.y;
  float fz = (p.z - gmin.z) * fac.z;
  int gx = (int)fx, gy = (int)fy, gz = (int)fz;
  uint32_t k = sentinel;
  if (gx >= 1 && gx < res.x - 1 && gy >= 1 && gy < res.y - 1 && gz >= 1 && gz < res.z - 1)
    k = (uint32_t)((gy * res.z + gz) * res.x + gx);
  key[i] = k;
  val[i] = (uint32_t)i;
}

// Lanes of the wave whose `digit` equals this lane's (only `active` lanes).
__device__ inline uint64_t match_digit(uint32_t digit, int bits, uint64_t active) {
  uint64_t m = active;
  for (int b = 0; b < bits; ++b) {
    uint64_t bal = __ballot((digit >> b) & 1u);
    m &= ((digit >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

__device__ inline uint64_t lanemask_lt() {
  int lane = threadIdx.x & 63;
  return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_upsweep(const uint32_t* __restrict__ keys,
                                                            int n, int shift, int bits,
                                                            uint32_t* __restrict__ hist,
                                                            int nblocks) {
  __shared__ uint32_t h[RS_MAX_BINS];
  const int bins = 1 << bits;
  const uint32_t mask = (uint32_t)bins - 1u;
  for (int t = threadIdx.x; t < bins; t += RS_THREADS) h[t] = 0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int base = blockIdx.x * RS_TILE + wave * (64 * RS_ITEMS);
  for (int it = 0; it < RS_ITEMS; ++it) {
    int idx = base + it * 64 + lane;
    bool valid = idx < n;
    uint32_t d = valid ? ((keys[idx] >> shift) & mask) : 0u;
    // Each run of equal digits in lane order adds its length once (the
    // keys arrive nearly sorted, so runs are long); the valid lanes are a
    // prefix of the wave.
    const int nvalid = __popcll(__ballot(valid));
    const uint32_t dprev = __shfl_up(d, 1, 64);
    const bool head = valid && (lane == 0 || d != dprev);
    const uint64_t above = __ballot(head) & ~((2ull << lane) - 1ull);
    const int next = above ? __builtin_ctzll(above) : 64;
    if (head) atomicAdd(&h[d], (uint32_t)(min(next, nvalid) - lane));
  }
  __syncthreads();
  for (int t = threadIdx.x; t < bins; t += RS_THREADS) hist[(size_t)t * nblocks + blockIdx.x] = h[t];
}

// Device-wide exclusive scan of uint32 in three coalesced launches:
// per-tile sums, a scan of the tile sums, per-tile scan plus tile offset.
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;   // 2048 entries per tile

// inclusive wave64 scan
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// exclusive block scan of one value per thread; returns the block total too
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t ws[SC_THREADS / 64];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan(v);
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < SC_THREADS / 64; ++q) {
    uint32_t s = ws[q];
    if (q < w) off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_reduce(const uint32_t* __restrict__ a,
                                                             int count,
                                                             uint32_t* __restrict__ tile_sum) {
  int base = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    int i = base + k;
    if (i < count) s += a[i];
  }
  uint32_t tot;
  (void)block_excl_scan(s, &tot);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

// exclusive scan of the tile sums (count <= SC_THREADS * SC_ITEMS * 64)
__global__ __launch_bounds__(SC_THREADS) void k_scan_top(uint32_t* __restrict__ t, int count) {
  uint32_t carry = 0;
  for (int base = 0; base < count; base += SC_TILE) {
    uint32_t v[SC_ITEMS], s = 0;
    int b = base + threadIdx.x * SC_ITEMS;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; ++k) {
      v[k] = (b + k < count) ? t[b + k] : 0u;
      s += v[k];
    }
    uint32_t tot;
    uint32_t run = carry + block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < SC_ITEMS; ++k) {
      if (b + k < count) t[b + k] = run;
      run += v[k];
    }
    carry += tot;
  }
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_down(uint32_t* __restrict__ a, int count,
                                                           const uint32_t* __restrict__ tile_off) {
  int base = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
  uint32_t v[SC_ITEMS], s = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    int i = base + k;
    v[k] = (i < count) ? a[i] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = tile_off[blockIdx.x] + block_excl_scan(s, &tot);
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    int i = base + k;
    if (i < count) a[i] = run;
    run += v[k];
  }
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_downsweep(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int n, int shift, int bits,
    const uint32_t* __restrict__ hist, int nblocks) {
  __shared__ uint32_t wcnt[RS_WAVES][RS_MAX_BINS];
  __shared__ uint32_t goff[RS_MAX_BINS];
  const int bins = 1 << bits;
  const uint32_t mask = (uint32_t)bins - 1u;
  for (int t = threadIdx.x; t < bins; t += RS_THREADS) {
    goff[t] = hist[(size_t)t * nblocks + blockIdx.x];
    for (int w = 0; w < RS_WAVES; ++w) wcnt[w][t] = 0;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int base = blockIdx.x * RS_TILE + wave * (64 * RS_ITEMS);
  uint32_t kk[RS_ITEMS], vv[RS_ITEMS], rk[RS_ITEMS];
  const uint64_t lt = lanemask_lt();
  for (int it = 0; it < RS_ITEMS; ++it) {
    int idx = base + it * 64 + lane;
    bool valid = idx < n;
    uint32_t k = valid ? kin[idx] : 0u;
    uint32_t v = valid ? vin[idx] : 0u;
    uint32_t d = (k >> shift) & mask;
    uint64_t active = __ballot(valid);
    // When the round's digits do not decrease in lane order (the common
    // case for nearly sorted keys), a lane's peers are its run of equal
    // digits: rank and count come from the run heads.  Otherwise the
    // ballot match finds them.
    const uint32_t dprev = __shfl_up(d, 1, 64);
    const bool mono = __ballot(valid && lane > 0 && d < dprev) == 0ull;
    uint32_t rank, count;
    bool leader;
    if (mono) {
      const bool head = valid && (lane == 0 || d != dprev);
      const uint64_t heads = __ballot(head);
      const uint64_t upto = (2ull << lane) - 1ull;   // lanes 0..lane
      const uint64_t le = heads & upto, above = heads & ~upto;
      const int start = le ? 63 - __builtin_clzll(le) : 0;
      const int next = above ? __builtin_ctzll(above) : 64;
      rank = (uint32_t)(lane - start);
      count = (uint32_t)(min(next, __popcll(active)) - lane);
      leader = head;
    } else {
      const uint64_t peers = match_digit(d, bits, active);
      rank = (uint32_t)__popcll(peers & lt);
      count = (uint32_t)__popcll(peers);
      leader = (peers & lt) == 0;
    }
    uint32_t prev = 0;
    if (valid) prev = wcnt[wave][d];
    // every peer has read `prev` before the group leader publishes the
    // new running count (same wave, program order on the LDS queue)
    __builtin_amdgcn_wave_barrier();
    if (valid && leader) wcnt[wave][d] = prev + count;
    __builtin_amdgcn_wave_barrier();
    kk[it] = k;
    vv[it] = v;
    rk[it] = valid ? prev + rank : 0xFFFFFFFFu;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < bins; t += RS_THREADS) {
    uint32_t s = 0;
    for (int w = 0; w < RS_WAVES; ++w) {
      uint32_t c = wcnt[w][t];
      wcnt[w][t] = s;
      s += c;
    }
  }
  __syncthreads();
  for (int it = 0; it < RS_ITEMS; ++it) {
    if (rk[it] == 0xFFFFFFFFu) continue;
    uint32_t d = (kk[it] >> shift) & mask;
    uint32_t dst = goff[d] + wcnt[wave][d] + rk[it];
    kout[dst] = kk[it];
    vout[dst] = vv[it];
  }
}

// cell_start[c] = lower_bound(sorted keys, c) for c in [0, gridtotal]: one
// thread per cell (no serial gaps).  The buffer's previous contents (the
// last grid build's ranges, or anything) are only the starting guess: a
// galloping search from it brackets the bound, and a binary search inside
// the bracket finds it.  Particles move little between steps, so the bound
// is usually found in two or three loads instead of log2(n).
__global__ void k_cell_start(const uint32_t* __restrict__ skey, int n, int gridtotal,
                             int* __restrict__ cell_start) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > gridtotal) return;
  const int g = min(max(cell_start[c], 0), n);
  // the bound lies in [lo, hi]: keys before lo are < c, the key at hi (if
  // hi < n) is >= c
  int lo, hi;
  if (g < n && (int)skey[g] < c) {
    lo = g + 1;
    hi = n;
    for (int step = 1;; step <<= 1) {
      const int p = g + step;
      if (p >= n) break;
      if ((int)skey[p] >= c) { hi = p; break; }
      lo = p + 1;
    }
  } else {
    lo = 0;
    hi = g;
    for (int step = 1;; step <<= 1) {
      const int p = g - step;
      if (p < 0) break;
      if ((int)skey[p] < c) { lo = p + 1; break; }
      hi = p;
    }
  }
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if ((int)skey[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  cell_start[c] = lo;
}

// RearrangeData (kern.cu:98-121): slot s takes the particle idsort[s]; an
// out-of-grid slot keeps its own pre-sort data with pos = (-1000)^3 (the
// reference does not copy calc data there, A9).  perm follows the same
// gather so cold fields stay reachable; mfid maps bornid -> slot (the
// intended inverse permutation, A8).
__global__ void k_reorder(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ idsort,
                          int n, uint32_t sentinel,
                          const float4* __restrict__ pos_in, float4* __restrict__ pos_out,
                          const float4* __restrict__ vel_in, float4* __restrict__ vel_out,
                          const float4* __restrict__ ve_in, float4* __restrict__ ve_out,
                          const float4* __restrict__ acc_in, float4* __restrict__ acc_out,
                          const float4* __restrict__ phys_in, float4* __restrict__ phys_out,
                          const int* __restrict__ born_in, int* __restrict__ born_out,
                          const int* __restrict__ perm_in, int* __restrict__ perm_out,
                          int* __restrict__ mfid) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  bool valid = skey[s] != sentinel;
  int src = valid ? (int)idsort[s] : s;
  float4 p = pos_in[src];
  if (!valid) { p.x = -1000.f; p.y = -1000.f; p.z = -1000.f; }
  pos_out[s] = p;
  vel_out[s] = vel_in[src];
  ve_out[s] = ve_in[src];
  acc_out[s] = acc_in[src];
  phys_out[s] = phys_in[src];
  int b = born_in[src];
  born_out[s] = b;
  perm_out[s] = perm_in[src];
  if (valid && b >= 0 && b < n) mfid[b] = s;
}

__global__ void k_fill_i32(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

int mf_reorder(sphx_context* c, int a, int b);
int mlp_reorder(sphx_context* c, int a, int b);
int mpm_reorder(sphx_context* c, int a, int b);

int grid_hash(sphx_context* c) {
  int n = c->num;
  c->sorted_valid = false;
  positions_changed(c);
  if (n == 0) return 0;
  KP k = make_kp(c);
  timer_begin(c, "hash");
  k_hash<<<blocks_for(n, 256), 256, 0, c->stream>>>(c->POS(), c->key.p, c->idsort.p, n, k.gridmin,
                                                      k.gridfac, k.gridres,
                                                      (uint32_t)c->P.gridtotal);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

int grid_sort(sphx_context* c) {
  int n = c->num;
  if (n == 0) {
    SPHX_HIP(hipMemsetAsync(c->cell_start.p, 0, ((size_t)c->P.gridtotal + 1) * sizeof(int),
                            c->stream));
    c->sorted_valid = true;
    return 0;
  }
  int nblocks = blocks_for(n, RS_TILE);
  int passes = (c->key_bits + RS_MAX_BITS - 1) / RS_MAX_BITS;
  int per = (c->key_bits + passes - 1) / passes;
  size_t hist_n = (size_t)nblocks * RS_MAX_BINS;
  int scan_tiles = blocks_for((int)hist_n, SC_TILE);
  SPHX_CHECK(scan_tiles <= SC_TILE * 64, "sphx_grid_sort", "too many particles for one scan");
  SPHX_TRY(c->block_hist.alloc(hist_n + scan_tiles + 1));
  uint32_t* tile_sum = c->block_hist.p + hist_n;
  uint32_t *ka = c->key.p, *va = c->idsort.p, *kb = c->key_tmp.p, *vb = c->val_tmp.p;
  int shift = 0;
  timer_begin(c, "radix_sort");
  for (int ps = 0; ps < passes; ++ps) {
    int bits = std::min(per, c->key_bits - shift);
    int bins = 1 << bits;
    int cnt = bins * nblocks;
    int tiles = blocks_for(cnt, SC_TILE);
    k_rs_upsweep<<<nblocks, RS_THREADS, 0, c->stream>>>(ka, n, shift, bits, c->block_hist.p, nblocks);
    k_scan_reduce<<<tiles, SC_THREADS, 0, c->stream>>>(c->block_hist.p, cnt, tile_sum);
    k_scan_top<<<1, SC_THREADS, 0, c->stream>>>(tile_sum, tiles);
    k_scan_down<<<tiles, SC_THREADS, 0, c->stream>>>(c->block_hist.p, cnt, tile_sum);
    k_rs_downsweep<<<nblocks, RS_THREADS, 0, c->stream>>>(ka, va, kb, vb, n, shift, bits,
                                                          c->block_hist.p, nblocks);
    std::swap(ka, kb);
    std::swap(va, vb);
    shift += bits;
  }
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  if (ka != c->key.p) {
    SPHX_HIP(hipMemcpyAsync(c->key.p, ka, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
    SPHX_HIP(hipMemcpyAsync(c->idsort.p, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
  }
  timer_begin(c, "cell_start");
  k_cell_start<<<blocks_for(c->P.gridtotal + 1, 256), 256, 0, c->stream>>>(
      c->key.p, n, c->P.gridtotal, c->cell_start.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  c->sorted_valid = true;
  return 0;
}

int grid_reorder(sphx_context* c) {
  int n = c->num;
  if (n == 0) return 0;
  SPHX_CHECK(c->sorted_valid, "sphx_grid_reorder", "grid not sorted (call hash and sort first)");
  int a = c->cur, b = 1 - c->cur;
  positions_changed(c);
  k_fill_i32<<<blocks_for(c->cap, 256), 256, 0, c->stream>>>(c->mfid.p, c->cap, -1);
  timer_begin(c, "reorder");
  k_reorder<<<blocks_for(n, 256), 256, 0, c->stream>>>(
      c->key.p, c->idsort.p, n, (uint32_t)c->P.gridtotal, c->pos[a].p, c->pos[b].p, c->vel[a].p,
      c->vel[b].p, c->veleval[a].p, c->veleval[b].p, c->accel[a].p, c->accel[b].p, c->phys[a].p,
      c->phys[b].p, c->bornid[a].p, c->bornid[b].p, c->perm[a].p, c->perm[b].p, c->mfid.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  SPHX_TRY(mf_reorder(c, a, b));
  SPHX_TRY(mlp_reorder(c, a, b));
  SPHX_TRY(mpm_reorder(c, a, b));
  c->cur = b;
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

int sphx_grid_hash(sphx_context* c) {
  SPHX_CHECK(c, "sphx_grid_hash", "null context");
  SPHX_TRY(ensure_device(c));
  return grid_hash(c);
}

int sphx_grid_sort(sphx_context* c) {
  SPHX_CHECK(c, "sphx_grid_sort", "null context");
  SPHX_TRY(ensure_device(c));
  return grid_sort(c);
}

int sphx_grid_reorder(sphx_context* c) {
  SPHX_CHECK(c, "sphx_grid_reorder", "null context");
  SPHX_TRY(ensure_device(c));
  return grid_reorder(c);
}

int sphx_grid_build(sphx_context* c) {
  SPHX_CHECK(c, "sphx_grid_build", "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_TRY(grid_hash(c));
  SPHX_TRY(grid_sort(c));
  return grid_reorder(c);
}

int sphx_download_grid(sphx_context* c, uint32_t* mgcell, uint32_t* midsort, uint32_t* mgndx,
                       int32_t* gridoff, int32_t* gridcnt, int32_t* MFidTable) {
  SPHX_CHECK(c, "sphx_download_grid", "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  int n = c->num, gt = c->P.gridtotal;
  std::vector<uint32_t> key(n);
  std::vector<int> start((size_t)gt + 1);
  if (n > 0) SPHX_HIP(hipMemcpy(key.data(), c->key.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  SPHX_HIP(hipMemcpy(start.data(), c->cell_start.p, ((size_t)gt + 1) * sizeof(int),
                     hipMemcpyDeviceToHost));
  if (mgcell)
    for (int i = 0; i < n; ++i) mgcell[i] = key[i] == (uint32_t)gt ? SPHX_GRID_UNDEF : key[i];
  if (midsort && n > 0)
    SPHX_HIP(hipMemcpy(midsort, c->idsort.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (mgndx)
    for (int i = 0; i < n; ++i)
      mgndx[i] = key[i] == (uint32_t)gt ? SPHX_GRID_UNDEF : (uint32_t)(i - start[key[i]]);
  for (int cidx = 0; cidx < gt; ++cidx) {
    int cnt = start[cidx + 1] - start[cidx];
    if (gridcnt) gridcnt[cidx] = cnt;
    if (gridoff) gridoff[cidx] = cnt > 0 ? start[cidx] : 0;
  }
  if (MFidTable && n > 0)
    SPHX_HIP(hipMemcpy(MFidTable, c->mfid.p, n * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
