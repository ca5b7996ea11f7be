// FP64 conservation totals of the fluid particles, on the device.
//
// Per scene (one for a single scene, every scene of a batched ensemble):
//   [0] sum m, [1..3] sum m v, [4] sum m |v|^2 / 2, [5..7] sum m x
// over the in-grid fluid particles (type 0), accumulated in FP64.  The
// centre of mass is [5..7] / [0].  For one scene, one workgroup per 256
// particles writes its sums (the integration pass writes the same rows
// for the state it produces) and one workgroup adds the rows in a fixed
// order, so a single scene's totals are reproducible bit for bit; an
// ensemble's per-scene sums (TOT_BLOCKS workgroups) go through LDS atomics
// and are exact only to FP64 rounding.
// While tracking is on, every integration appends its totals to a device
// ring (the ring index is a device counter, so a captured step advances it
// on every replay).
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace sphx {

constexpr int TOT_Q = SPHX_TOTALS_Q;
constexpr int TOT_BLOCKS = 512;
constexpr int TOT_THREADS = 256;
constexpr int TOT_HIST = 4096;

__global__ __launch_bounds__(TOT_THREADS) void k_totals_partial(
    int n, int gridtotal, const uint32_t* __restrict__ key, const float4* __restrict__ pos,
    const float4* __restrict__ vel, const float4* __restrict__ phys, int nscenes,
    double* __restrict__ partial) {
  __shared__ double sh[SPHX_MAX_SCENES * TOT_Q];
  const int S = nscenes > 1 ? nscenes : 1;
  const int stride = TOT_BLOCKS * TOT_THREADS;
  if (S == 1) {
    // one particle per thread over blocks_for(n) workgroups, the layout and
    // arithmetic of the integration pass's totals (the same bits)
    double a[TOT_Q] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int i = xcd_index();
    if (i < n) {
      const float4 p = pos[i];
      if (type_of(p) == SPHX_TYPE_FLUID && (int)key[i] != gridtotal) {
        const float4 v = vel[i];
        particle_totals(phys[i].x, v.x, v.y, v.z, p.x, p.y, p.z, a);
      }
    }
    block_totals(a, partial);
    return;
  }
  for (int t = threadIdx.x; t < S * TOT_Q; t += TOT_THREADS) sh[t] = 0.0;
  __syncthreads();
  for (int i = blockIdx.x * TOT_THREADS + threadIdx.x; i < n; i += stride) {
    const float4 p = pos[i];
    if (type_of(p) != SPHX_TYPE_FLUID || (int)key[i] == gridtotal) continue;
    const int s = scene_of(p);
    if (s >= S) continue;
    const float4 v = vel[i];
    const double m = phys[i].x, vx = v.x, vy = v.y, vz = v.z;
    double* o = sh + s * TOT_Q;
    atomicAdd(o + 0, m);
    atomicAdd(o + 1, m * vx); atomicAdd(o + 2, m * vy); atomicAdd(o + 3, m * vz);
    atomicAdd(o + 4, 0.5 * m * (vx * vx + vy * vy + vz * vz));
    atomicAdd(o + 5, m * (double)p.x); atomicAdd(o + 6, m * (double)p.y);
    atomicAdd(o + 7, m * (double)p.z);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < S * TOT_Q; t += TOT_THREADS)
    partial[(size_t)blockIdx.x * S * TOT_Q + t] = sh[t];
}

// One workgroup: entry t of [S][Q] = the sum over the blocks' partials in a
// fixed order (TOT_PARTS interleaved partial sums over blocks b = part,
// part + TOT_PARTS, ..., then the parts in order); with a ring, also stored
// at slot (*count % cap), and the count advanced.
constexpr int TOT_PARTS = 128;
constexpr int TOT_FINAL_THREADS = TOT_Q * TOT_PARTS;
__global__ __launch_bounds__(TOT_FINAL_THREADS) void k_totals_final(
    const double* __restrict__ partial, int nblocks, int nscenes, double* __restrict__ out,
    int* __restrict__ count, double* __restrict__ hist, int cap) {
  __shared__ double ps[TOT_PARTS][TOT_Q];
  const int S = nscenes > 1 ? nscenes : 1;
  const int slot = hist ? (*count % cap) : 0;
  const int q = threadIdx.x % TOT_Q, part = threadIdx.x / TOT_Q;   // one (quantity, part) each
  for (int t0 = 0; t0 < S * TOT_Q; t0 += TOT_Q) {
    const int t = t0 + q;
    double a = 0.0;
#pragma unroll 4
    for (int b = part; b < nblocks; b += TOT_PARTS) a += partial[(size_t)b * S * TOT_Q + t];
    ps[part][q] = a;
    __syncthreads();
    if (part == 0) {
      double v = 0.0;
      for (int u = 0; u < TOT_PARTS; ++u) v += ps[u][q];
      out[t] = v;
      if (hist) hist[(size_t)slot * S * TOT_Q + t] = v;
    }
    __syncthreads();
  }
  if (hist && threadIdx.x == 0) *count += 1;
}


static int totals_alloc(sphx_context* c) {
  const int S = std::max(1, c->nscenes);
  if (c->tot_scenes != S || !c->tot_part.p) {
    // partials of TOT_BLOCKS workgroups, or of every integration workgroup
    // (a single scene's totals are summed by the integration pass)
    SPHX_TRY(c->tot_part.alloc((size_t)std::max(TOT_BLOCKS, blocks_for(c->cap, TOT_THREADS)) * S *
                               TOT_Q));
    SPHX_TRY(c->tot_out.alloc((size_t)S * TOT_Q));
    SPHX_TRY(c->tot_hist.alloc((size_t)TOT_HIST * S * TOT_Q));
    SPHX_TRY(c->tot_count.alloc(1));
    SPHX_HIP(hipMemsetAsync(c->tot_count.p, 0, sizeof(int), c->stream));
    c->tot_scenes = S;
  }
  return 0;
}

// The totals of the current state, enqueued on the context's stream (no
// allocation, no synchronisation: safe inside a captured step once the
// buffers exist).
int totals_enqueue(sphx_context* c, bool ring) {
  SPHX_CHECK(c->tot_part.p && c->tot_scenes == std::max(1, c->nscenes), "conservation totals",
             "buffers not set up for this scene count");
  KP k = make_kp(c);
  // one scene: a workgroup per 256 particles; an ensemble: TOT_BLOCKS
  const int nb = c->nscenes > 1 ? TOT_BLOCKS : blocks_for(k.n, TOT_THREADS);
  SPHX_CHECK((size_t)nb * c->tot_scenes * TOT_Q <= c->tot_part.n, "conservation totals",
             "more particles than the totals were set up for");
  timer_begin(c, "totals");
  k_totals_partial<<<nb, TOT_THREADS, 0, c->stream>>>(
      k.n, k.gridtotal, c->key.p, c->POS(), c->VEL(), c->PHYS(), c->nscenes, c->tot_part.p);
  k_totals_final<<<1, TOT_FINAL_THREADS, 0, c->stream>>>(c->tot_part.p, nb, c->nscenes, c->tot_out.p,
                                                   c->tot_count.p, ring ? c->tot_hist.p : nullptr,
                                                   TOT_HIST);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// The ring entry of a single scene whose integration pass wrote `nblocks`
// per-workgroup partials into tot_part.
int totals_enqueue_final(sphx_context* c, int nblocks) {
  SPHX_CHECK(c->tot_part.p && c->tot_scenes == 1, "conservation totals",
             "buffers not set up for one scene");
  k_totals_final<<<1, TOT_FINAL_THREADS, 0, c->stream>>>(c->tot_part.p, nblocks, 1, c->tot_out.p,
                                                   c->tot_count.p, c->tot_hist.p, TOT_HIST);
  SPHX_HIP(hipGetLastError());
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

int sphx_conservation_totals(sphx_context* c, double* out) {
  SPHX_CHECK(c && out, "sphx_conservation_totals", "null argument");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(c->sorted_valid, "sphx_conservation_totals", "grid not built (call sphx_grid_build first)");
  SPHX_CHECK(c->nscenes <= SPHX_MAX_SCENES, "sphx_conservation_totals", "too many scenes");
  SPHX_TRY(totals_alloc(c));
  SPHX_TRY(totals_enqueue(c, false));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  const size_t m = (size_t)c->tot_scenes * TOT_Q;
  SPHX_HIP(hipMemcpy(out, c->tot_out.p, m * sizeof(double), hipMemcpyDeviceToHost));
  return timer_flush(c);
}

int sphx_conservation_track(sphx_context* c, int on) {
  SPHX_CHECK(c, "sphx_conservation_track", "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(c->nscenes <= SPHX_MAX_SCENES, "sphx_conservation_track", "too many scenes");
  if (on) {
    SPHX_TRY(totals_alloc(c));
    SPHX_HIP(hipMemsetAsync(c->tot_count.p, 0, sizeof(int), c->stream));
    SPHX_HIP(hipStreamSynchronize(c->stream));
  }
  c->tot_track = on != 0;
  return 0;
}

int sphx_conservation_history(sphx_context* c, double* out, int max_steps, int* steps) {
  SPHX_CHECK(c && steps && max_steps >= 0 && (out || max_steps == 0), "sphx_conservation_history",
             "bad argument");
  SPHX_TRY(ensure_device(c));
  *steps = 0;
  if (!c->tot_count.p) return 0;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  int count = 0;
  SPHX_HIP(hipMemcpy(&count, c->tot_count.p, sizeof(int), hipMemcpyDeviceToHost));
  const int kept = std::min(count, TOT_HIST);
  const int take = std::min(kept, max_steps);
  const size_t row = (size_t)c->tot_scenes * TOT_Q;
  std::vector<double> h((size_t)TOT_HIST * row);
  SPHX_HIP(hipMemcpy(h.data(), c->tot_hist.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  // the last `take` steps, oldest first
  for (int s = 0; s < take; ++s) {
    const int step = count - take + s;
    std::copy(h.begin() + (size_t)(step % TOT_HIST) * row,
              h.begin() + (size_t)(step % TOT_HIST + 1) * row, out + (size_t)s * row);
  }
  *steps = take;
  return 0;
}

}  // extern "C"
