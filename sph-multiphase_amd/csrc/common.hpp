// Internal definitions shared by the HIP translation units of libsphx.
//
// Device layout (private; the public boundary is the AoS of sphx.h):
//   particles live in cell-sorted order as structure-of-arrays; the fields
//   read by neighbour loops are packed into 16-byte records so one
//   global_load_dwordx4 fetches what a neighbour visit needs:
//     pos  = {x, y, z, type bits}        vel  = {vx, vy, vz, visc}
//     phys = {mass, dens, pressure, restdens}
//   Fields no hot-path kernel reads (color, deformGrad, X, invA and, until
//   MPM is enabled, stress/B/solidG/solidK) stay in upload order and are
//   reached through `perm` (sorted slot -> upload slot), so the per-step
//   reorder moves ~100 B per particle instead of the reference's 512 B
//   (fluid_system_host.cu:192-202).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>
#include <map>

#include "../../include/sphx.h"

namespace sphx {

// ---------------------------------------------------------------- errors --
void set_error(const std::string& msg);
int fail(const char* where, const char* what);
int hip_fail(const char* where, hipError_t e);

#define SPHX_HIP(call)                                                  \
  do {                                                                  \
    hipError_t e__ = (call);                                            \
    if (e__ != hipSuccess) return ::sphx::hip_fail(#call, e__);         \
  } while (0)

#define SPHX_CHECK(cond, where, what)                                   \
  do {                                                                  \
    if (!(cond)) return ::sphx::fail(where, what);                      \
  } while (0)

#define SPHX_TRY(call)                                                  \
  do {                                                                  \
    int r__ = (call);                                                   \
    if (r__ != 0) return r__;                                           \
  } while (0)

// ------------------------------------------------------------ device math --
__host__ __device__ inline float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__host__ __device__ inline float3 xyz(const float4& a) { return make_float3(a.x, a.y, a.z); }
__host__ __device__ inline float3 add(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ inline float3 sub(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ inline float3 mul(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__host__ __device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// pos.w holds the displayPack type word: particle type in bits 0-7, the
// scene of a batched ensemble in bits 8-23 (0 for a single scene).
__device__ inline int type_of(const float4& p) { return __float_as_int(p.w) & 0xFF; }
__device__ inline int scene_of(const float4& p) { return (__float_as_int(p.w) >> 8) & 0xFFFF; }
__host__ __device__ inline int type_bits(int t) { return t & 0xFF; }

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Give XCD x the contiguous tile range [x*q + min(x,r), ...) (q, r =
// gridDim / 8, gridDim % 8) so the cell-sorted neighbours its waves gather
// stay in its L2 instead of every XCD streaming the whole domain.
__device__ __forceinline__ int xcd_tile() {
  const int nb = (int)gridDim.x, b = (int)blockIdx.x;
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}
__device__ __forceinline__ int xcd_index() { return xcd_tile() * (int)blockDim.x + (int)threadIdx.x; }

// FP64 conservation totals of one fluid particle: m, m v, m |v|^2 / 2, m x.
// No contraction (the same bits in every kernel that calls it).
__device__ __forceinline__ void particle_totals(double m, float vx, float vy, float vz, float px,
                                                float py, float pz, double* t) {
#pragma clang fp contract(off)
  const double x = vx, y = vy, z = vz;
  t[0] = m;
  t[1] = m * x; t[2] = m * y; t[3] = m * z;
  t[4] = 0.5 * m * (x * x + y * y + z * z);
  t[5] = m * (double)px; t[6] = m * (double)py; t[7] = m * (double)pz;
}
// The workgroup's sums of t (one particle per thread, zeros elsewhere) into
// row blockIdx.x of `rows`, in a fixed order; every thread calls it.
__device__ __forceinline__ void block_totals(double* t, double* __restrict__ rows) {
  __shared__ double sh[16][SPHX_TOTALS_Q];   // up to 16 waves
#pragma unroll
  for (int q = 0; q < SPHX_TOTALS_Q; ++q)
    for (int off = 32; off > 0; off >>= 1) t[q] += __shfl_down(t[q], off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < SPHX_TOTALS_Q; ++q) sh[w][q] = t[q];
  __syncthreads();
  if (threadIdx.x < SPHX_TOTALS_Q) {
    double s = 0.0;
    for (int u = 0; u < (int)(blockDim.x + 63) / 64; ++u) s += sh[u][threadIdx.x];
    rows[(size_t)blockIdx.x * SPHX_TOTALS_Q + threadIdx.x] = s;
  }
}

// Cubic spline kernel W and gradient, support 2h (fluid_system_kern.cu:637-671).
// `r` is x_i - x_j in metres.
__device__ inline float cubic_w(float rmod, float inv_h, float kspline) {
  float q = rmod * inv_h;
  if (q > 2.f) return 0.f;
  if (q > 1.f) { float t = 2.f - q; return 0.25f * t * t * t * kspline; }
  return (1.f - 1.5f * q * q + 0.75f * q * q * q) * kspline;
}
// Returns the scalar f with grad W = r * f (0 outside the support).
__device__ inline float cubic_gradf(float rmod, float inv_rmod, float inv_h, float kspline) {
  float q = rmod * inv_h;
  if (q > 2.f) return 0.f;
  if (q > 1.f) { float t = 2.f - q; return -0.75f * t * t * inv_h * inv_rmod * kspline; }
  return (-3.f * q + 2.25f * q * q) * inv_h * inv_rmod * kspline;
}

// ------------------------------------------------------- kernel constants --
// Passed by value to every kernel (no __device__ parameter symbol, so
// several contexts can coexist and launches are graph-capturable).
struct KP {
  int n;
  int gridtotal;
  float ss;          // simscale
  float h, h2, inv_h;
  float kpoly6, kspikydiff, kspline;
  float dt;
  float3 g;
  float bvisc;
  float3 gridmin, gridfac;
  int3 gridres;
  int rowoff[9];     // neighborid[3r]: the dx=-1 cell of neighbour row r
  // options
  float omega, warm;
  int full_support, boundary_pressure;
  int wcsph_kernel;
  float eos_B, eos_gamma, art_alpha, c0;
  int clamp;
  int clamp_p;
  int nscenes;       // > 0: batched ensemble, per-scene convergence
  float solidK;      // elastic solid stiffness (ParamCarrier::solidK)
  float st_K, st_C;  // surface tension radius factor and strength
};

constexpr int SPHX_MAX_SCENES = 256;

// Workgroup size of each neighbour-loop kernel class, set by the autotuner
// (multiples of 64 up to 256; the kernels do not depend on it otherwise).
enum TuneId {
  T_DENSITY, T_DII, T_AII, T_DP, T_ITER, T_MF_DRIFT, T_MF_ALPHA, T_MF_FORCE,
  T_MLP_FEAT, T_MLP_INFER, T_DENS_DII, T_COUNT
};
// timer names of the kernels, in TuneId order
constexpr const char* TUNE_NAMES[T_COUNT] = {
    "iisph_density", "iisph_dii", "iisph_aii", "iisph_dp", "iisph_iter",
    "mf_drift", "mf_alpha", "mf_force", "mlp_features", "mlp_infer", "iisph_density_dii"};

// Per-scene solver state of a batched ensemble (device arrays).
struct SceneCtl {
  int n;             // number of scenes (0: single scene)
  double* acc;       // [n][3] per-iteration sums: signed residual, error, count
  int* state;        // [n][2] done, iterations
  double* err;       // [n] last error
};

// ----------------------------------------------------------------- buffer --
// Device array owned by its holder: freed by release() or the destructor,
// movable (std::swap works), not copyable.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  int alloc(size_t count) {
    if (count <= n && p) return 0;
    release();
    if (count == 0) return 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    if (e != hipSuccess) { p = nullptr; n = 0; return hip_fail("hipMalloc", e); }
    n = count;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct Timer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<std::string> names;
  std::map<std::string, std::pair<double, int>> acc;
  std::vector<hipEvent_t> pool;   // events reused across flushes
  std::string only;               // time only this kernel (empty: all)
  size_t ended = 0;               // pending entries whose end is recorded
};

}  // namespace sphx

// ---------------------------------------------------------------- context --
struct sphx_context {
  sphx_ParamCarrier P;
  sphx_SolverOptions O;
  int num = 0;
  int cap = 0;
  int device = 0;
  hipStream_t stream = nullptr;

  // hot particle fields, double-buffered for the reorder gather
  sphx::DevBuf<float4> pos[2], vel[2], veleval[2], accel[2], phys[2];
  sphx::DevBuf<int> bornid[2], perm[2];
  int cur = 0;
  // cold fields in upload order (reached through perm)
  sphx::DevBuf<float4> cold_color;
  sphx::DevBuf<float> cold_mat;   // 9*4 mats: deformGrad, stress, B, invA
  sphx::DevBuf<float4> cold_X;    // X, w unused
  sphx::DevBuf<float2> cold_GK;   // solidG, solidK

  // grid
  sphx::DevBuf<uint32_t> key, idsort, key_tmp, val_tmp, block_hist;
  sphx::DevBuf<int> cell_start;   // gridtotal + 1 entries (dense lower bounds)
  sphx::DevBuf<int> mfid;         // bornid -> slot
  int key_bits = 1;
  bool sorted_valid = false;

  // IISPH per-slot buffers (not permuted, as in the reference)
  sphx::DevBuf<float4> dii;       // d_ii.xyz, m
  sphx::DevBuf<float4> cst;       // a_ii, rho0 - rho_adv, m / rho^2, rho0 (the Jacobi pass's constants)
  sphx::DevBuf<float4> dijpj;     // sum_j d_ij p_j .xyz, unused
  sphx::DevBuf<float4> mqa;       // {x, y, z, m q.x}, q = d_ii p + sum d_ij p_j (0 for boundary)
  sphx::DevBuf<float2> mqb;       // {m q.y, m q.z}
  sphx::DevBuf<float4> gh;        // sum m_j gradW_ij (+ boundary), sum m_j |gradW_ij|^2
  bool gh_valid = false;          // gh matches the current positions and d_ii inputs
  sphx::DevBuf<float> press_l, press_l1, rho_adv, residue, sfac;
  sphx::DevBuf<double> red;       // reduction partials
  sphx::DevBuf<int> solve_state;  // [0]=done [1]=iters
  sphx::DevBuf<double> solve_err; // [0]=last mean |rho-rho0|/rho0 [1]=mean signed
  double* host_pinned = nullptr;  // 4 doubles: done, iters, err, signed
  int last_iters = 0;             // iterations of the previous solve
  // per-step neighbour list (sliced, 16-bit entries packed in 16-byte groups); rebuilt after any
  // change of positions or order
  sphx::DevBuf<uint4> nbr_idx;   // packed neighbour list groups
  sphx::DevBuf<int> nbr_cnt;
  sphx::DevBuf<float4> ps;        // {x, y, z, s = m p / rho^2} for the DP pass
  int nbr_cap = 0;
  bool nbr_valid = false;

  // multiphase (drift-flux mixture); nphase == 0 when disabled
  int nphase = 0;
  float mf_rho[3] = {0, 0, 0}, mf_nu[3] = {0, 0, 0};
  sphx::DevBuf<float4> mf_alpha[2], mf_force[2];  // permuted with particles
  sphx::DevBuf<float4> mf_alpha_new, mf_drift[3], mf_amf;
  // drift-stress sum of the last alpha pass; valid while its inputs
  // (positions, densities, drifts, fractions before the step) are unchanged
  sphx::DevBuf<float4> mf_stress;
  bool mf_stress_ok = false;
  // packed neighbour records of the drift-flux loops ({x, y, z, m / rho},
  // {alpha, p}); mf_rec_ok while positions and densities are unchanged
  sphx::DevBuf<float4> mf_rec, mf_reca;
  bool mf_rec_ok = false;
  bool mf_pending = false;        // alpha_new holds the corrected fractions

  // MPM transfer (particle-centric scatter into per-workgroup node tiles,
  // summed per touched 4^3-node block; the touched blocks are listed)
  struct MPM {
    bool on = false;
    int kernel = 0;        // 0 cubic (reference), 1 quadratic
    int physical = 0;      // 1: gradients in metres, momentum-conserving node force
    bool loaded = false;   // hot B / sigma hold the particles' values
    sphx::DevBuf<float> B[2], S[2];          // [9][cap] component-major, permuted
    sphx::DevBuf<float4> acc, force, nodev;  // per node: {mv, m}, {f}, {v, m}
    sphx::DevBuf<int> flags, list, count;    // per node block; active list
    sphx::DevBuf<float4> tiles;              // [workgroup][tile] node sums
    sphx::DevBuf<int4> hdr;                  // [workgroup][3] tile boxes
    sphx::DevBuf<int> direct;                // [2] a pass added directly
    size_t tiles_cap = 0;                    // workgroups the tiles hold
    int nblk[3] = {0, 0, 0};
    size_t nodes = 0, blocks = 0;
  } mpm;

  // learned warm start / drift closure (24-64-64-4 MLP)
  struct MLP {
    bool ready = false;
    int warm = 0, drift = 0, record = 0;
    float p_scale = 1.f, u_scale = 1.f, drift_blend = 1.f;
    int adam_t = 0;
    size_t nsamples = 0, max_samples = 0, wpos = 0;
    // online refresh: every `online_every` solves, `online_iters` Adam steps
    // on batches drawn from the sample ring
    int online_every = 0, online_iters = 0, online_batch = 0, solves = 0;
    int record_stride = 1;        // online recording keeps every stride-th slot
    float online_lr = 0.f, online_lw[4] = {1.f, 1.f, 1.f, 1.f};
    uint64_t online_seed = 0;
    // per particle, permuted with the particles: {drift correction xyz,
    // pressure increment of the last solve}
    sphx::DevBuf<float4> hist[2];
    long long learned_solves = 0, fallbacks = 0;
    int last_fallback = 0;
    // guard of the learned start (single scene): after the solve's first
    // iteration, a probe iteration from the plain start when that iteration's
    // error exceeds guard_factor * tolerance, and the better first iterate kept
    int guard = 1;                      // on by default
    float guard_factor = 10.f;
    bool guard_armed = false;           // the next solve's first iteration gets the guard
    sphx::DevBuf<int> gstate;           // [0] skip probe, [1] restore learned, [2] probes, [3] plain picks
    sphx::DevBuf<float> pkeep;          // the learned start's first iterate
    sphx::DevBuf<float> w, m, v, g;
    sphx::DevBuf<float> gpart;          // per-workgroup weight gradients
    sphx::DevBuf<uint16_t> w16;         // FP16 bits of w
    sphx::DevBuf<uint4> wmf;            // w16 in the MFMA operand layout
    bool wmf_ok = false;                // wmf packed from the current w16
    sphx::DevBuf<float> X, pprev;       // per slot, current step
    sphx::DevBuf<float4> unew;          // drift with the converged pressure
    sphx::DevBuf<float> sX, sT, sM;     // recorded samples
    sphx::DevBuf<int> order;
    sphx::DevBuf<double> loss;
    // asynchronous online refresh: the Adam steps run on their own stream
    // (default priority) while the next step is solved, into staged FP16
    // weights that the step after that adopts
    bool feat_ok = false;        // X and pprev hold this step's features (a_ii pass)
    bool start_ok = false;       // the a_ii pass also wrote the learned start
    bool settled = false;        // this step's due refresh is adopted
    bool fuse_start = true;      // evaluate the network inside the a_ii pass
    int online_async = 0;
    int train_pending = 0;       // solves until the staged weights are adopted
    bool train_launch = false;   // a refresh is due, its samples recorded (ev_main)
    bool train_inflight = false; // a refresh was launched and not yet waited for
    hipStream_t tstream = nullptr;
    hipEvent_t ev_main = nullptr, ev_train = nullptr;
    sphx::DevBuf<uint16_t> w16_stage;
  } mlp;

  // batched ensemble: scenes in disjoint boxes of one grid
  int nscenes = 0;
  sphx::DevBuf<double> sc_acc, sc_err;
  sphx::DevBuf<int> sc_state;
  std::vector<int> scene_first, scene_count;   // upload order
  sphx::SceneCtl scenes() {
    return sphx::SceneCtl{nscenes, sc_acc.p, sc_state.p, sc_err.p};
  }

  // FP64 conservation totals (per scene) and, while tracking, the totals
  // of every integration in a device ring
  sphx::DevBuf<double> tot_part, tot_out, tot_hist;
  sphx::DevBuf<int> tot_count;
  int tot_scenes = 0;             // scenes the buffers were sized for
  bool tot_track = false;

  // captured two-phase step: graph A (sort .. a fixed budget of pressure
  // iterations) and graph B (integration), one of each per buffer parity
  hipGraphExec_t graph_a[2] = {nullptr, nullptr}, graph_b[2] = {nullptr, nullptr};
  int graph_budget = 0, graph_num = -1;

  // persistent solve: the whole pressure solve as one cooperative launch
  bool coop = false;
  int coop_max = 0;                // co-resident workgroups (0: not queried yet)
  int coop_fallbacks = 0;          // device-side fallbacks of the last solve
  // per persistent solve {iterations, error}, appended on the device
  sphx::DevBuf<double> solve_log;
  sphx::DevBuf<int> solve_log_count;

  // workgroup sizes per kernel class (TuneId)
  // defaults: the autotuner's picks on the 1M-particle two-phase tank
  // (TuneId order: density, dii, aii, dp, iter, drift, alpha, force,
  // features, infer, density+dii)
  int tune[sphx::T_COUNT] = {256, 256, 64, 128, 128, 64, 64, 256, 256, 256, 64};

  // timing
  bool timing = false;
  sphx::Timer timer;

  float4* POS() { return pos[cur].p; }
  float4* VEL() { return vel[cur].p; }
  float4* VELEVAL() { return veleval[cur].p; }
  float4* ACCEL() { return accel[cur].p; }
  float4* PHYS() { return phys[cur].p; }
  int* BORN() { return bornid[cur].p; }
  int* PERM() { return perm[cur].p; }
};

namespace sphx {

KP make_kp(const sphx_context* c);
int ensure_device(sphx_context* c);
// Time the enclosed launch(es) when timing is on.
void timer_begin(sphx_context* c, const char* name);
void timer_end(sphx_context* c);
int timer_flush(sphx_context* c);
int timer_collect(sphx_context* c);

inline int blocks_for(int n, int threads) { return (n + threads - 1) / threads; }

// Particle positions or order changed: the neighbour list and the packed
// drift-flux records are stale.
inline void positions_changed(sphx_context* c) {
  c->nbr_valid = false;
  c->gh_valid = false;
  c->mf_stress_ok = false;
  c->mf_rec_ok = false;
  c->mlp.feat_ok = false;
}

}  // namespace sphx
