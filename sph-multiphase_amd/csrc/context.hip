// Context lifecycle, parameter block, AoS <-> SoA transfers, kernel timing.
// Reference counterparts: FluidSetupCUDA / FluidClearCUDA
// (fluid_system_host.cu:79-144), CarryParam (fluid_system_kern.cu:23-25),
// TransferToCUDA / TransferFromCUDA / EmitUpdateCUDA
// (fluid_system.cpp:979-999).
#include "common.hpp"

#include <cstring>
#include <cmath>

namespace sphx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(const char* where, const char* what) {
  set_error(std::string(where) + ": " + what);
  return -1;
}

int hip_fail(const char* where, hipError_t e) {
  set_error(std::string(where) + ": " + hipGetErrorString(e));
  return -2;
}

KP make_kp(const sphx_context* c) {
  const sphx_ParamCarrier& P = c->P;
  const sphx_SolverOptions& O = c->O;
  KP k;
  std::memset(&k, 0, sizeof(k));
  k.n = c->num;
  k.gridtotal = P.gridtotal;
  k.ss = P.simscale;
  k.h = P.smoothradius;
  k.h2 = P.smoothradius * P.smoothradius;
  k.inv_h = 1.0f / P.smoothradius;
  k.kpoly6 = P.kpoly6;
  k.kspikydiff = P.kspikydiff;
  k.kspline = P.kspline;
  k.dt = P.dt;
  k.g = make_float3(P.gravity.x, P.gravity.y, P.gravity.z);
  k.bvisc = P.bvisc;
  k.gridmin = make_float3(P.gridmin.x, P.gridmin.y, P.gridmin.z);
  k.gridfac = make_float3(P.gridIdfac.x, P.gridIdfac.y, P.gridIdfac.z);
  k.gridres = make_int3(P.gridres.x, P.gridres.y, P.gridres.z);
  for (int r = 0; r < 9; ++r) k.rowoff[r] = P.neighborid[3 * r];
  k.omega = O.omega;
  k.warm = O.warm_start;
  k.full_support = O.full_support;
  k.boundary_pressure = O.boundary_pressure;
  k.wcsph_kernel = O.wcsph_kernel;
  k.eos_B = O.eos_B > 0.f ? O.eos_B : -1.f;
  k.eos_gamma = O.eos_gamma;
  k.art_alpha = O.art_visc_alpha;
  k.c0 = O.sound_speed;
  k.clamp = O.clamp_to_grid;
  k.clamp_p = O.clamp_pressure;
  k.nscenes = c->nscenes;
  k.solidK = P.solidK;
  k.st_K = P.surfaceTensionK;
  k.st_C = P.surfaceTensionFluidC;
  return k;
}

int ensure_device(sphx_context* c) {
  SPHX_HIP(hipSetDevice(c->device));
  return 0;
}

static bool timer_event(sphx_context* c, hipEvent_t* e) {
  auto& pool = c->timer.pool;
  if (!pool.empty()) {
    *e = pool.back();
    pool.pop_back();
    return true;
  }
  // timing only: no system-scope fence (with it each record writes back and
  // invalidates the caches, which slows the work that follows)
  return hipEventCreateWithFlags(e, hipEventDisableSystemFence) == hipSuccess;
}

void timer_begin(sphx_context* c, const char* name) {
  if (!c->timing) return;
  if (!c->timer.only.empty() && c->timer.only != name) return;
  hipEvent_t a, b;
  if (!timer_event(c, &a)) return;
  if (!timer_event(c, &b)) { c->timer.pool.push_back(a); return; }
  (void)hipEventRecord(a, c->stream);
  c->timer.pending.push_back({a, b});
  c->timer.names.push_back(name);
}

void timer_end(sphx_context* c) {
  if (!c->timing || c->timer.pending.empty()) return;
  if (c->timer.pending.size() == c->timer.ended) return;   // its begin was filtered out
  (void)hipEventRecord(c->timer.pending.back().second, c->stream);
  c->timer.ended = c->timer.pending.size();
}

// Adds the first `count` timed launches to the totals and returns their
// events to the pool.
static void timer_take(sphx_context* c, size_t count) {
  auto& T = c->timer;
  for (size_t i = 0; i < count; ++i) {
    float ms = 0.f;
    hipEvent_t a = T.pending[i].first, b = T.pending[i].second;
    if (hipEventElapsedTime(&ms, a, b) == hipSuccess) {
      auto& e = T.acc[T.names[i]];
      e.first += ms;
      e.second += 1;
    }
    T.pool.push_back(a);
    T.pool.push_back(b);
  }
  T.pending.erase(T.pending.begin(), T.pending.begin() + count);
  T.names.erase(T.names.begin(), T.names.begin() + count);
  T.ended = T.ended > count ? T.ended - count : 0;
}

// Every timed launch, after the stream has drained (the readers of the
// totals call this).
int timer_flush(sphx_context* c) {
  if (c->timer.pending.empty()) return 0;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  timer_take(c, c->timer.pending.size());
  return 0;
}

// The timed launches that have already completed, without waiting: the end
// of a step calls this, so that timing a run adds no synchronisation to it.
// The events of one stream complete in order, so the first end event still
// pending stops the scan; pairs from `ended` on may lack their end event.
int timer_collect(sphx_context* c) {
  auto& T = c->timer;
  size_t done = 0;
  while (done < T.ended) {
    const hipError_t q = hipEventQuery(T.pending[done].second);
    if (q != hipSuccess) {
      // hipErrorNotReady is left as the thread's last error: clear it so
      // that the next launch check does not report it
      (void)hipGetLastError();
      break;
    }
    ++done;
  }
  if (done) timer_take(c, done);
  return 0;
}

// --------------------------------------------------------------- kernels --
__global__ void k_iota(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// cold[s] <- cold_old[perm[s]] for every cold field (materialise slot order)
__global__ void k_cold_gather(const int* __restrict__ perm, int n,
                              const float4* __restrict__ color_in, float4* __restrict__ color_out,
                              const float* __restrict__ mat_in, float* __restrict__ mat_out,
                              const float4* __restrict__ X_in, float4* __restrict__ X_out,
                              const float2* __restrict__ gk_in, float2* __restrict__ gk_out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int src = perm[s];
  color_out[s] = color_in[src];
  X_out[s] = X_in[src];
  gk_out[s] = gk_in[src];
  for (int k = 0; k < 36; ++k) mat_out[(size_t)s * 36 + k] = mat_in[(size_t)src * 36 + k];
}

template <class T>
__global__ void k_gather(const int* __restrict__ perm, int n, const T* __restrict__ in,
                         T* __restrict__ out, int width) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  size_t src = (size_t)perm[s];
  for (int k = 0; k < width; ++k) out[(size_t)s * width + k] = in[src * width + k];
}

int mpm_sync_cold(sphx_context* c);
void mpm_invalidate(sphx_context* c);

static int materialize_cold(sphx_context* c) {
  int n = c->num;
  if (n == 0) return 0;
  DevBuf<float4> color, X;
  DevBuf<float> mat;
  DevBuf<float2> gk;
  SPHX_TRY(color.alloc(c->cap));
  SPHX_TRY(X.alloc(c->cap));
  SPHX_TRY(mat.alloc((size_t)c->cap * 36));
  SPHX_TRY(gk.alloc(c->cap));
  k_cold_gather<<<blocks_for(n, 256), 256, 0, c->stream>>>(
      c->PERM(), n, c->cold_color.p, color.p, c->cold_mat.p, mat.p, c->cold_X.p, X.p,
      c->cold_GK.p, gk.p);
  SPHX_HIP(hipGetLastError());
  k_iota<<<blocks_for(c->cap, 256), 256, 0, c->stream>>>(c->PERM(), c->cap);
  SPHX_HIP(hipGetLastError());
  SPHX_HIP(hipStreamSynchronize(c->stream));
  std::swap(c->cold_color, color);
  std::swap(c->cold_X, X);
  std::swap(c->cold_mat, mat);
  std::swap(c->cold_GK, gk);
  color.release();
  X.release();
  mat.release();
  gk.release();
  return 0;
}

// The neighbour loops read the rows [c-1, c+2) of the dense cell table; that
// is in range only for the grid SetupGrid builds (fluid_system.cpp:136-143).
static int validate_grid(const sphx_ParamCarrier& P) {
  const sphx_int3 r = P.gridres;
  SPHX_CHECK(r.x >= 3 && r.y >= 3 && r.z >= 3, "grid", "gridres must be >= 3 on every axis");
  SPHX_CHECK((long long)r.x * r.y * r.z == (long long)P.gridtotal && P.gridtotal < (1 << 30),
             "grid", "gridtotal must equal gridres.x*gridres.y*gridres.z (< 2^30)");
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        SPHX_CHECK(P.neighborid[i * 9 + j * 3 + k] ==
                       (i - 1) * r.x * r.z + (j - 1) * r.x + (k - 1),
                   "grid", "neighborid does not match gridres (use sphx_setup_grid)");
  SPHX_CHECK(P.simscale > 0.f && P.smoothradius > 0.f, "grid",
             "simscale and smoothradius must be positive");
  return 0;
}

static int alloc_grid(sphx_context* c) {
  int gt = c->P.gridtotal;
  SPHX_TRY(validate_grid(c->P));
  SPHX_TRY(c->cell_start.alloc((size_t)gt + 1));
  SPHX_HIP(hipMemsetAsync(c->cell_start.p, 0, ((size_t)gt + 1) * sizeof(int), c->stream));
  int bits = 1;
  while ((1u << bits) <= (unsigned)gt) ++bits;   // keys in [0, gt]
  c->key_bits = bits;
  c->sorted_valid = false;
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

const char* sphx_last_error(void) { return g_err.c_str(); }

int sphx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* sphx_version(void) { return "sphx 0.1 (gfx950)"; }

// Device buffers for callers of the entry points that take device pointers
// (the reference host allocated them with cudaMalloc; an FFI host has no
// other way to).  kind: 1 host->device, 2 device->host, 3 device->device.
int sphx_dev_alloc(void** out, size_t bytes) {
  SPHX_CHECK(out, "sphx_dev_alloc", "null argument");
  *out = nullptr;
  if (bytes == 0) return 0;
  SPHX_HIP(hipMalloc(out, bytes));
  return 0;
}
int sphx_dev_free(void* p) {
  if (p) SPHX_HIP(hipFree(p));
  return 0;
}
int sphx_dev_copy(void* dst, const void* src, size_t bytes, int kind) {
  SPHX_CHECK(dst && src, "sphx_dev_copy", "null argument");
  SPHX_CHECK(kind >= 1 && kind <= 3, "sphx_dev_copy", "kind must be 1, 2 or 3");
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                              : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  SPHX_HIP(hipMemcpy(dst, src, bytes, k));
  return 0;
}

void sphx_default_options(sphx_SolverOptions* o) {
  std::memset(o, 0, sizeof(*o));
  o->omega = 0.1f;
  o->warm_start = 0.5f;
  o->full_support = 0;
  o->boundary_pressure = 0;
  o->tol = 1e-3f;
  o->min_iters = 2;
  o->max_iters = 1000;
  o->clamp_to_grid = 0;
  o->wcsph_kernel = 0;
  o->eos_B = -1.f;
  o->eos_gamma = 7.f;
  o->art_visc_alpha = 0.f;
  o->sound_speed = 0.f;
  o->drift_relax = 0.f;
  o->drift_diffusion = 0.f;
  o->drift_pressure_scale = 1.f;
  o->drift_diffusion_scale = 1.f;
  o->boussinesq = 0;
  o->rho_ref = 1000.f;
}

int sphx_create(sphx_context** out, const sphx_ParamCarrier* p) {
  SPHX_CHECK(out && p, "sphx_create", "null argument");
  *out = nullptr;
  SPHX_CHECK(p->maxNum > 0, "sphx_create", "maxNum must be positive");
  SPHX_CHECK(p->num >= 0 && p->num <= p->maxNum, "sphx_create", "num outside [0, maxNum]");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail("sphx_create", "no HIP device visible (the HIP path has no CPU fallback)");
  sphx_context* c = new sphx_context();
  c->P = *p;
  sphx_default_options(&c->O);
  c->cap = p->maxNum;
  c->num = p->num;
  int r = 0;
  auto chk = [&](int x) { if (x != 0 && r == 0) r = x; };
  if (hipGetDevice(&c->device) != hipSuccess) c->device = 0;
  if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return fail("sphx_create", "hipStreamCreate failed"); }
  size_t cap = (size_t)c->cap;
  for (int b = 0; b < 2; ++b) {
    chk(c->pos[b].alloc(cap));
    chk(c->vel[b].alloc(cap));
    chk(c->veleval[b].alloc(cap));
    chk(c->accel[b].alloc(cap));
    chk(c->phys[b].alloc(cap));
    chk(c->bornid[b].alloc(cap));
    chk(c->perm[b].alloc(cap));
  }
  chk(c->cold_color.alloc(cap));
  chk(c->cold_mat.alloc(cap * 36));
  chk(c->cold_X.alloc(cap));
  chk(c->cold_GK.alloc(cap));
  chk(c->key.alloc(cap));
  chk(c->idsort.alloc(cap));
  chk(c->key_tmp.alloc(cap));
  chk(c->val_tmp.alloc(cap));
  chk(c->mfid.alloc(cap));
  chk(c->dii.alloc(cap));
  chk(c->cst.alloc(cap));
  chk(c->dijpj.alloc(cap));
  chk(c->mqa.alloc(cap));
  chk(c->mqb.alloc(cap));
  chk(c->gh.alloc(cap));
  chk(c->press_l.alloc(cap));
  chk(c->press_l1.alloc(cap));
  chk(c->rho_adv.alloc(cap));
  chk(c->residue.alloc(cap));
  chk(c->sfac.alloc(cap));
  chk(c->red.alloc(4096));
  chk(c->solve_state.alloc(4));
  chk(c->solve_err.alloc(4));
  if (r == 0) r = alloc_grid(c);
  if (r == 0 && hipHostMalloc(reinterpret_cast<void**>(&c->host_pinned), 8 * sizeof(double)) != hipSuccess)
    r = fail("sphx_create", "hipHostMalloc failed");
  if (r == 0) {
    for (int b = 0; b < 2; ++b) {
      k_iota<<<blocks_for(c->cap, 256), 256, 0, c->stream>>>(c->perm[b].p, c->cap);
    }
    if (hipMemsetAsync(c->dii.p, 0, cap * sizeof(float4), c->stream) != hipSuccess ||
        hipMemsetAsync(c->cst.p, 0, cap * sizeof(float4), c->stream) != hipSuccess ||
        hipMemsetAsync(c->dijpj.p, 0, cap * sizeof(float4), c->stream) != hipSuccess ||
        hipMemsetAsync(c->press_l.p, 0, cap * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->press_l1.p, 0, cap * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->residue.p, 0, cap * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->rho_adv.p, 0, cap * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->sfac.p, 0, cap * sizeof(float), c->stream) != hipSuccess ||
        hipMemsetAsync(c->cold_mat.p, 0, cap * 36 * sizeof(float), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      r = fail("sphx_create", "device initialisation failed");
  }
  if (r != 0) {
    sphx_destroy(c);
    return r;
  }
  *out = c;
  return 0;
}

int sphx_destroy(sphx_context* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->mlp.tstream) {
    (void)hipStreamSynchronize(c->mlp.tstream);
    (void)hipStreamDestroy(c->mlp.tstream);
    (void)hipEventDestroy(c->mlp.ev_main);
    (void)hipEventDestroy(c->mlp.ev_train);
  }
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)timer_flush(c);
  for (hipEvent_t e : c->timer.pool) (void)hipEventDestroy(e);
  c->timer.pool.clear();
  (void)sphx_graph_release(c);
  for (int b = 0; b < 2; ++b) {
    c->pos[b].release(); c->vel[b].release(); c->veleval[b].release();
    c->accel[b].release(); c->phys[b].release(); c->bornid[b].release();
    c->perm[b].release();
  }
  c->cold_color.release(); c->cold_mat.release(); c->cold_X.release(); c->cold_GK.release();
  c->key.release(); c->idsort.release(); c->key_tmp.release(); c->val_tmp.release();
  c->block_hist.release(); c->cell_start.release(); c->mfid.release();
  c->dii.release(); c->cst.release(); c->dijpj.release(); c->mqa.release(); c->mqb.release(); c->gh.release();
  c->press_l.release(); c->press_l1.release(); c->rho_adv.release(); c->residue.release();
  c->sfac.release(); c->red.release(); c->solve_state.release(); c->solve_err.release();
  if (c->host_pinned) (void)hipHostFree(c->host_pinned);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int sphx_set_params(sphx_context* c, const sphx_ParamCarrier* p) {
  SPHX_CHECK(c && p, "sphx_set_params", "null argument");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(p->num >= 0 && p->num <= c->cap, "sphx_set_params", "num outside [0, capacity]");
  SPHX_TRY(validate_grid(*p));
  bool regrid = p->gridtotal != c->P.gridtotal;
  // the sorted keys and cell ranges describe the old particle count and
  // grid; neighbour loops must not read them until the grid is rebuilt
  const sphx_ParamCarrier& o = c->P;
  if (regrid || p->num != c->num || p->gridres.x != o.gridres.x || p->gridres.y != o.gridres.y ||
      p->gridres.z != o.gridres.z || p->gridmin.x != o.gridmin.x || p->gridmin.y != o.gridmin.y ||
      p->gridmin.z != o.gridmin.z || p->gridIdfac.x != o.gridIdfac.x ||
      p->gridIdfac.y != o.gridIdfac.y || p->gridIdfac.z != o.gridIdfac.z ||
      p->smoothradius != o.smoothradius || p->simscale != o.simscale)
    c->sorted_valid = false;
  c->P = *p;
  c->P.maxNum = c->cap;
  c->num = p->num;
  if (regrid) SPHX_TRY(alloc_grid(c));
  return 0;
}

int sphx_get_params(const sphx_context* c, sphx_ParamCarrier* p) {
  SPHX_CHECK(c && p, "sphx_get_params", "null argument");
  *p = c->P;
  p->num = c->num;
  return 0;
}

int sphx_set_options(sphx_context* c, const sphx_SolverOptions* o) {
  SPHX_CHECK(c && o, "sphx_set_options", "null argument");
  SPHX_CHECK(o->min_iters >= 1 && o->max_iters >= o->min_iters, "sphx_set_options",
             "need 1 <= min_iters <= max_iters");
  c->O = *o;
  return 0;
}

int sphx_sync(sphx_context* c) {
  SPHX_CHECK(c, "sphx_sync", "null context");
  if (c->mlp.tstream) SPHX_HIP(hipStreamSynchronize(c->mlp.tstream));   // training stream
  SPHX_HIP(hipStreamSynchronize(c->stream));
  return timer_flush(c);
}

int sphx_num_particles(const sphx_context* c) { return c ? c->num : -1; }

int sphx_upload(sphx_context* c, const sphx_displayPack* disp, const sphx_calculationPack* calc,
                int start, int count, int grow) {
  SPHX_CHECK(c && disp && calc, "sphx_upload", "null argument");
  SPHX_CHECK(start >= 0 && count >= 0 && start + count <= c->cap, "sphx_upload",
             "range exceeds capacity (maxNum)");
  SPHX_TRY(ensure_device(c));
  if (count == 0) return 0;
  SPHX_TRY(mpm_sync_cold(c));   // particles not in this upload keep their MPM state
  SPHX_TRY(materialize_cold(c));
  mpm_invalidate(c);
  positions_changed(c);
  c->sorted_valid = false;   // keys and cell ranges no longer match the particles
  std::vector<float4> pos(count), vel(count), ve(count), acc(count), phys(count), color(count), X(count);
  std::vector<int> born(count);
  std::vector<float> mat((size_t)count * 36);
  std::vector<float2> gk(count);
  for (int i = 0; i < count; ++i) {
    const sphx_displayPack& d = disp[i];
    const sphx_calculationPack& q = calc[i];
    float tw;
    std::memcpy(&tw, &d.type, 4);
    pos[i] = make_float4(d.pos.x, d.pos.y, d.pos.z, tw);
    vel[i] = make_float4(q.vel.x, q.vel.y, q.vel.z, q.visc);
    ve[i] = make_float4(q.veleval.x, q.veleval.y, q.veleval.z, 0.f);
    acc[i] = make_float4(q.accel.x, q.accel.y, q.accel.z, 0.f);
    phys[i] = make_float4(q.mass, q.dens, q.pressure, q.restdens);
    born[i] = q.bornid;
    color[i] = make_float4(d.color.x, d.color.y, d.color.z, d.color.w);
    X[i] = make_float4(q.X.x, q.X.y, q.X.z, 0.f);
    gk[i] = make_float2(q.solidG, q.solidK);
    float* m = &mat[(size_t)i * 36];
    std::memcpy(m, q.deformGrad.data, 36);
    std::memcpy(m + 9, q.stress.data, 36);
    std::memcpy(m + 18, q.B.data, 36);
    std::memcpy(m + 27, q.invA.data, 36);
  }
  int b = c->cur;
  size_t off = (size_t)start;
  SPHX_HIP(hipMemcpy(c->pos[b].p + off, pos.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->vel[b].p + off, vel.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->veleval[b].p + off, ve.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->accel[b].p + off, acc.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->phys[b].p + off, phys.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->bornid[b].p + off, born.data(), count * sizeof(int), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->cold_color.p + off, color.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->cold_X.p + off, X.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->cold_GK.p + off, gk.data(), count * sizeof(float2), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->cold_mat.p + off * 36, mat.data(), (size_t)count * 36 * sizeof(float),
                     hipMemcpyHostToDevice));
  if (grow && start + count > c->num) {
    c->num = start + count;
    c->P.num = c->num;
  }
  return 0;
}

int sphx_download(sphx_context* c, sphx_displayPack* disp, sphx_calculationPack* calc, int start,
                  int count) {
  SPHX_CHECK(c, "sphx_download", "null context");
  SPHX_CHECK(start >= 0 && count >= 0 && start + count <= c->num, "sphx_download",
             "range exceeds num");
  SPHX_TRY(ensure_device(c));
  if (count == 0) return 0;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  int n = c->num;
  int b = c->cur;
  SPHX_TRY(mpm_sync_cold(c));
  // gather cold fields into slot order on the device
  DevBuf<float4> color, X;
  DevBuf<float> mat;
  DevBuf<float2> gk;
  SPHX_TRY(color.alloc(n));
  k_gather<float4><<<blocks_for(n, 256), 256, 0, c->stream>>>(c->PERM(), n, c->cold_color.p, color.p, 1);
  if (calc) {
    SPHX_TRY(X.alloc(n));
    SPHX_TRY(mat.alloc((size_t)n * 36));
    SPHX_TRY(gk.alloc(n));
    k_gather<float4><<<blocks_for(n, 256), 256, 0, c->stream>>>(c->PERM(), n, c->cold_X.p, X.p, 1);
    k_gather<float2><<<blocks_for(n, 256), 256, 0, c->stream>>>(c->PERM(), n, c->cold_GK.p, gk.p, 1);
    k_gather<float><<<blocks_for(n, 256), 256, 0, c->stream>>>(c->PERM(), n, c->cold_mat.p, mat.p, 36);
  }
  SPHX_HIP(hipGetLastError());
  SPHX_HIP(hipStreamSynchronize(c->stream));
  size_t off = (size_t)start;
  std::vector<float4> pos(count), vcol(count);
  SPHX_HIP(hipMemcpy(pos.data(), c->pos[b].p + off, count * sizeof(float4), hipMemcpyDeviceToHost));
  SPHX_HIP(hipMemcpy(vcol.data(), color.p + off, count * sizeof(float4), hipMemcpyDeviceToHost));
  if (disp) {
    for (int i = 0; i < count; ++i) {
      sphx_displayPack& d = disp[i];
      d.pos = {pos[i].x, pos[i].y, pos[i].z};
      d.color = {vcol[i].x, vcol[i].y, vcol[i].z, vcol[i].w};
      std::memcpy(&d.type, &pos[i].w, 4);
    }
  }
  if (calc) {
    std::vector<float4> vel(count), ve(count), acc(count), phys(count), vX(count);
    std::vector<int> born(count);
    std::vector<float> vmat((size_t)count * 36);
    std::vector<float2> vgk(count);
    SPHX_HIP(hipMemcpy(vel.data(), c->vel[b].p + off, count * sizeof(float4), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(ve.data(), c->veleval[b].p + off, count * sizeof(float4), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(acc.data(), c->accel[b].p + off, count * sizeof(float4), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(phys.data(), c->phys[b].p + off, count * sizeof(float4), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(born.data(), c->bornid[b].p + off, count * sizeof(int), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(vX.data(), X.p + off, count * sizeof(float4), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(vgk.data(), gk.p + off, count * sizeof(float2), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(vmat.data(), mat.p + off * 36, (size_t)count * 36 * sizeof(float),
                       hipMemcpyDeviceToHost));
    for (int i = 0; i < count; ++i) {
      sphx_calculationPack& q = calc[i];
      q.vel = {vel[i].x, vel[i].y, vel[i].z};
      q.visc = vel[i].w;
      q.veleval = {ve[i].x, ve[i].y, ve[i].z};
      q.accel = {acc[i].x, acc[i].y, acc[i].z};
      q.mass = phys[i].x;
      q.dens = phys[i].y;
      q.pressure = phys[i].z;
      q.restdens = phys[i].w;
      q.bornid = born[i];
      q.X = {vX[i].x, vX[i].y, vX[i].z};
      q.solidG = vgk[i].x;
      q.solidK = vgk[i].y;
      const float* m = &vmat[(size_t)i * 36];
      std::memcpy(q.deformGrad.data, m, 36);
      std::memcpy(q.stress.data, m + 9, 36);
      std::memcpy(q.B.data, m + 18, 36);
      std::memcpy(q.invA.data, m + 27, 36);
    }
  }
  return 0;
}

int sphx_timing_enable(sphx_context* c, int on) {
  SPHX_CHECK(c, "sphx_timing_enable", "null context");
  SPHX_TRY(timer_flush(c));
  c->timing = on != 0;
  return 0;
}

// Time only the named kernel (null or empty: every kernel).
int sphx_timing_only(sphx_context* c, const char* kernel) {
  SPHX_CHECK(c, "sphx_timing_only", "null context");
  SPHX_TRY(timer_flush(c));
  c->timer.only = kernel ? kernel : "";
  return 0;
}

int sphx_timing_reset(sphx_context* c) {
  SPHX_CHECK(c, "sphx_timing_reset", "null context");
  SPHX_TRY(timer_flush(c));
  c->timer.acc.clear();
  return 0;
}

int sphx_timing_get(sphx_context* c, const char* kernel, double* total_ms, int* launches) {
  SPHX_CHECK(c && kernel, "sphx_timing_get", "null argument");
  SPHX_TRY(timer_flush(c));
  auto it = c->timer.acc.find(kernel);
  double t = 0.0;
  int l = 0;
  if (it != c->timer.acc.end()) { t = it->second.first; l = it->second.second; }
  if (total_ms) *total_ms = t;
  if (launches) *launches = l;
  return 0;
}

}  // extern "C"
