// Drift-flux mixture model: drift velocities, volume-fraction advection,
// fraction correction and the drift-stress / buoyancy acceleration.
//
// Formula specification: SPH_Solver/fluids/multiphase.cu (not compiled in
// the reference; it targets a removed buffer layout):
//   contributeDriftVel / mfComputeDriftVel     :19-166
//   contributeAlphaChange / mfComputeAlphaAdvance :169-272
//   mfComputeCorrection                        :275-349
//   drift-stress term of contributeForce_projectu :555-562
// Kernel: cubic spline with support 2h (the IISPH kernel of this build)
// over the step's neighbour list; V_j = m_j / rho_j; dist = x_i - x_j so
// V_j gradW_ij = dist * V_j * f(r) as in the spec's cmterm * dist.
//
// Per-particle multiphase state (cell-sorted, permuted with the particles):
//   alpha  = {alpha_0, alpha_1, alpha_2, rho_m}   (fractions before the step)
//   mforce = {g - a (last step), 0}               ("mforce" of :99)
// Per-slot scratch: drift[k] = {u_mk, 0}, alpha_new.
#include "nbr.hpp"

namespace sphx {

struct MFP {
  int nphase;
  float rho[3], nu[3];
  float relax;         // tau = relax / nu_i (:116)
  float sigma;         // diffusion coefficient (:94)
  float cpress;        // cont * cont2 (:133)
  float cdiff;         // cont * cont1 (:162)
  int boussinesq;
  float rho_ref;
};

constexpr int MF_THREADS = 256;

__device__ __forceinline__ float comp(const float4& a, int k) {
  return k == 0 ? a.x : (k == 1 ? a.y : a.z);
}

// Packed neighbour records of the drift-flux loops: rec = {x, y, z, V =
// m / rho} (V = -1 marks a boundary particle, which those loops skip) and
// reca = {alpha_0, alpha_1, alpha_2, p}.  One 16-byte gather per record
// instead of pos + phys (+ alpha).
__global__ void k_mf_pack(int n, const float4* __restrict__ pos, const float4* __restrict__ phys,
                          const float4* __restrict__ alpha, float4* __restrict__ rec,
                          float4* __restrict__ reca) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = pos[i], q = phys[i];
  rec[i] = make_float4(p.x, p.y, p.z, type_of(p) == SPHX_TYPE_BOUNDARY ? -1.f : q.x / q.y);
  if (reca) {
    const float4 a = alpha[i];
    reca[i] = make_float4(a.x, a.y, a.z, q.z);
  }
}

// mfComputeDriftVel (multiphase.cu:19-166)
__global__ __launch_bounds__(MF_THREADS) void k_mf_drift(
    KP k, MFP m, NbrList L, const uint32_t* __restrict__ key, const int* __restrict__ cs,
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float4* __restrict__ phys, const float4* __restrict__ alpha,
    const float4* __restrict__ mforce, const float4* __restrict__ dcorr,
    const float4* __restrict__ rec, const float4* __restrict__ reca,
    float4* __restrict__ drift0, float4* __restrict__ drift1, float4* __restrict__ drift2) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  uint32_t gc = key[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY || (int)gc == k.gridtotal) return;
  const float4 qi = phys[i];
  const float4 ai = alpha[i];
  const float ipress = qi.z;
  const float relax = m.relax / vel[i].w;       // relax_coef /= mf_visc[i]
  const float invdens = 1.f / ai.w;             // 1 / rho_m,i
  float apre[3], conc[3];
  float3 dv[3], ag[3];
  for (int f = 0; f < 3; ++f) {
    apre[f] = f < m.nphase ? comp(ai, f) : 0.f;
    conc[f] = apre[f] * m.rho[f] * invdens;     // mass concentration (:122)
    dv[f] = f3(0.f, 0.f, 0.f);
    ag[f] = f3(0.f, 0.f, 0.f);
  }
  // Branch-free loop over every slot of each batch: a boundary neighbour
  // (V_j < 0) or a slot past the count gets an infinite distance, which
  // the support test turns into a zero gradient.  Every x_ij below is
  // multiplied by the gradient factor cm, so both are taken in the stored
  // position units (grad W = x_u g, solve_gradf).
  const SolveGrad sg = solve_grad(k, true);
  struct Rec { float4 r, a; };
  for_each_listed_m<NBR_BATCH, true>(
      k, L, cs, gc, i, [&](uint32_t j) { return Rec{load_at(rec, j * 16u), load_at(reca, j * 16u)}; },
      [&](uint32_t, const Rec& o, bool valid) {
    const float4 rj = o.r;
    const float dx = pi.x - rj.x, dy = pi.y - rj.y, dz = pi.z - rj.z;
    const float dsq = (valid && rj.w >= 0.f) ? pair_dsq(dx, dy, dz) : INFINITY;
    const float4 aj = o.a;                                     // alpha_j, p_j
    const float cm = solve_gradf(sg, dsq) * rj.w;              // V_j f(r)
    float pg[3];
    float3 pgs = f3(0.f, 0.f, 0.f);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      if (f >= m.nphase) break;
      float ajf = comp(aj, f);
      pg[f] = cm * (-apre[f] * ipress + ajf * aj.w);           // :62
      pgs.x += pg[f] * dx * conc[f];
      pgs.y += pg[f] * dy * conc[f];
      pgs.z += pg[f] * dz * conc[f];
      float ga = (ajf - apre[f]) * cm;                         // :67
      ag[f].x += ga * dx; ag[f].y += ga * dy; ag[f].z += ga * dz;
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) {                              // :70-73
      if (f >= m.nphase) break;
      dv[f].x -= relax * (pg[f] * dx - pgs.x);
      dv[f].y -= relax * (pg[f] * dy - pgs.y);
      dv[f].z -= relax * (pg[f] * dz - pgs.z);
    }
  });
  // :138-163
  float dsum = 0.f;
  float3 ssum = f3(0.f, 0.f, 0.f), st[3] = {f3(0.f, 0.f, 0.f), f3(0.f, 0.f, 0.f), f3(0.f, 0.f, 0.f)};
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (f >= m.nphase) break;
    float t = apre[f];
    dsum += t * m.rho[f] * m.rho[f] * invdens;
    st[f] = t > 0.0001f ? mul(ag[f], 1.f / t) : f3(0.f, 0.f, 0.f);
    ssum = add(ssum, mul(st[f], t * m.rho[f] * invdens));
  }
  const float4 acc = mforce[i];
  // learned closure: dispersed phase 1 gets +d, phase 0 -(c_1/c_0) d so
  // that sum_k c_k u_mk stays zero
  float3 dc = f3(0.f, 0.f, 0.f);
  float dc0 = 0.f;
  if (dcorr && m.nphase >= 2 && conc[0] > 1e-6f) {
    float4 d = dcorr[i];
    dc = f3(d.x, d.y, d.z);
    dc0 = -conc[1] / conc[0];
  }
  float4* out[3] = {drift0, drift1, drift2};
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (f >= m.nphase) break;
    float tdiff = (m.rho[f] - dsum) * relax;
    float3 u = mul(dv[f], m.cpress);
    u.x += acc.x * tdiff - m.cdiff * m.sigma * (st[f].x - ssum.x);
    u.y += acc.y * tdiff - m.cdiff * m.sigma * (st[f].y - ssum.y);
    u.z += acc.z * tdiff - m.cdiff * m.sigma * (st[f].z - ssum.z);
    float cf = f == 1 ? 1.f : (f == 0 ? dc0 : 0.f);
    u = add(u, mul(dc, cf));
    out[f][i] = make_float4(u.x, u.y, u.z, apre[f]);   // .w: alpha_f for the next loops
  }
}

// mfComputeAlphaAdvance (multiphase.cu:169-272); advects with the particle
// velocity `vel` (the spec's mveleval).  The same neighbour pass also sums
// the drift-stress acceleration of contributeForce_projectu (:555-562),
// which reads a subset of its records (positions, V_j, drifts and the
// fractions before the step) and none of its outputs; it is stored in
// `stress` for the force stage, which then only adds it (and buoyancy).
// NP = the number of phases (1-3): each neighbour record holds NP drifts.
template <int NP>
__global__ __launch_bounds__(MF_THREADS) void k_mf_alpha(
    KP k, MFP m, NbrList L, const uint32_t* __restrict__ key, const int* __restrict__ cs,
    const float4* __restrict__ pos, const float4* __restrict__ vel,
    const float4* __restrict__ phys, const float4* __restrict__ alpha,
    const float4* __restrict__ drift0, const float4* __restrict__ drift1,
    const float4* __restrict__ drift2, const float4* __restrict__ rec,
    float4* __restrict__ alpha_new, float4* __restrict__ stress) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  uint32_t gc = key[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY || (int)gc == k.gridtotal) {
    alpha_new[i] = alpha[i];
    return;
  }
  const float4 vi = vel[i];
  const float4 ai = alpha[i];
  const float idens = 1.f / phys[i].y;
  const float4* dr[3] = {drift0, drift1, drift2};
  float apre[3], ach[3];
  float3 ui[3];
  for (int f = 0; f < 3; ++f) {
    apre[f] = f < NP ? comp(ai, f) : 0.f;
    ach[f] = 0.f;
    ui[f] = f < NP ? xyz(dr[f][i]) : f3(0.f, 0.f, 0.f);
  }
  float fx = 0.f, fy = 0.f, fz = 0.f;
  // Branch-free as the drift loop: boundary neighbours and slots past the
  // count get an infinite distance (zero gradient); every x_ij is
  // multiplied by cm, so both are in the stored position units.
  const SolveGrad sg = solve_grad(k, true);
  struct Rec { float4 r, v, u[NP]; };
  for_each_listed_m<NBR_BATCH, false>(k, L, cs, gc, i, [&](uint32_t j) {
    Rec o;
    o.r = load_at(rec, j * 16u);
    o.v = load_at(vel, j * 16u);
#pragma unroll
    for (int f = 0; f < NP; ++f) o.u[f] = load_at(dr[f], j * 16u);
    return o;
  }, [&](uint32_t, const Rec& o, bool valid) {
    const float4 rj = o.r;
    const float dx = pi.x - rj.x, dy = pi.y - rj.y, dz = pi.z - rj.z;
    const float dsq = (valid && rj.w >= 0.f) ? pair_dsq(dx, dy, dz) : INFINITY;
    const float4 vj = o.v;
    const float cm = solve_gradf(sg, dsq) * rj.w;
    float vmrd = (vj.x - vi.x) * dx + (vj.y - vi.y) * dy + (vj.z - vi.z) * dz;
#pragma unroll
    for (int f = 0; f < NP; ++f) {
      const float4 uj = o.u[f];                                          // u_f, alpha_f
      float ajf = uj.w;
      ach[f] -= 0.5f * cm * (ajf + apre[f]) * vmrd;                        // :207
      float vkx = ajf * uj.x + apre[f] * ui[f].x;
      float vky = ajf * uj.y + apre[f] * ui[f].y;
      float vkz = ajf * uj.z + apre[f] * ui[f].z;
      ach[f] -= cm * (vkx * dx + vky * dy + vkz * dz);                     // :209-212
      // drift stress (:555-562)
      float sj = cm * (uj.x * dx + uj.y * dy + uj.z * dz) * uj.w;
      float si = cm * (ui[f].x * dx + ui[f].y * dy + ui[f].z * dz) * apre[f];
      float w = m.rho[f] * idens;
      fx += (sj * uj.x + si * ui[f].x) * w;
      fy += (sj * uj.y + si * ui[f].y) * w;
      fz += (sj * uj.z + si * ui[f].z) * w;
    }
  });
  stress[i] = make_float4(fx, fy, fz, 0.f);
  float o[3];
  for (int f = 0; f < 3; ++f) {
    float d = ach[f] * k.dt;
    if (d < -0.99f) d = -0.99f;                                            // :258-261
    o[f] = f < NP ? d + apre[f] : 0.f;
  }
  alpha_new[i] = make_float4(o[0], o[1], o[2], ai.w);
}

// mfComputeCorrection (multiphase.cu:275-349): clamp at 0, normalise, or
// one-hot the phase that dominated before the step; mixture density and
// viscosity.  (The WCSPH pressure modification of :330-331 has no IISPH
// counterpart: the pressure solve enforces the density.)  One particle;
// returns its mixture density rho_m.
__device__ __forceinline__ float mf_correct_one(const MFP& m, int i,
                                                const float4* __restrict__ alpha_pre,
                                                float4* __restrict__ alpha_new,
                                                float4* __restrict__ vel) {
  float4 a = alpha_new[i];
  float4 ap = alpha_pre[i];
  float v[3] = {a.x, a.y, a.z};
  float sum = 0.f;
  for (int f = 0; f < m.nphase; ++f) {
    if (v[f] < 0.f) v[f] = 0.f;
    sum += v[f];
  }
  int flag = sum > 0.f;
  int maxcat = 0;
  for (int f = 1; f < m.nphase; ++f)
    if (comp(ap, f) > comp(ap, maxcat)) maxcat = f;
  float inv = 1.f / (flag ? sum : 1.f);
  float rho_m = 0.f, nu_m = 0.f;
  for (int f = 0; f < m.nphase; ++f) {
    v[f] = flag ? v[f] * inv : (f == maxcat ? 1.f : 0.f);
    rho_m += v[f] * m.rho[f];
    nu_m += v[f] * m.nu[f];
  }
  for (int f = m.nphase; f < 3; ++f) v[f] = 0.f;
  alpha_new[i] = make_float4(v[0], v[1], v[2], rho_m);
  vel[i].w = nu_m;
  return rho_m;
}

__global__ void k_mf_correction(KP k, MFP m, const float4* __restrict__ pos,
                                const uint32_t* __restrict__ key,
                                const float4* __restrict__ alpha_pre,
                                float4* __restrict__ alpha_new, float4* __restrict__ vel) {
  int i = xcd_index();
  if (i >= k.n) return;
  if (type_of(pos[i]) == SPHX_TYPE_BOUNDARY || (int)key[i] == k.gridtotal) return;
  mf_correct_one(m, i, alpha_pre, alpha_new, vel);
}

// Non-pressure mixture acceleration added to the advected velocity:
// drift stress (contributeForce_projectu, multiphase.cu:555-562) with the
// fractions before the step, plus Boussinesq buoyancy (rho_m - rho_ref) /
// rho_ref * g.  Stores a_mf in `amf` for the record of the total.
__global__ __launch_bounds__(MF_THREADS) void k_mf_force(
    KP k, MFP m, NbrList L, const uint32_t* __restrict__ key, const int* __restrict__ cs,
    const float4* __restrict__ pos, const float4* __restrict__ phys,
    const float4* __restrict__ alpha, const float4* __restrict__ alpha_new,
    const float4* __restrict__ drift0, const float4* __restrict__ drift1,
    const float4* __restrict__ drift2, const float4* __restrict__ rec,
    float4* __restrict__ veleval, float4* __restrict__ amf) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  uint32_t gc = key[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY || (int)gc == k.gridtotal) return;
  const float4 qi = phys[i];
  const float idens = 1.f / qi.y;
  const float4 ai = alpha[i];
  const float4* dr[3] = {drift0, drift1, drift2};
  float3 ui[3];
  float apre[3];
  for (int f = 0; f < 3; ++f) {
    ui[f] = f < m.nphase ? xyz(dr[f][i]) : f3(0.f, 0.f, 0.f);
    apre[f] = f < m.nphase ? comp(ai, f) : 0.f;
  }
  float fx = 0.f, fy = 0.f, fz = 0.f;
  struct Rec { float4 r, u[3]; };
  const bool three = m.nphase > 2;
  // the fraction-advection loop's arithmetic (branch-free, stored position
  // units), so its fused drift stress and this pass agree to rounding
  const SolveGrad sg = solve_grad(k, true);
  for_each_listed_m<NBR_BATCH, true>(k, L, cs, gc, i, [&](uint32_t j) {
    Rec o;
    o.r = load_at(rec, j * 16u);
    o.u[0] = load_at(drift0, j * 16u);
    o.u[1] = load_at(drift1, j * 16u);
    o.u[2] = three ? load_at(drift2, j * 16u) : make_float4(0.f, 0.f, 0.f, 0.f);
    return o;
  }, [&](uint32_t, const Rec& o, bool valid) {
    const float4 rj = o.r;
    const float dx = pi.x - rj.x, dy = pi.y - rj.y, dz = pi.z - rj.z;
    const float dsq = (valid && rj.w >= 0.f) ? pair_dsq(dx, dy, dz) : INFINITY;
    const float cm = solve_gradf(sg, dsq) * rj.w;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      if (f >= m.nphase) break;
      const float4 uj = o.u[f];                                          // u_f, alpha_f
      float sj = cm * (uj.x * dx + uj.y * dy + uj.z * dz) * uj.w;
      float si = cm * (ui[f].x * dx + ui[f].y * dy + ui[f].z * dz) * apre[f];
      float w = m.rho[f] * idens;
      fx += (sj * uj.x + si * ui[f].x) * w;
      fy += (sj * uj.y + si * ui[f].y) * w;
      fz += (sj * uj.z + si * ui[f].z) * w;
    }
  });
  if (m.boussinesq) {
    float b = (alpha_new[i].w - m.rho_ref) / m.rho_ref;
    fx += b * k.g.x; fy += b * k.g.y; fz += b * k.g.z;
  }
  float4 ve = veleval[i];
  veleval[i] = make_float4(ve.x + fx * k.dt, ve.y + fy * k.dt, ve.z + fz * k.dt, ve.w);
  amf[i] = make_float4(fx, fy, fz, 0.f);
}

// The force stage after a fused alpha pass: drift stress from `stress`,
// plus buoyancy; same operations and order as k_mf_force's tail.
// The drift stress summed by the alpha pass plus buoyancy, added to one
// particle's advected velocity; rho_m = its corrected mixture density.
__device__ __forceinline__ void mf_force_add_one(const KP& k, const MFP& m, int i, float rho_m,
                                                 const float4* __restrict__ stress,
                                                 float4* __restrict__ veleval,
                                                 float4* __restrict__ amf) {
  const float4 st = stress[i];
  float fx = st.x, fy = st.y, fz = st.z;
  if (m.boussinesq) {
    float b = (rho_m - m.rho_ref) / m.rho_ref;
    fx += b * k.g.x; fy += b * k.g.y; fz += b * k.g.z;
  }
  float4 ve = veleval[i];
  veleval[i] = make_float4(ve.x + fx * k.dt, ve.y + fy * k.dt, ve.z + fz * k.dt, ve.w);
  amf[i] = make_float4(fx, fy, fz, 0.f);
}

__global__ void k_mf_force_add(KP k, MFP m, const float4* __restrict__ pos,
                               const uint32_t* __restrict__ key,
                               const float4* __restrict__ alpha_new,
                               const float4* __restrict__ stress,
                               float4* __restrict__ veleval, float4* __restrict__ amf) {
  int i = xcd_index();
  if (i >= k.n) return;
  if (type_of(pos[i]) == SPHX_TYPE_BOUNDARY || (int)key[i] == k.gridtotal) return;
  mf_force_add_one(k, m, i, m.boussinesq ? alpha_new[i].w : 0.f, stress, veleval, amf);
}

// k_mf_correction then k_mf_force_add (same particles skipped), in one pass.
// `cnt`, when given, is the step's neighbour list counts (NBR_INVALID marks
// boundary and out-of-grid slots): one 4-byte read instead of pos and key.
__global__ void k_mf_correct_force(KP k, MFP m, const int* __restrict__ cnt,
                                   const float4* __restrict__ pos,
                                   const uint32_t* __restrict__ key,
                                   const float4* __restrict__ alpha_pre,
                                   float4* __restrict__ alpha_new, float4* __restrict__ vel,
                                   const float4* __restrict__ stress,
                                   float4* __restrict__ veleval, float4* __restrict__ amf) {
  int i = xcd_index();
  if (i >= k.n) return;
  if (cnt ? cnt[i] == NBR_INVALID
          : (type_of(pos[i]) == SPHX_TYPE_BOUNDARY || (int)key[i] == k.gridtotal))
    return;
  const float rho_m = mf_correct_one(m, i, alpha_pre, alpha_new, vel);
  mf_force_add_one(k, m, i, rho_m, stress, veleval, amf);
}

__global__ void k_mf_reorder(const uint32_t* __restrict__ skey, const uint32_t* __restrict__ idsort,
                             int n, uint32_t sentinel, const float4* __restrict__ a_in,
                             float4* __restrict__ a_out, const float4* __restrict__ f_in,
                             float4* __restrict__ f_out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int src = skey[s] != sentinel ? (int)idsort[s] : s;
  a_out[s] = a_in[src];
  f_out[s] = f_in[src];
}

MFP make_mfp(const sphx_context* c) {
  MFP m;
  m.nphase = c->nphase;
  for (int f = 0; f < 3; ++f) { m.rho[f] = c->mf_rho[f]; m.nu[f] = c->mf_nu[f]; }
  m.relax = c->O.drift_relax;
  m.sigma = c->O.drift_diffusion;
  m.cpress = c->O.drift_pressure_scale;
  m.cdiff = c->O.drift_diffusion_scale;
  m.boussinesq = c->O.boussinesq;
  m.rho_ref = c->O.rho_ref;
  return m;
}

__global__ void k_mf_copy(const float4* __restrict__ a, float4* __restrict__ b, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Make the corrected fractions the current ones (alpha_new is per slot).
// A copy, not a pointer swap: the buffers bound to every launch depend only
// on the particle-buffer parity, so a captured step graph stays valid.
int iisph_density_dii(sphx_context* c);

int mf_commit(sphx_context* c) {
  if (!c->mf_pending) return 0;
  c->mf_stress_ok = false;   // the fractions before the step change
  if (c->num > 0) {
    k_mf_copy<<<blocks_for(c->num, 256), 256, 0, c->stream>>>(c->mf_alpha_new.p,
                                                              c->mf_alpha[c->cur].p, c->num);
    SPHX_HIP(hipGetLastError());
  }
  c->mf_pending = false;
  return 0;
}

// Called by the grid reorder after it gathered the core fields from buffer
// `a` into buffer `b`; pending corrected fractions are gathered directly.
int mf_reorder(sphx_context* c, int a, int b) {
  if (c->nphase == 0 || c->num == 0) return 0;
  const float4* src = c->mf_pending ? c->mf_alpha_new.p : c->mf_alpha[a].p;
  c->mf_pending = false;
  int n = c->num;
  k_mf_reorder<<<blocks_for(n, 256), 256, 0, c->stream>>>(
      c->key.p, c->idsort.p, n, (uint32_t)c->P.gridtotal, src, c->mf_alpha[b].p,
      c->mf_force[a].p, c->mf_force[b].p);
  SPHX_HIP(hipGetLastError());
  return 0;
}

int nbr_list(sphx_context* c, NbrList* out);

// Drift velocities of the current state (fractions alpha[cur], pressure
// phys.z, g - a of the last step) into out0..2, plus the optional learned
// correction.  Does not commit pending fractions.
int mf_drift_into(sphx_context* c, const float4* dcorr, float4* out0, float4* out1,
                  float4* out2) {
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  MFP m = make_mfp(c);
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  // records with this call's pressures and fractions (also when the
  // recorder re-evaluates the drift after the solve)
  k_mf_pack<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(k.n, c->POS(), c->PHYS(),
                                                          c->mf_alpha[c->cur].p, c->mf_rec.p,
                                                          c->mf_reca.p);
  c->mf_rec_ok = true;
  k_mf_drift<<<blocks_for(k.n, c->tune[T_MF_DRIFT]), c->tune[T_MF_DRIFT], 0, c->stream>>>(
      k, m, L, c->key.p, c->cell_start.p, c->POS(), c->VEL(), c->PHYS(), c->mf_alpha[c->cur].p,
      c->mf_force[c->cur].p, dcorr, c->mf_rec.p, c->mf_reca.p, out0, out1, out2);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// {x, y, z, m / rho} for the loops after the drift pass, when a caller ran
// them without it since the positions or densities last changed.
static int mf_need_rec(sphx_context* c) {
  if (c->mf_rec_ok || c->num == 0) return 0;
  k_mf_pack<<<blocks_for(c->num, 256), 256, 0, c->stream>>>(c->num, c->POS(), c->PHYS(), nullptr,
                                                             c->mf_rec.p, nullptr);
  SPHX_HIP(hipGetLastError());
  c->mf_rec_ok = true;
  return 0;
}

static int need_mf(sphx_context* c, const char* who) {
  SPHX_CHECK(c, who, "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(c->nphase > 0, who, "multiphase not set up (call sphx_mf_setup)");
  SPHX_CHECK(c->sorted_valid, who, "grid not built (call sphx_grid_build first)");
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

int sphx_mf_setup(sphx_context* c, int nphase, const float* rho, const float* nu) {
  SPHX_CHECK(c && rho && nu, "sphx_mf_setup", "null argument");
  SPHX_CHECK(nphase >= 1 && nphase <= SPHX_MAX_FLUIDNUM, "sphx_mf_setup",
             "nphase must be in [1, 3] (MAX_FLUIDNUM)");
  SPHX_TRY(ensure_device(c));
  size_t cap = (size_t)c->cap;
  for (int b = 0; b < 2; ++b) {
    SPHX_TRY(c->mf_alpha[b].alloc(cap));
    SPHX_TRY(c->mf_force[b].alloc(cap));
  }
  SPHX_TRY(c->mf_alpha_new.alloc(cap));
  SPHX_TRY(c->mf_amf.alloc(cap));
  SPHX_TRY(c->mf_stress.alloc(cap));
  c->mf_stress_ok = false;
  for (int f = 0; f < 3; ++f) SPHX_TRY(c->mf_drift[f].alloc(cap));
  SPHX_TRY(c->mf_rec.alloc(cap));
  SPHX_TRY(c->mf_reca.alloc(cap));
  c->mf_rec_ok = false;
  for (int b = 0; b < 2; ++b) {
    SPHX_HIP(hipMemsetAsync(c->mf_alpha[b].p, 0, cap * sizeof(float4), c->stream));
    SPHX_HIP(hipMemsetAsync(c->mf_force[b].p, 0, cap * sizeof(float4), c->stream));
  }
  SPHX_HIP(hipMemsetAsync(c->mf_alpha_new.p, 0, cap * sizeof(float4), c->stream));
  SPHX_HIP(hipMemsetAsync(c->mf_amf.p, 0, cap * sizeof(float4), c->stream));
  for (int f = 0; f < 3; ++f)
    SPHX_HIP(hipMemsetAsync(c->mf_drift[f].p, 0, cap * sizeof(float4), c->stream));
  c->nphase = nphase;
  for (int f = 0; f < 3; ++f) {
    c->mf_rho[f] = f < nphase ? rho[f] : 0.f;
    c->mf_nu[f] = f < nphase ? nu[f] : 0.f;
  }
  c->mf_pending = false;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// alpha: [count][nphase] volume fractions for slots [start, start+count);
// rho_m = sum alpha_k rho_k, visc = sum alpha_k nu_k, mforce = g.
int sphx_mf_upload(sphx_context* c, const float* alpha, int start, int count) {
  SPHX_CHECK(c && alpha, "sphx_mf_upload", "null argument");
  SPHX_CHECK(c->nphase > 0, "sphx_mf_upload", "multiphase not set up");
  SPHX_CHECK(start >= 0 && count >= 0 && start + count <= c->num, "sphx_mf_upload",
             "range exceeds num");
  SPHX_TRY(ensure_device(c));
  c->mf_stress_ok = false;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  SPHX_TRY(mf_commit(c));
  if (count == 0) return 0;
  int np = c->nphase;
  std::vector<float4> a(count), f(count), v(count);
  SPHX_HIP(hipMemcpy(v.data(), c->VEL() + start, count * sizeof(float4), hipMemcpyDeviceToHost));
  std::vector<float4> ps(count);
  SPHX_HIP(hipMemcpy(ps.data(), c->POS() + start, count * sizeof(float4), hipMemcpyDeviceToHost));
  for (int i = 0; i < count; ++i) {
    float al[3] = {0, 0, 0}, rm = 0.f, nm = 0.f;
    for (int k = 0; k < np; ++k) {
      al[k] = alpha[(size_t)i * np + k];
      rm += al[k] * c->mf_rho[k];
      nm += al[k] * c->mf_nu[k];
    }
    a[i] = make_float4(al[0], al[1], al[2], rm);
    f[i] = make_float4(c->P.gravity.x, c->P.gravity.y, c->P.gravity.z, 0.f);
    int t;
    std::memcpy(&t, &ps[i].w, 4);
    if (type_bits(t) != SPHX_TYPE_BOUNDARY) v[i].w = nm;
  }
  SPHX_HIP(hipMemcpy(c->mf_alpha[c->cur].p + start, a.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->mf_force[c->cur].p + start, f.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  SPHX_HIP(hipMemcpy(c->VEL() + start, v.data(), count * sizeof(float4), hipMemcpyHostToDevice));
  return 0;
}

// Any pointer may be NULL.  alpha [num][nphase] (corrected fractions when a
// correction is pending), rho_m [num], drift [num][nphase][3], mforce [num][3].
int sphx_mf_download(sphx_context* c, float* alpha, float* rho_m, float* drift, float* mforce) {
  SPHX_CHECK(c, "sphx_mf_download", "null context");
  SPHX_CHECK(c->nphase > 0, "sphx_mf_download", "multiphase not set up");
  SPHX_TRY(ensure_device(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  int n = c->num, np = c->nphase;
  if (n == 0) return 0;
  std::vector<float4> t(n);
  if (alpha || rho_m) {
    const float4* src = c->mf_pending ? c->mf_alpha_new.p : c->mf_alpha[c->cur].p;
    SPHX_HIP(hipMemcpy(t.data(), src, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
      float v[3] = {t[i].x, t[i].y, t[i].z};
      if (alpha) for (int k = 0; k < np; ++k) alpha[(size_t)i * np + k] = v[k];
      if (rho_m) rho_m[i] = t[i].w;
    }
  }
  if (drift) {
    for (int k = 0; k < np; ++k) {
      SPHX_HIP(hipMemcpy(t.data(), c->mf_drift[k].p, n * sizeof(float4), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) {
        float* d = drift + ((size_t)i * np + k) * 3;
        d[0] = t[i].x; d[1] = t[i].y; d[2] = t[i].z;
      }
    }
  }
  if (mforce) {
    SPHX_HIP(hipMemcpy(t.data(), c->mf_force[c->cur].p, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) { mforce[3 * i] = t[i].x; mforce[3 * i + 1] = t[i].y; mforce[3 * i + 2] = t[i].z; }
  }
  return 0;
}

int sphx_mf_drift_velocity(sphx_context* c) {
  SPHX_TRY(need_mf(c, "sphx_mf_drift_velocity"));
  SPHX_TRY(mf_commit(c));
  c->mf_stress_ok = false;
  if (c->num == 0) return 0;
  const float4* dcorr =
      c->mlp.ready && c->mlp.drift ? c->mlp.hist[c->cur].p : nullptr;
  timer_begin(c, "mf_drift");
  SPHX_TRY(mf_drift_into(c, dcorr, c->mf_drift[0].p, c->mf_drift[1].p, c->mf_drift[2].p));
  timer_end(c);
  return 0;
}

int sphx_mf_alpha_advance(sphx_context* c) {
  SPHX_TRY(need_mf(c, "sphx_mf_alpha_advance"));
  SPHX_TRY(mf_commit(c));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  MFP m = make_mfp(c);
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  SPHX_TRY(mf_need_rec(c));
  timer_begin(c, "mf_alpha");
  auto* kern = m.nphase == 1 ? k_mf_alpha<1> : (m.nphase == 2 ? k_mf_alpha<2> : k_mf_alpha<3>);
  kern<<<blocks_for(k.n, c->tune[T_MF_ALPHA]), c->tune[T_MF_ALPHA], 0, c->stream>>>(
      k, m, L, c->key.p, c->cell_start.p, c->POS(), c->VEL(), c->PHYS(), c->mf_alpha[c->cur].p,
      c->mf_drift[0].p, c->mf_drift[1].p, c->mf_drift[2].p, c->mf_rec.p, c->mf_alpha_new.p,
      c->mf_stress.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  c->mf_stress_ok = true;
  return 0;
}

int sphx_mf_correction(sphx_context* c) {
  SPHX_TRY(need_mf(c, "sphx_mf_correction"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  MFP m = make_mfp(c);
  timer_begin(c, "mf_correction");
  k_mf_correction<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, m, c->POS(), c->key.p, c->mf_alpha[c->cur].p, c->mf_alpha_new.p, c->VEL());
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  c->mf_pending = true;
  return 0;
}

// Needs the advected velocity of sphx_iisph_predict_advection's first
// kernel; adds dt * a_mf to it.
int sphx_mf_force(sphx_context* c) {
  SPHX_TRY(need_mf(c, "sphx_mf_force"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  MFP m = make_mfp(c);
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  const float4* an = c->mf_pending ? c->mf_alpha_new.p : c->mf_alpha[c->cur].p;
  timer_begin(c, "mf_force");
  if (c->mf_stress_ok) {
    // the alpha pass summed the drift stress from these same inputs
    k_mf_force_add<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
        k, m, c->POS(), c->key.p, an, c->mf_stress.p, c->VELEVAL(), c->mf_amf.p);
  } else {
    SPHX_TRY(mf_need_rec(c));
    k_mf_force<<<blocks_for(k.n, c->tune[T_MF_FORCE]), c->tune[T_MF_FORCE], 0, c->stream>>>(
        k, m, L, c->key.p, c->cell_start.p, c->POS(), c->PHYS(), c->mf_alpha[c->cur].p, an,
        c->mf_drift[0].p, c->mf_drift[1].p, c->mf_drift[2].p, c->mf_rec.p, c->VELEVAL(),
        c->mf_amf.p);
  }
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

int sphx_grid_build(sphx_context* c);
int sphx_iisph_predict_aii(sphx_context* c);

}  // extern "C"

namespace sphx {

// sphx_mf_correction then sphx_mf_force: one pointwise pass when the alpha
// pass summed the drift stress (the composite steps), else the two calls.
int mf_correction_force(sphx_context* c) {
  SPHX_TRY(need_mf(c, "sphx_mf_correction"));
  if (!c->mf_stress_ok) {
    SPHX_TRY(sphx_mf_correction(c));
    return sphx_mf_force(c);
  }
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  MFP m = make_mfp(c);
  timer_begin(c, "mf_correction_force");
  k_mf_correct_force<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, m, c->nbr_valid ? c->nbr_cnt.p : nullptr, c->POS(), c->key.p, c->mf_alpha[c->cur].p,
      c->mf_alpha_new.p, c->VEL(),
      c->mf_stress.p, c->VELEVAL(), c->mf_amf.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  c->mf_pending = true;
  return 0;
}

}  // namespace sphx

extern "C" {
int sphx_iisph_pressure_solve(sphx_context* c, int* iters, double* err);
int sphx_iisph_integrate(sphx_context* c);

// One two-phase step: grid build, density with the advected velocity and
// d_ii (one pass), drift velocity, fraction advection (+ drift stress),
// correction with the drift stress and buoyancy added to the advected
// velocity (one pointwise pass), a_ii, pressure solve to tolerance,
// integration.  The same arithmetic as the calls density, drift, alpha,
// correction, predict_dii, force, ...: the fraction passes change neither
// x, v.xyz, m nor rho0, which are all d_ii and v_adv read.
int sphx_mf_step(sphx_context* c, int* iters, double* err) {
  SPHX_CHECK(c, "sphx_mf_step", "null context");
  SPHX_CHECK(c->nphase > 0, "sphx_mf_step", "multiphase not set up");
  SPHX_TRY(sphx_grid_build(c));
  SPHX_TRY(iisph_density_dii(c));
  SPHX_TRY(sphx_mf_drift_velocity(c));
  SPHX_TRY(sphx_mf_alpha_advance(c));
  SPHX_TRY(mf_correction_force(c));
  SPHX_TRY(sphx_iisph_predict_aii(c));
  SPHX_TRY(sphx_iisph_pressure_solve(c, iters, err));
  return sphx_iisph_integrate(c);
}

}  // extern "C"

// ------------------------------------------------------------ checkpoint --
namespace sphx {

// Multiphase state of a context in slot order (the order sphx_download
// uses): nphase, rho[3], nu[3], then fractions {alpha_0..2, rho_m} and g - a
// of the last step, one float4 each per particle.
int mf_state_get(sphx_context* c, std::vector<char>& out) {
  out.clear();
  if (!c || c->nphase == 0) return 0;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  const int n = c->num;
  std::vector<float4> a((size_t)n), f((size_t)n);
  if (n > 0) {
    const float4* src = c->mf_pending ? c->mf_alpha_new.p : c->mf_alpha[c->cur].p;
    SPHX_HIP(hipMemcpy(a.data(), src, n * sizeof(float4), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(f.data(), c->mf_force[c->cur].p, n * sizeof(float4), hipMemcpyDeviceToHost));
  }
  const size_t head = sizeof(int32_t) * 2 + sizeof(float) * 6;
  out.resize(head + 2 * (size_t)n * sizeof(float4));
  char* p = out.data();
  const int32_t hn[2] = {c->nphase, n};
  std::memcpy(p, hn, sizeof(hn));
  std::memcpy(p + 8, c->mf_rho, 12);
  std::memcpy(p + 20, c->mf_nu, 12);
  std::memcpy(p + head, a.data(), n * sizeof(float4));
  std::memcpy(p + head + n * sizeof(float4), f.data(), n * sizeof(float4));
  return 0;
}

// Restore what mf_state_get wrote, after the particles were uploaded (the
// particles' mixture viscosity came with them and is left as it is).
int mf_state_set(sphx_context* c, const char* p, size_t bytes) {
  SPHX_CHECK(c && p, "sphx_checkpoint_restore", "null multiphase section");
  const size_t head = sizeof(int32_t) * 2 + sizeof(float) * 6;
  SPHX_CHECK(bytes >= head, "sphx_checkpoint_restore", "truncated multiphase section");
  int32_t hn[2];
  std::memcpy(hn, p, sizeof(hn));
  const int np = hn[0], n = hn[1];
  SPHX_CHECK(np >= 1 && np <= 3 && n == c->num &&
                 bytes == head + 2 * (size_t)n * sizeof(float4),
             "sphx_checkpoint_restore", "multiphase section does not match the particles");
  float rho[3], nu[3];
  std::memcpy(rho, p + 8, 12);
  std::memcpy(nu, p + 20, 12);
  SPHX_TRY(sphx_mf_setup(c, np, rho, nu));
  if (n > 0) {
    SPHX_HIP(hipStreamSynchronize(c->stream));
    SPHX_HIP(hipMemcpy(c->mf_alpha[c->cur].p, p + head, n * sizeof(float4), hipMemcpyHostToDevice));
    SPHX_HIP(hipMemcpy(c->mf_force[c->cur].p, p + head + n * sizeof(float4), n * sizeof(float4),
                       hipMemcpyHostToDevice));
  }
  c->mf_pending = false;
  c->mf_stress_ok = false;
  c->mf_rec_ok = false;
  return 0;
}

}  // namespace sphx
