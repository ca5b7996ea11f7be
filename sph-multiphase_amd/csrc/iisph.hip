// IISPH: density, advection prediction (d_ii, a_ii, rho_adv), relaxed
// Jacobi pressure iterations with an on-device FP64 convergence test, and
// integration.
//
// Reference: fluid_system_kern.cu:635-1115 (kern, nablakern, ComputeDii,
// ComputeAii, Pressure_DP, Pressure_Iter, IntegrateIISPH,
// ComputeDensityIISPH) and fluid_system_host.cu:505-556 (PressureSolve,
// whose loop never terminates; here it stops at the density tolerance).
//
// Data flow of one iteration l (same arithmetic as Pressure_DP +
// Pressure_Iter, regrouped so no kernel reads a value another thread of the
// same kernel writes):
//   A  dijpj_i = -dt^2 sum_j s_j gradW_ij      with s_j = m_j p_j / rho_j^2
//      q_i     = d_ii p_i + dijpj_i            (= a_p,i dt^2)
//   B  S_i     = sum_j m_j (dijpj_i - q_j + dt^2 m_i p_i / rho_i^2 gradW_ij) . gradW_ij
//      p_i    <- (1-w) p_i + w/a_ii (rho0 - rho_adv - S_i)
//      residue = rho_adv + S_i + a_ii p_i(old) - rho0,   s_i <- m_i p_i(new)/rho_i^2
// B touches p only at its own index, so the update is in place and the
// reference's press_l1 -> press_l copy (host.cu:536) is not needed.
#include "mlp_mfma.hpp"

#include <hip/hip_cooperative_groups.h>

#include <cmath>
#include <type_traits>
#include <vector>

namespace sphx {

constexpr int NB_THREADS = 256;

struct IterCtl {
  int min_iters, max_iters;
  float tol;
};

__device__ __forceinline__ bool in_support(const KP& k, float dsq) {
  float lim = k.full_support ? 4.f * k.h2 : k.h2;
  return dsq < lim && dsq > 0.f;
}

// Per-pair terms of the density and d_ii passes, branch-free in the stored
// position units (see solve_gradf): x_u = x_i - x_j, dsq = |x_u|^2 (+inf for
// an unused slot).  The fused pass and the two separate ones call the same
// functions, so they produce the same bits.
struct DensDii {
  SolveGrad g;          // support 2h
  uint32_t lim_sup_m1;  // d_ii's support (SolverOptions::full_support): bits - 1
  float kw;             // kspline
  float lim2h;          // (2h)^2 / s^2
};
__device__ __forceinline__ DensDii dens_dii_consts(const KP& k) {
  DensDii c;
  c.g = solve_grad(k, true);
  c.lim_sup_m1 = k.full_support ? c.g.lim_m1 : __float_as_uint(k.h2 / (k.ss * k.ss)) - 1u;
  c.kw = k.kspline;
  c.lim2h = 4.f * k.h2 / (k.ss * k.ss);
  return c;
}
// sum + W(r) m_j of contributeIIDensity (kern.cu:1085-1100) for r < 2h
// (r = 0 included: a coincident neighbour adds W(0)).  The accumulation is
// an explicit FMA (as the d_ii sum below), so every kernel that calls these
// helpers rounds the same way whatever the compiler fuses around them.
__device__ __forceinline__ float pair_density(const DensDii& c, float dsq, float mj, float sum) {
#pragma clang fp contract(off)
  const float ir = __builtin_amdgcn_rsqf(dsq);
  const float q = dsq > 0.f ? dsq * ir * c.g.q_of_r : 0.f;
  const float t = 2.f - q;
  const float w = q > 1.f ? 0.25f * t * t * t * c.kw : (1.f - 1.5f * q * q + 0.75f * q * q * q) * c.kw;
  const bool in = dsq < c.lim2h;
  return __builtin_fmaf(in ? w : 0.f, in ? mj : 0.f, sum);
}
// ComputeDii's per-pair terms (kern.cu:673-798): grad W = x_u g for
// 0 < r < 2h; the artificial boundary viscosity of a boundary neighbour
// (:712-723, only for an approaching pair) is added to a; returns the
// factor f with d_ii -= x_u f (zero outside d_ii's support, or for a
// boundary neighbour without boundary pressure).
__device__ __forceinline__ float pair_dii(const KP& k, const DensDii& c, float dx, float dy,
                                          float dz, float dsq, bool bnd, float mj_fluid,
                                          float vb, float rest, const float4& vi,
                                          const float4* __restrict__ vel, uint32_t j, float& ax,
                                          float& ay, float& az) {
#pragma clang fp contract(off)
  const float g = solve_gradf(c.g, dsq);   // 0 outside 0 < r < 2h
  if (bnd && g != 0.f) {
    const float4 vj = vel[j];
    float pij = k.ss * (dx * (vi.x - vj.x) + dy * (vi.y - vj.y) + dz * (vi.z - vj.z));
    if (pij < 0.f) {
      pij = pij / (k.ss * k.ss * dsq + k.h2 * 0.01f);
      pij = pij * k.bvisc * k.h / 2.f / rest;
      pij = rest * vb * pij;
      const float gp = g * pij;
      ax = __builtin_fmaf(dx, gp, ax);
      ay = __builtin_fmaf(dy, gp, ay);
      az = __builtin_fmaf(dz, gp, az);
    }
  }
  const float mj = bnd ? (k.boundary_pressure ? rest * vb : 0.f) : mj_fluid;   // :763-764
  return (__float_as_uint(dsq) - 1u) < c.lim_sup_m1 ? g * mj : 0.f;
}

// ComputeDensityIISPH + contributeIIDensity (kern.cu:1067-1115): rho = sum m W
__global__ void k_ii_density(KP k, NbrList L, const uint32_t* __restrict__ key, const int* __restrict__ cs,
                             const float4* __restrict__ pos, float4* __restrict__ phys) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  const float4 qi = phys[i];
  const float rest = qi.w;
  // self term (the reference's scan visits i itself, W(0) = kspline)
  float sum = k.kspline * qi.x;
  const DensDii dc = dens_dii_consts(k);
  struct Rec { float4 p, q; };
  for_each_listed_m<NBR_BATCH_SMALL, false>(
      k, L, cs, gc, i, [&](uint32_t j) { return Rec{load_at(pos, j * 16u), load_at(phys, j * 16u)}; },
      [&](uint32_t, const Rec& o, bool valid) {
    const float4 pj = o.p, qj = o.q;
    const float mj = (type_of(pj) == SPHX_TYPE_BOUNDARY) ? rest * qj.y : qj.x;
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float dsq = valid ? pair_dsq(dx, dy, dz) : INFINITY;
    sum = pair_density(dc, dsq, mj, sum);
  });
  phys[i].y = sum;
}

// ComputeDii (kern.cu:673-798): v_adv = v + dt (a_bvisc + g);
// d_ii = -dt^2 / rho_i^2 sum_j m_j gradW_ij
__global__ void k_ii_dii(KP k, NbrList L, const uint32_t* __restrict__ key, const int* __restrict__ cs,
                         const float4* __restrict__ pos, const float4* __restrict__ vel,
                         const float4* __restrict__ phys, float4* __restrict__ veleval,
                         float4* __restrict__ dii) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  const float4 vi = vel[i];
  const float4 qi = phys[i];
  const float rest = qi.w;
  float ax = 0.f, ay = 0.f, az = 0.f;
  float dx_ = 0.f, dy_ = 0.f, dz_ = 0.f;
  const DensDii dc = dens_dii_consts(k);
  struct Rec { float4 p, q; };
  for_each_listed_m<NBR_BATCH, true>(
      k, L, cs, gc, i, [&](uint32_t j) { return Rec{load_at(pos, j * 16u), load_at(phys, j * 16u)}; },
      [&](uint32_t j, const Rec& o, bool valid) {
    const float4 pj = o.p, qj = o.q;
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float dsq = valid ? pair_dsq(dx, dy, dz) : INFINITY;
    const bool bnd = type_of(pj) == SPHX_TYPE_BOUNDARY;
    const float f = pair_dii(k, dc, dx, dy, dz, dsq, bnd, qj.x, qj.y, rest, vi, vel, j, ax, ay, az);
    dx_ = __builtin_fmaf(-dx, f, dx_);
    dy_ = __builtin_fmaf(-dy, f, dy_);
    dz_ = __builtin_fmaf(-dz, f, dz_);
  });
  ax += k.g.x; ay += k.g.y; az += k.g.z;
  veleval[i] = make_float4(vi.x + ax * k.dt, vi.y + ay * k.dt, vi.z + az * k.dt, 0.f);
  float rho = qi.y;
  float f = 1.f / rho / rho * k.dt * k.dt;
  dii[i] = make_float4(dx_ * f, dy_ * f, dz_ * f, qi.x);   // d_ii, m (the DP pass's record)
}

// k_ii_density and k_ii_dii in one neighbour loop, for steps that run them
// back to back (the fraction passes between them in the two-phase step
// leave x, v.xyz, m and rho0 alone).  Both read the same {x, phys} record;
// d_ii's sum needs no density, so rho_i only scales it at the end.  The
// density sum visits the entries in the same order as k_ii_density, so
// rho_i, v_adv and d_ii are bitwise those of the two passes.
__global__ void k_ii_density_dii(KP k, NbrList L, const uint32_t* __restrict__ key,
                                 const int* __restrict__ cs, const float4* __restrict__ pos,
                                 const float4* __restrict__ vel, float4* __restrict__ phys,
                                 float4* __restrict__ veleval, float4* __restrict__ dii) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  const float4 vi = vel[i];
  const float4 qi = phys[i];
  const float rest = qi.w;
  float sum = k.kspline * qi.x;
  float ax = 0.f, ay = 0.f, az = 0.f;
  float dx_ = 0.f, dy_ = 0.f, dz_ = 0.f;
  const DensDii dc = dens_dii_consts(k);
  struct Rec { float4 p, q; };
  for_each_listed_m<NBR_BATCH, true>(
      k, L, cs, gc, i, [&](uint32_t j) { return Rec{load_at(pos, j * 16u), load_at(phys, j * 16u)}; },
      [&](uint32_t j, const Rec& o, bool valid) {
    const float4 pj = o.p, qj = o.q;
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float dsq = valid ? pair_dsq(dx, dy, dz) : INFINITY;
    const bool bnd = type_of(pj) == SPHX_TYPE_BOUNDARY;
    sum = pair_density(dc, dsq, bnd ? rest * qj.y : qj.x, sum);
    const float f = pair_dii(k, dc, dx, dy, dz, dsq, bnd, qj.x, qj.y, rest, vi, vel, j, ax, ay, az);
    dx_ = __builtin_fmaf(-dx, f, dx_);
    dy_ = __builtin_fmaf(-dy, f, dy_);
    dz_ = __builtin_fmaf(-dz, f, dz_);
  });
  phys[i].y = sum;
  ax += k.g.x; ay += k.g.y; az += k.g.z;
  veleval[i] = make_float4(vi.x + ax * k.dt, vi.y + ay * k.dt, vi.z + az * k.dt, 0.f);
  float f = 1.f / sum / sum * k.dt * k.dt;
  dii[i] = make_float4(dx_ * f, dy_ * f, dz_ * f, qi.x);   // d_ii, m (the DP pass's record)
}

// ComputeAii (kern.cu:801-908): rho_adv, warm start p0 = w p(t-1), a_ii;
// also the position-only sums G_i, H_i of the pressure pass, and, with
// FEAT, the learned start's 24 features (the same map as k_mlp_features,
// from the same neighbour records).  With START the pass then evaluates
// the network on its own features: each wave runs its 64 particles' rows
// from LDS on MFMA and writes the learned start (as k_mlp_infer would
// from the stored rows, with the same bits).
struct FeatOut {
  MLPP mp;
  const float4* alpha;   // fractions before the step, or null (one phase)
  const float4* hist;    // per particle: last solve's pressure increment in w
  float* X;              // [n][24]
  float* pprev;          // p(t-1)
  int xstride, xphase;   // rows of X written: slots i with i % xstride == xphase
};
struct StartOut {
  const uint4* wmf;      // the weights in the MFMA operand layout
  int warm;              // write the learned start
  float4* dcorr;         // drift correction for the next step, or null
};
struct AiiStartLds {
  MlpMfmaShared S;
  float xs[NB_THREADS * MLP_NF];            // the workgroup's feature rows
  float ys[NB_THREADS / 64][64 * MLP_NO];   // network outputs per wave
};

template <bool FEAT, bool START>
__global__ __launch_bounds__(NB_THREADS) void k_ii_aii(KP k, NbrList L, const uint32_t* __restrict__ key, const int* __restrict__ cs,
                         const float4* __restrict__ pos, const float4* __restrict__ veleval,
                         const float4* __restrict__ phys, const float4* __restrict__ dii,
                         float* __restrict__ rho_adv, float* __restrict__ press_l,
                         float4* __restrict__ ps, float4* __restrict__ gh,
                         float4* __restrict__ cst, FeatOut fo, StartOut so) {
  static_assert(FEAT || !START, "the learned start needs the features");
  __shared__ typename std::conditional<START, AiiStartLds, int>::type lds;
  if constexpr (START) {
    load_mfma(lds.S, so.wmf);
    __syncthreads();
  }
  const int i = xcd_index();
  bool fluid = false;
  float4 qi = make_float4(0.f, 0.f, 0.f, 0.f);
  if (i < k.n) {
    const float4 pi = pos[i];
    const uint32_t gc = key[i];
    qi = phys[i];
    fluid = !(type_of(pi) == SPHX_TYPE_BOUNDARY || (int)gc == k.gridtotal);
    if (!fluid) {
      if (FEAT) {
        fo.pprev[i] = qi.z;
        if (i % fo.xstride == fo.xphase) {
          float* x = fo.X + (size_t)i * MLP_NF;
          for (int q = 0; q < MLP_NF; ++q) x[q] = 0.f;
        }
      }
      if (START && so.dcorr) so.dcorr[i] = make_float4(0.f, 0.f, 0.f, so.dcorr[i].w);
    } else {
      const float4 vi = veleval[i];
      const float rest = qi.w, mi = qi.x, rhoi = qi.y;
      const float4 di = dii[i];
      const float cji = mi / rhoi / rhoi * k.dt * k.dt;
      const float p_i = fmaxf(0.f, qi.z);
      float dchange = 0.f, aii = 0.f;
      float Gx = 0.f, Gy = 0.f, Gz = 0.f, H = 0.f;
      FeatAcc fa(p_i);
      struct Rec { float4 p, q, v; };
      for_each_listed_b<NBR_BATCH, false>(k, L, cs, gc, i, [&](int j) { return Rec{pos[j], phys[j], veleval[j]}; },
                        [&](int, const Rec& o) {
        const float4 pj = o.p, qj = o.q, vj = o.v;
        float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
        float dsq = dx * dx + dy * dy + dz * dz;
        float r, ir;
        dist_rinv(dsq, r, ir);
        if (FEAT) fa.add(k, dx, dy, dz, r, ir, pj, qj, p_i);
        if (!in_support(k, dsq)) return;
        float gf = cubic_gradf(r, ir, k.inv_h, k.kspline);
        float gx = dx * gf, gy = dy * gf, gz = dz * gf;
        float vdg = (vi.x - vj.x) * gx + (vi.y - vj.y) * gy + (vi.z - vj.z) * gz;
        // fluid j: weight m_j and d_ji = gradW m_i / rho_i^2 dt^2, so
        // a_ii += (d_ii - d_ji) . gradW m_j; boundary j (when boundary
        // pressure is on, :833-834): weight rho0_i V_j and no d_ji term;
        // else nothing.  Written with selects, one expression for both kinds.
        const bool fl = type_of(pj) != SPHX_TYPE_BOUNDARY;
        const float w = fl ? qj.x : (k.boundary_pressure ? rest * qj.y : 0.f);
        const float wh = fl ? qj.x : 0.f;
        const float cj = fl ? cji : 0.f;
        dchange += vdg * w * k.dt;
        float tx = di.x - gx * cj, ty = di.y - gy * cj, tz = di.z - gz * cj;
        aii += (tx * gx + ty * gy + tz * gz) * w;
        Gx += gx * w; Gy += gy * w; Gz += gz * w;
        H += (gx * gx + gy * gy + gz * gz) * wh;
      });
      gh[i] = make_float4(Gx, Gy, Gz, H);
      const float radv = rhoi + dchange;
      rho_adv[i] = radv;
      if (!(START && so.warm)) {   // else the learned start below
        float p0 = k.warm * qi.z;
        press_l[i] = p0;
        ps[i].w = mi * p0 / rhoi / rhoi;
      }
      cst[i] = make_float4(aii, rest - radv, mi / rhoi / rhoi, rest);
      if (FEAT) {
        fo.pprev[i] = qi.z;
        float x[MLP_NF];
        feat_write(k, fo.mp, fa, qi, radv, di, aii, vi, fo.alpha, fo.hist, i, x);
        if (i % fo.xstride == fo.xphase) {
          float4* xg = reinterpret_cast<float4*>(fo.X + (size_t)i * MLP_NF);
#pragma unroll
          for (int q = 0; q < MLP_NF / 4; ++q)
            xg[q] = make_float4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
        }
        if constexpr (START) {
          float4* xl = reinterpret_cast<float4*>(lds.xs + threadIdx.x * MLP_NF);
#pragma unroll
          for (int q = 0; q < MLP_NF / 4; ++q)
            xl[q] = make_float4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
        }
      }
    }
  }
  if constexpr (START) {
    // Each wave evaluates its own 64 rows as soon as it has written them
    // (one wave's LDS accesses complete in order; the weights were staged
    // before the barrier above).
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    asm volatile("" ::: "memory");
    if (i - lane < k.n) {   // the wave holds particles (the same for all lanes)
      // rows of lanes past n or off the fluid are not written; each
      // particle's outputs depend on its own row only
      mlp_forward_mfma(lds.S, lds.xs + wave * 64 * MLP_NF, 64, lds.ys[wave]);
      asm volatile("" ::: "memory");
      if (fluid)
        mlp_start_out(k, fo.mp, so.warm, i, lds.ys[wave] + lane * MLP_NO, qi, press_l, ps,
                      so.dcorr);
    }
  }
}

// The position-only sums of the a_ii pass, G_i = sum_j m_j gradW_ij (plus
// rho0_i V_j gradW_ij over boundary j when boundary pressure is on) and
// H_i = sum_fluid m_j |gradW_ij|^2, for iterations that start from uploaded
// solver state instead of a fresh a_ii pass.
__global__ void k_ii_gh(KP k, NbrList L, const uint32_t* __restrict__ key, const int* __restrict__ cs,
                        const float4* __restrict__ pos, const float4* __restrict__ phys,
                        float4* __restrict__ gh) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  const float rest = phys[i].w;
  float Gx = 0.f, Gy = 0.f, Gz = 0.f, H = 0.f;
  struct Rec { float4 p, q; };
  for_each_listed_b<NBR_BATCH, false>(k, L, cs, gc, i, [&](int j) { return Rec{pos[j], phys[j]}; },
                                      [&](int, const Rec& o) {
    const float4 pj = o.p, qj = o.q;
    float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
    float dsq = dx * dx + dy * dy + dz * dz;
    if (!in_support(k, dsq)) return;
    float r, ir;
    dist_rinv(dsq, r, ir);
    float gf = cubic_gradf(r, ir, k.inv_h, k.kspline);
    float gx = dx * gf, gy = dy * gf, gz = dz * gf;
    const bool fl = type_of(pj) != SPHX_TYPE_BOUNDARY;
    const float w = fl ? qj.x : (k.boundary_pressure ? rest * qj.y : 0.f);
    const float wh = fl ? qj.x : 0.f;
    Gx += gx * w; Gy += gy * w; Gz += gz * w;
    H += (gx * gx + gy * gy + gz * gz) * wh;
  });
  gh[i] = make_float4(Gx, Gy, Gz, H);
}

// Pressure_DP (kern.cu:910-958) plus q_i = d_ii p_i + dijpj_i, stored as
// the record the Jacobi pass gathers ({x, m q.x}, {m q.y, m q.z}).  The
// pressure acceleration q_i / dt^2 (kern.cu:1044-1047) of the solve's last
// DP pass is written once, when the solve ends (k_ii_publish), from that
// record.  Per particle the pass reads its list, key, own {x, s} record,
// {d_ii, m} and p; boundary and out-of-grid slots are marked in the list.
struct DpArgs {
  const uint32_t* __restrict__ key;
  const int* __restrict__ cs;
  const float4* __restrict__ pos;   // for the scene of an ensemble's particle
  const float4* __restrict__ ps;
  const float4* __restrict__ dii;
  const float* __restrict__ press_l;
  float4* __restrict__ dijpj;
  float4* __restrict__ mqa;
  float2* __restrict__ mqb;
};

__device__ __forceinline__ void dp_particle(const KP& k, const NbrList& L, const DpArgs& a,
                                            const SceneCtl& sc, int i) {
  if (i >= k.n) return;
  if (L.cnt[i] == NBR_INVALID) return;   // boundary, or outside the grid
  if (sc.n && sc.state[2 * scene_of(a.pos[i])]) return;   // converged scene: p is final
  const uint32_t gc = a.key[i];
  const float4* __restrict__ ps = a.ps;
  const float4 pi = ps[i];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  // one 16-byte gather per neighbour: {x_j, s_j}; boundary s_j = 0 adds
  // nothing, as the reference's skip of boundary j (:925-926)
  const SolveGrad sg = solve_grad(k);
  for_each_listed_m<NBR_BATCH, true>(
      k, L, a.cs, gc, i, [&](uint32_t j) { return load_at(ps, j * 16u); },
      [&](uint32_t, const float4& pj, bool valid) {
    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const float dsq = valid ? pair_dsq(dx, dy, dz) : INFINITY;
    const float gf = solve_gradf(sg, dsq) * pj.w;
    sx += dx * gf; sy += dy * gf; sz += dz * gf;
  });
  float f = -k.dt * k.dt;
  float4 d = make_float4(sx * f, sy * f, sz * f, 0.f);
  a.dijpj[i] = d;
  const float4 di = a.dii[i];
  const float p = a.press_l[i];
  const float qx = di.x * p + d.x, qy = di.y * p + d.y, qz = di.z * p + d.z;
  const float mi = di.w;
  a.mqa[i] = make_float4(pi.x, pi.y, pi.z, mi * qx);
  a.mqb[i] = make_float2(mi * qy, mi * qz);
}

__global__ void k_ii_dp(KP k, NbrList L, const int* __restrict__ done, DpArgs a, SceneCtl sc) {
  if (*done) return;
  dp_particle(k, L, a, sc, xcd_index());
}

// Block reduction of (signed residue, |residue|/rho0, fluid count) in FP64.
__device__ void block_reduce3(double a, double b, double c, double* out) {
  __shared__ double sh[3][NB_THREADS / 64];   // blockDim <= NB_THREADS
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
    c += __shfl_down(c, off, 64);
  }
  if (lane == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0, y = 0, z = 0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) { x += sh[0][q]; y += sh[1][q]; z += sh[2][q]; }
    out[3 * blockIdx.x + 0] = x;
    out[3 * blockIdx.x + 1] = y;
    out[3 * blockIdx.x + 2] = z;
  }
}

// Pressure_Iter (kern.cu:960-1050) for particle i: the relaxed Jacobi
// update and the residual; rs, ra, rc receive its (signed residue,
// error / rho0, 1) when it counts toward the solve's error, scene its scene.
// Per particle the pass reads its list, key, own record {x, m q.x} (for x),
// the constants {a_ii, rho0 - rho_adv, m / rho^2, rho0} of the a_ii pass,
// dijpj, {G, H} and p, and writes p, the residual and the DP pass's record
// {x, s}; the pressure field and press_l1 are written when the solve ends.
struct IterArgs {
  const uint32_t* __restrict__ key;
  const int* __restrict__ cs;
  const float4* __restrict__ pos;   // for the scene of an ensemble's particle
  const float4* __restrict__ cst;
  const float4* __restrict__ dijpj;
  const float4* __restrict__ mqa;
  const float2* __restrict__ mqb;
  const float4* __restrict__ gh;
  float* __restrict__ press_l;
  float4* __restrict__ ps;
  float* __restrict__ residue;
};

__device__ __forceinline__ void iter_particle(const KP& k, const NbrList& L, const IterArgs& a,
                                              const SceneCtl& sc, int i, double& rs, double& ra,
                                              double& rc, int& scene) {
  rs = ra = rc = 0.0;
  scene = 0;
  if (i >= k.n) return;
  if (L.cnt[i] == NBR_INVALID) {   // boundary, or outside the grid
    a.residue[i] = 0.f;
    return;
  }
  if (sc.n) {
    scene = scene_of(a.pos[i]);
    if (sc.state[2 * scene]) return;   // converged scene: pressures final
  }
  const uint32_t gc = a.key[i];
  const float4* __restrict__ mqa = a.mqa;
  const float2* __restrict__ mqb = a.mqb;
  const float4 pi = mqa[i];
  const float4 ci = a.cst[i];
  const float aii = ci.x, b = ci.y, mr2 = ci.z, rest = ci.w;
  const float4 dpi = a.dijpj[i];
  const float pl = a.press_l[i];
  const float cji = k.dt * k.dt * mr2 * pl;
  // S_i = sum_j m_j (dijpj_i - q_j + d_ji p_i) . gradW_ij over fluid j
  // (kern.cu:992-999), plus m_b dijpj_i . gradW_ij over boundary j when
  // boundary pressure is on (:1000-1002).  The position-only parts are
  // sums of the a_ii pass: S_i = dijpj_i . G_i + (d_ji p_i / gradW) H_i -
  // sum_j m_j q_j . gradW_ij, so the loop gathers one record pair per
  // neighbour, {x_j, m_j q_j.x} and {m_j q_j.y, m_j q_j.z} (zero q for
  // boundary j).
  float acc = 0.f;
  const SolveGrad sg = solve_grad(k);
  struct Rec { float4 a; float2 b; };
  for_each_listed_m<NBR_BATCH, true>(
      k, L, a.cs, gc, i,
      [&](uint32_t j) { return Rec{load_at(mqa, j * 16u), load_at(mqb, j * 8u)}; },
      [&](uint32_t, const Rec& o, bool valid) {
    const float dx = pi.x - o.a.x, dy = pi.y - o.a.y, dz = pi.z - o.a.z;
    const float dsq = valid ? pair_dsq(dx, dy, dz) : INFINITY;
    acc += solve_gradf(sg, dsq) * (dx * o.a.w + dy * o.b.x + dz * o.b.y);
  });
  const float4 G = a.gh[i];
  const float S = dpi.x * G.x + dpi.y * G.y + dpi.z * G.z + cji * G.w - acc;
  float pn;
  if (fabsf(aii) < 1e-6f) pn = pl;
  else pn = (1.f - k.omega) * pl + k.omega / aii * (b - S);
  if (k.clamp_p && pn < 0.f) pn = 0.f;   // kern.cu:1035-1036 (commented there)
  const float res = (S + aii * pl) - b;   // rho_adv + S + a_ii p - rho0
  a.residue[i] = res;
  a.press_l[i] = pn;
  a.ps[i] = make_float4(pi.x, pi.y, pi.z, mr2 * pn);
  rs = res;
  // With the clamp the solve is a complementarity problem (p > 0 and
  // rho = rho0, or p = 0 and rho <= rho0): the error is |res| where the
  // residual's pressure p_l is positive, compression only where it is
  // zero.  NaN must propagate into the sum (fmax would drop it and
  // report convergence), so the one-sided part is written as a branch.
  double rr = (double)res;
  ra = (k.clamp_p && !(pl > 0.f) ? (rr < 0.0 ? 0.0 : rr) : fabs(rr)) / (double)rest;
  rc = 1.0;
}

// Per-scene sums of a workgroup: LDS atomics, then one global atomic per
// touched entry (sacc holds 3 * sc.n doubles, zeroed by the caller).
__device__ __forceinline__ void scene_accumulate(const SceneCtl& sc, double* sacc, int scene,
                                                 double rs, double ra, double rc) {
  __syncthreads();
  if (rc != 0.0) {
    atomicAdd(&sacc[3 * scene], rs);
    atomicAdd(&sacc[3 * scene + 1], ra);
    atomicAdd(&sacc[3 * scene + 2], rc);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 3 * sc.n; t += blockDim.x)
    if (sacc[t] != 0.0) atomicAdd(&sc.acc[t], sacc[t]);
}

__global__ __launch_bounds__(NB_THREADS) void k_ii_iter(KP k, NbrList L,
                                                        const int* __restrict__ done, IterArgs a,
                                                        double* __restrict__ partial,
                                                        SceneCtl sc) {
  if (*done) return;
  extern __shared__ double sacc[];   // 3 * sc.n doubles, ensembles only
  if (sc.n)
    for (int t = threadIdx.x; t < 3 * sc.n; t += blockDim.x) sacc[t] = 0.0;
  double rs, ra, rc;
  int scene;
  iter_particle(k, L, a, sc, xcd_index(), rs, ra, rc, scene);
  if (sc.n) {
    scene_accumulate(sc, sacc, scene, rs, ra, rc);
    return;
  }
  block_reduce3(rs, ra, rc, partial);
}

// Sum the partials, count the iteration and decide convergence (one
// workgroup of up to 1024 threads).
__device__ void finalize_single(int* __restrict__ state, double* __restrict__ err,
                                const double* __restrict__ partial, int nblocks, int num,
                                IterCtl ctl) {
  __shared__ double sh[3][16];
  double a = 0, b = 0, c = 0;
  for (int q = threadIdx.x; q < nblocks; q += blockDim.x) {
    a += partial[3 * q];
    b += partial[3 * q + 1];
    c += partial[3 * q + 2];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
    c += __shfl_down(c, off, 64);
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0, y = 0, z = 0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) { x += sh[0][q]; y += sh[1][q]; z += sh[2][q]; }
    double e = z > 0 ? y / z : 0.0;
    err[0] = e;
    err[1] = num > 0 ? x / num : 0.0;
    int it = state[1] + 1;
    state[1] = it;
    if ((it >= ctl.min_iters && e <= (double)ctl.tol) || it >= ctl.max_iters) state[0] = 1;
  }
}

__global__ __launch_bounds__(1024) void k_ii_finalize(int* __restrict__ state,
                                                      double* __restrict__ err,
                                                      const double* __restrict__ partial,
                                                      int nblocks, int num, IterCtl ctl,
                                                      int respect_done) {
  if (respect_done && state[0]) return;
  finalize_single(state, err, partial, nblocks, num, ctl);
}

// Ensemble version: per-scene error, iteration count and convergence; the
// solve is done when every scene is.  err[0] = largest scene error,
// state[1] = most iterations of any scene.  One workgroup of 256 threads.
__device__ void finalize_scenes(int* __restrict__ state, double* __restrict__ err,
                                const SceneCtl& sc, IterCtl ctl) {
  __shared__ int s_open[256], s_it[256];
  __shared__ double s_err[256];
  int open = 0, maxit = 0;
  double maxerr = 0.0;
  for (int s = threadIdx.x; s < sc.n; s += blockDim.x) {
    int done = sc.state[2 * s];
    int it = sc.state[2 * s + 1];
    if (!done) {
      double cnt = sc.acc[3 * s + 2];
      double e = cnt > 0 ? sc.acc[3 * s + 1] / cnt : 0.0;
      sc.err[s] = e;
      ++it;
      sc.state[2 * s + 1] = it;
      // a NaN error stops the scene, which then reports as unconverged
      if ((it >= ctl.min_iters && e <= (double)ctl.tol) || it >= ctl.max_iters || !(e == e))
        done = 1;
      sc.state[2 * s] = done;
    }
    sc.acc[3 * s] = 0.0;
    sc.acc[3 * s + 1] = 0.0;
    sc.acc[3 * s + 2] = 0.0;
    open += !done;
    maxit = max(maxit, it);
    double e = sc.err[s];
    if (!(e == e) || e > maxerr) maxerr = e;   // NaN wins
  }
  s_open[threadIdx.x] = open;
  s_it[threadIdx.x] = maxit;
  s_err[threadIdx.x] = maxerr;
  __syncthreads();
  if (threadIdx.x == 0) {
    int o = 0, m = 0;
    double e = 0.0;
    for (int t = 0; t < (int)blockDim.x; ++t) {
      o += s_open[t];
      m = max(m, s_it[t]);
      if (!(s_err[t] == s_err[t]) || s_err[t] > e) e = s_err[t];
    }
    state[0] = o == 0;
    state[1] = m;
    err[0] = e;
    err[1] = 0.0;
  }
}

__global__ __launch_bounds__(256) void k_ii_finalize_scenes(int* __restrict__ state,
                                                            double* __restrict__ err,
                                                            SceneCtl sc, IterCtl ctl,
                                                            int respect_done) {
  if (respect_done && state[0]) return;
  finalize_scenes(state, err, sc, ctl);
}

// Start of a solve: clear the done flag, iteration count and (ensemble)
// every scene's state.  A kernel, not a memset, so it can sit in a graph.
__global__ void k_solve_reset(int* __restrict__ state, SceneCtl sc) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) { state[0] = 0; state[1] = 0; state[2] = 0; state[3] = 0; }
  if (t < sc.n) {
    sc.state[2 * t] = 0;
    sc.state[2 * t + 1] = 0;
    sc.acc[3 * t] = 0.0;
    sc.acc[3 * t + 1] = 0.0;
    sc.acc[3 * t + 2] = 0.0;
  }
}

// The whole pressure solve as one cooperative launch of G co-resident
// workgroups (SolverOptions-independent: the same passes and arithmetic as
// the DP / Jacobi / finalize launches).  Per iteration: the DP pass over
// every particle, a grid barrier, the Jacobi pass with per-workgroup sums,
// a grid barrier, the convergence test by workgroup 0, and a grid barrier
// after which every workgroup reads the same done flag; at most max_iters
// iterations.  Workgroups walk 256-particle tiles: XCD x (workgroup b goes
// to XCD b % 8) owns the contiguous tile range [x q + min(x, r), ...) of the
// cell-sorted particles (q, r = tiles / 8, tiles % 8), so neighbour gathers
// stay in its L2.
__device__ __forceinline__ void coop_tiles(int ntiles, int& first, int& count, int& lane0,
                                           int& step) {
  const int G = gridDim.x, b = blockIdx.x;
  if (G % 8 == 0) {
    const int x = b & 7, q = ntiles >> 3, r = ntiles & 7;
    first = x * q + min(x, r);
    count = q + (x < r ? 1 : 0);
    lane0 = b >> 3;
    step = G >> 3;
  } else {
    first = 0;
    count = ntiles;
    lane0 = b;
    step = G;
  }
}

__global__ __launch_bounds__(NB_THREADS) void k_ii_solve_coop(KP k, NbrList L, DpArgs da,
                                                              IterArgs ia,
                                                              int* __restrict__ state,
                                                              double* __restrict__ err,
                                                              double* __restrict__ partial,
                                                              SceneCtl sc, IterCtl ctl,
                                                              int ntiles) {
  // a solve already done (a fallback launch with nothing to redo): every
  // workgroup leaves before the first barrier
  if (state[0]) return;
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  extern __shared__ double sacc[];   // 3 * sc.n doubles, ensembles only
  int first, count, lane0, step;
  coop_tiles(ntiles, first, count, lane0, step);
  for (int it = 0; it < ctl.max_iters; ++it) {
    for (int t = lane0; t < count; t += step)
      dp_particle(k, L, da, sc, (first + t) * NB_THREADS + (int)threadIdx.x);
    grid.sync();
    if (sc.n) {
      for (int t = threadIdx.x; t < 3 * sc.n; t += blockDim.x) sacc[t] = 0.0;
      __syncthreads();
    }
    double rs = 0.0, ra = 0.0, rc = 0.0;
    for (int t = lane0; t < count; t += step) {
      double a, b, c;
      int scene;
      iter_particle(k, L, ia, sc, (first + t) * NB_THREADS + (int)threadIdx.x, a, b, c, scene);
      if (sc.n) {
        if (c != 0.0) {
          atomicAdd(&sacc[3 * scene], a);
          atomicAdd(&sacc[3 * scene + 1], b);
          atomicAdd(&sacc[3 * scene + 2], c);
        }
      } else {
        rs += a; ra += b; rc += c;
      }
    }
    if (sc.n) {
      __syncthreads();
      for (int t = threadIdx.x; t < 3 * sc.n; t += blockDim.x)
        if (sacc[t] != 0.0) atomicAdd(&sc.acc[t], sacc[t]);
    } else {
      block_reduce3(rs, ra, rc, partial);
    }
    grid.sync();
    if (blockIdx.x == 0) {
      if (sc.n) finalize_scenes(state, err, sc, ctl);
      else finalize_single(state, err, partial, gridDim.x, k.n, ctl);
    }
    grid.sync();
    if (state[0]) break;   // the same value in every workgroup after the barrier
  }
}

// Device-side fallback of a learned start, between two persistent solves:
// where the solve missed the tolerance (the whole scene, or each scene of an
// ensemble) the pressures restart from the plain start p0 = warm * p(t-1)
// (fluid_system_kern.cu:899) and the solve is reopened; the iterations run
// so far are kept in state[2] and the fallbacks counted in state[3].
__global__ void k_fallback_prep(KP k, int* __restrict__ state, const double* __restrict__ err,
                                SceneCtl sc, float tol, const float4* __restrict__ pos,
                                const uint32_t* __restrict__ key, float4* __restrict__ phys,
                                const float* __restrict__ pprev, float* __restrict__ press_l,
                                float4* __restrict__ ps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc.n) {
    if (blockIdx.x == 0) {
      __shared__ int any;
      if (threadIdx.x == 0) any = 0;
      __syncthreads();
      for (int t = threadIdx.x; t < sc.n; t += blockDim.x)
        if (!(sc.err[t] <= (double)tol)) {
          sc.state[2 * t] = 0;
          sc.state[2 * t + 1] = 0;
          atomicOr(&any, 1);
        }
      __syncthreads();
      if (threadIdx.x == 0 && any) {
        state[2] += state[1];
        state[3] += 1;
        state[1] = 0;
        state[0] = 0;
      }
    }
  } else if (i == 0 && !(err[0] <= (double)tol)) {
    state[2] += state[1];
    state[3] += 1;
    state[1] = 0;
    state[0] = 0;
  }
  if (i >= k.n) return;
  const float4 pp = pos[i];
  if (type_of(pp) == SPHX_TYPE_BOUNDARY || (int)key[i] == k.gridtotal) return;
  const double e = sc.n ? sc.err[scene_of(pp)] : err[0];
  if (e <= (double)tol) return;
  float4 q = phys[i];
  q.z = pprev[i];
  phys[i] = q;
  const float p0 = k.warm * q.z;
  press_l[i] = p0;
  ps[i].w = q.x * p0 / (q.y * q.y);
}

// Appends {iterations (with those a device fallback redid), error} of the
// solve just finished to a ring of `cap` entries.
constexpr int SOLVE_LOG = 4096;
__global__ void k_solve_log(const int* __restrict__ state, const double* __restrict__ err,
                            double* __restrict__ ring, int* __restrict__ count) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int slot = *count % SOLVE_LOG;
  ring[2 * slot] = (double)(state[1] + state[2]);
  ring[2 * slot + 1] = err[0];
  *count += 1;
}

// Guard of the learned warm start, enqueued after the solve's first
// iteration (which ran from the learned start).  When that iteration's
// error exceeds `limit` = factor * tolerance (or is NaN), the learned first iterate is
// kept aside and one probe iteration runs from the plain start p0 = warm *
// p(t-1) (fluid_system_kern.cu:899); the solve then continues from whichever
// first iterate has the smaller error, and the probe counts as an iteration.
// Every decision is made on the device, so a captured step can hold it.
// gs: [0] skip the probe, [1] restore the learned iterate, [2] probes run,
// [3] probes that kept the plain start.
__global__ void k_guard_prep(KP k, const double* __restrict__ err, double limit, int* __restrict__ gs,
                             const float4* __restrict__ pos, const uint32_t* __restrict__ key,
                             const float4* __restrict__ phys, const float* __restrict__ pprev,
                             float* __restrict__ press_l, float4* __restrict__ ps,
                             float* __restrict__ keep) {
  const double e = err[0];
  const bool probe = !(e <= limit);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) gs[0] = probe ? 0 : 1;
  if (!probe || i >= k.n) return;
  const float4 pp = pos[i];
  if (type_of(pp) == SPHX_TYPE_BOUNDARY || (int)key[i] == k.gridtotal) return;
  const float4 q = phys[i];
  keep[i] = press_l[i];
  const float p0 = k.warm * pprev[i];
  press_l[i] = p0;
  ps[i].w = q.x * p0 / (q.y * q.y);
}

__global__ __launch_bounds__(1024) void k_guard_pick(int* __restrict__ gs, int* __restrict__ state,
                                                     double* __restrict__ err,
                                                     const double* __restrict__ partial,
                                                     int nblocks, int num, IterCtl ctl) {
  if (gs[0]) return;
  __shared__ double sh[3][16];
  double a = 0, b = 0, c = 0;
  for (int q = threadIdx.x; q < nblocks; q += 1024) {
    a += partial[3 * q];
    b += partial[3 * q + 1];
    c += partial[3 * q + 2];
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
    c += __shfl_down(c, off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0, y = 0, z = 0;
    for (int q = 0; q < 16; ++q) { x += sh[0][q]; y += sh[1][q]; z += sh[2][q]; }
    const double ep = z > 0 ? y / z : 0.0;
    const double el = err[0];
    const int it = state[1] + 1;
    state[1] = it;
    gs[2] += 1;
    if (ep < el || !(el == el)) {   // the plain start's first iterate is better
      err[0] = ep;
      err[1] = num > 0 ? x / num : 0.0;
      gs[1] = 0;
      gs[3] += 1;
      if ((it >= ctl.min_iters && ep <= (double)ctl.tol) || it >= ctl.max_iters) state[0] = 1;
    } else {
      gs[1] = 1;
      if (it >= ctl.max_iters) state[0] = 1;
    }
  }
}

__global__ void k_guard_restore(KP k, const int* __restrict__ gs, const float4* __restrict__ pos,
                                const uint32_t* __restrict__ key, const float4* __restrict__ phys,
                                const float* __restrict__ keep, float* __restrict__ press_l,
                                float4* __restrict__ ps) {
  if (gs[0] || !gs[1]) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.n) return;
  const float4 pp = pos[i];
  if (type_of(pp) == SPHX_TYPE_BOUNDARY || (int)key[i] == k.gridtotal) return;
  const float p = keep[i];
  const float4 q = phys[i];
  press_l[i] = p;
  ps[i].w = q.x * p / (q.y * q.y);
}

// End of a solve: the pressure field and press_l1 take the final iterate
// (the reference's Pressure_Iter writes both every iteration, :1037-1040),
// and the pressure acceleration q / dt^2 of the last DP pass (:1044-1047)
// is recovered from the record it stored, {x, m q.x}, {m q.y, m q.z}.
__global__ void k_ii_publish(KP k, const int* __restrict__ cnt, const float* __restrict__ press_l,
                             const float4* __restrict__ dii, const float4* __restrict__ mqa,
                             const float2* __restrict__ mqb, float4* __restrict__ phys,
                             float* __restrict__ press_l1, float4* __restrict__ accel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k.n || cnt[i] == NBR_INVALID) return;
  const float p = press_l[i];
  phys[i].z = p;
  press_l1[i] = p;
  const float f = 1.f / (dii[i].w * k.dt * k.dt);   // 1 / (m dt^2)
  const float4 a = mqa[i];
  const float2 b = mqb[i];
  accel[i] = make_float4(a.w * f, b.x * f, b.y * f, 0.f);
}

// IntegrateIISPH (kern.cu:1052-1064); optional clamp into the valid cells.
// With `tot`, also the FP64 conservation totals of the new state over the
// in-grid fluid particles (sum m, m v, m |v|^2 / 2, m x), one row of
// SPHX_TOTALS_Q partials per workgroup, summed by k_totals_final.
__global__ __launch_bounds__(NB_THREADS) void k_ii_integrate(
    KP k, const uint32_t* __restrict__ key, float4* __restrict__ pos, float4* __restrict__ vel,
    const float4* __restrict__ veleval, const float4* __restrict__ accel,
    float4* __restrict__ mforce, const float4* __restrict__ phys, double* __restrict__ tot) {
  const int i = xcd_index();
  double t[SPHX_TOTALS_Q] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i < k.n) {
    float4 p = pos[i];
    if (type_of(p) != SPHX_TYPE_BOUNDARY && (int)key[i] != k.gridtotal) {
      float4 ve = veleval[i], a = accel[i], v = vel[i];
      float vx = ve.x + a.x * k.dt, vy = ve.y + a.y * k.dt, vz = ve.z + a.z * k.dt;
      p.x += vx * k.dt / k.ss;
      p.y += vy * k.dt / k.ss;
      p.z += vz * k.dt / k.ss;
      if (k.clamp) {
        // keep 1 <= cell < res-1 on each axis (reflect the normal velocity)
        float lo[3], hi[3];
        lo[0] = k.gridmin.x + 1.0001f / k.gridfac.x; hi[0] = k.gridmin.x + (k.gridres.x - 1.0001f) / k.gridfac.x;
        lo[1] = k.gridmin.y + 1.0001f / k.gridfac.y; hi[1] = k.gridmin.y + (k.gridres.y - 1.0001f) / k.gridfac.y;
        lo[2] = k.gridmin.z + 1.0001f / k.gridfac.z; hi[2] = k.gridmin.z + (k.gridres.z - 1.0001f) / k.gridfac.z;
        if (p.x < lo[0]) { p.x = lo[0]; vx = fmaxf(vx, 0.f); }
        if (p.x > hi[0]) { p.x = hi[0]; vx = fminf(vx, 0.f); }
        if (p.y < lo[1]) { p.y = lo[1]; vy = fmaxf(vy, 0.f); }
        if (p.y > hi[1]) { p.y = hi[1]; vy = fminf(vy, 0.f); }
        if (p.z < lo[2]) { p.z = lo[2]; vz = fmaxf(vz, 0.f); }
        if (p.z > hi[2]) { p.z = hi[2]; vz = fminf(vz, 0.f); }
      }
      pos[i] = p;
      vel[i] = make_float4(vx, vy, vz, v.w);
      if (mforce) {
        // g - a of this step, read by the next drift-velocity pass
        float inv = 1.f / k.dt;
        mforce[i] = make_float4(k.g.x - (vx - v.x) * inv, k.g.y - (vy - v.y) * inv,
                                k.g.z - (vz - v.z) * inv, 0.f);
      }
      if (tot && type_of(p) == SPHX_TYPE_FLUID) particle_totals(phys[i].x, vx, vy, vz, p.x, p.y, p.z, t);
    }
  }
  if (tot) block_totals(t, tot);   // the same for the whole grid
}

// Neighbour list of every fluid particle: all j != i with |x_ij| < 2h, in
// the reference's 27-cell visiting order.
// Also seeds ps[i] = {x, y, z, 0}: the pressure pass gathers position and
// s_j = m_j p_j / rho_j^2 as one record (s stays 0 for boundary particles,
// whose pressure is not an unknown).
// Candidate positions staged per workgroup, as 12-byte {x, y, z} (37.5 KB:
// four workgroups per CU).
constexpr int NB_STAGE = 3200;
struct StagedPos { float x, y, z; };
constexpr int NB_THREADS_BUILD = 256;

__device__ __forceinline__ int nb_wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int nb_wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Neighbour list of the step (every j != i with |x_ij| < 2h, in the
// reference's visiting order).  A workgroup's 256 sorted particles lie in
// one or two (y, z) cell rows; for each such row the candidates of all its
// particles are its 9 neighbour rows over the row's x-range plus one cell,
// each one contiguous sorted range.  Those ranges are staged in LDS with
// coalesced loads and every thread tests its own 9 sub-ranges there.  A
// workgroup over three or more rows, or with more candidates than the
// stage holds, tests straight from global memory.
__global__ __launch_bounds__(NB_THREADS_BUILD) void k_nbr_build(
    KP k, const uint32_t* __restrict__ key, const int* __restrict__ cs,
    const float4* __restrict__ pos, uint4* __restrict__ idx, int* __restrict__ cnt,
    float4* __restrict__ ps, float4* __restrict__ mqa, float2* __restrict__ mqb, int cap) {
  __shared__ StagedPos spos[NB_STAGE + 3];   // + the tests past a range's end
  __shared__ int sh[7];           // row r0, row r1, key min/max of r0, of r1, bad
  __shared__ int rs[19], ro[19];  // range (row, neighbour row) begin; LDS offsets
  const int t = threadIdx.x, lane = t & 63;
  const int i = xcd_index();
  bool valid = false;
  float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t gc = 0;
  int row = 0;
  if (i < k.n) {
    pi = pos[i];
    ps[i] = make_float4(pi.x, pi.y, pi.z, 0.f);
    // the Jacobi pass's record: q = 0 until the pressure pass sets a fluid
    // particle's, and for good for a boundary particle
    mqa[i] = make_float4(pi.x, pi.y, pi.z, 0.f);
    mqb[i] = make_float2(0.f, 0.f);
    gc = key[i];
    valid = type_of(pi) != SPHX_TYPE_BOUNDARY && (int)gc != k.gridtotal;
    row = (int)(gc / (uint32_t)k.gridres.x);
  }
  if (t < 7) sh[t] = (t == 0 || t == 2 || t == 4) ? INT_MAX : (t == 6 ? 0 : INT_MIN);
  __syncthreads();
  {
    const int mn = nb_wave_min(valid ? row : INT_MAX), mx = nb_wave_max(valid ? row : INT_MIN);
    if (lane == 0 && mn != INT_MAX) {
      atomicMin(&sh[0], mn);
      atomicMax(&sh[1], mx);
    }
  }
  __syncthreads();
  const int r0 = sh[0], r1 = sh[1];
  const int R = row == r0 ? 0 : 1;
  {
    const int g = (int)gc;
    const int mn0 = nb_wave_min(valid && R == 0 ? g : INT_MAX);
    const int mx0 = nb_wave_max(valid && R == 0 ? g : INT_MIN);
    const int mn1 = nb_wave_min(valid && R == 1 ? g : INT_MAX);
    const int mx1 = nb_wave_max(valid && R == 1 ? g : INT_MIN);
    const bool bad = __ballot(valid && row != r0 && row != r1) != 0ull;
    if (lane == 0) {
      if (mn0 != INT_MAX) { atomicMin(&sh[2], mn0); atomicMax(&sh[3], mx0); }
      if (mn1 != INT_MAX) { atomicMin(&sh[4], mn1); atomicMax(&sh[5], mx1); }
      if (bad) sh[6] = 1;
    }
  }
  __syncthreads();
  if (t < 18) {
    const int g = t / 9, r = t % 9;
    const int kmin = sh[2 + 2 * g], kmax = sh[3 + 2 * g];
    int b = 0, e = 0;
    if (kmin != INT_MAX) {
      b = cs[kmin + k.rowoff[r]];
      e = cs[kmax + k.rowoff[r] + 3];
    }
    rs[t] = b;
    ro[t] = e - b;   // count for now
  }
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int q = 0; q < 18; ++q) {
      const int c = ro[q];
      ro[q] = run;
      run += c;
    }
    ro[18] = run;
  }
  __syncthreads();
  const int total = ro[18];
  const bool staged = sh[6] == 0 && total <= NB_STAGE;
  // candidates are tested on unscaled differences: |x_u|^2 < 4 h^2 / s^2
  const float rc2u = 4.f * k.h2 / (k.ss * k.ss);
  uint4* __restrict__ out = idx + nbr_gslot(i, 0, cap / NBR_GROUP);
  int c = 0;
  bool ovf = false;
  NbrPacker pk;
  // entry of candidate j in neighbour row r; sd = start of row 3 (r / 3)
  auto add = [&](int j, uint32_t dtag, int sd) {
    const uint32_t t = (uint32_t)(j - sd);
    ovf |= t > NBR_TMAX;
    if (c < cap) {
      pk.push(dtag | (t & NBR_TMAX));
      if (((c + 1) & (NBR_GROUP - 1)) == 0) out[(size_t)(c / NBR_GROUP) * NBR_SLICE] = pk.get();
    }
    ++c;
  };
  // Candidates of one row are tested 32 at a time into a hit mask, whose
  // set bits are then appended in order: the append runs once per hit of
  // the lane with the most hits, not once per candidate.
  auto scan_rows = [&](auto&& row_base, auto&& fetch) {
    int sd = 0, s0 = 0;
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
      const int c0 = (int)gc + k.rowoff[r];
      const int jb = cs[c0], je = cs[c0 + 3];
      if (r % 3 == 0) {
        // the decoder adds S_d - S_0 with a 24-bit multiply
        if (r == 0) s0 = jb;
        else ovf |= (uint32_t)(jb - s0) >= (1u << 24);
        sd = jb;
      }
      const uint32_t dtag = (uint32_t)(r / 3) << NBR_TBITS;
      const int base = row_base(r);
#pragma unroll 1
      for (int j0 = jb; j0 < je; j0 += 32) {
        const int m = min(32, je - j0);
        const int m4 = (m + 3) & ~3;
        // Candidates are tested four at a time, up to 3 past the range.
        // A test enters the mask at bit 0 and moves up, so candidate u
        // ends at bit m4 - 1 - u; its bit is the sign of d2 - rc2u, which
        // is d2 < rc2u exactly (the difference of two finite floats is
        // negative iff the first is smaller), shifted in by v_alignbit.
        uint32_t hit = 0;
        for (int u0 = 0; u0 < m4; u0 += 4) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float4 pj = fetch(base + j0 + u0 + v);
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
            hit = __builtin_amdgcn_alignbit(hit, __float_as_uint(d2 - rc2u), 31);
          }
        }
        hit >>= (uint32_t)(m4 - m);   // candidate u now at bit m - 1 - u
        // the particle itself is not its own neighbour
        const uint32_t self = (uint32_t)(i - j0);
        if (self < (uint32_t)m) hit &= ~(1u << (m - 1 - (int)self));
        if (hit == 0u) continue;
        // hits are appended in candidate order, so the last has the largest
        // offset from the layer start: one overflow test per chunk
        const int tlast = j0 + (m - 1 - __builtin_ctz(hit)) - sd;
        ovf |= (uint32_t)tlast > NBR_TMAX;
        if (c + __popc(hit) <= cap) {
          // the chunk fits: no capacity test per hit.  Entry of the hit at
          // bit b: dtag + (its offset), the offset below 2^14 unless the
          // list overflowed (then it is not used)
          const uint32_t top = dtag + (uint32_t)(j0 + m - 1 - sd);
          while (hit) {
            const int b = 31 - __builtin_clz(hit);
            hit ^= 1u << b;
            pk.push(top - (uint32_t)b);
            if (((c + 1) & (NBR_GROUP - 1)) == 0) out[(size_t)(c / NBR_GROUP) * NBR_SLICE] = pk.get();
            ++c;
          }
        } else {
          while (hit) {
            const int b = 31 - __builtin_clz(hit);
            hit ^= 1u << b;
            add(j0 + (m - 1 - b), dtag, sd);
          }
        }
      }
    }
  };
  if (staged) {
    for (int u = t; u < total; u += blockDim.x) {
      int q = 0;
      while (u >= ro[q + 1]) ++q;
      const float4 pj = pos[rs[q] + (u - ro[q])];
      spos[u] = StagedPos{pj.x, pj.y, pj.z};
    }
    __syncthreads();
    if (valid) {
      const int* rsR = rs + 9 * R;
      const int* roR = ro + 9 * R;
      // row r's candidates sit at LDS offset roR[r] - rsR[r] + j
      scan_rows([&](int r) { return roR[r] - rsR[r]; }, [&](int u) {
        const StagedPos q = spos[u];
        return make_float4(q.x, q.y, q.z, 0.f);
      });
    }
  } else if (valid) {
    scan_rows([&](int) { return 0; }, [&](int j) { return pos[min(j, k.n - 1)]; });
  }
  if (c <= cap && (c & (NBR_GROUP - 1))) {
    // zero entries pad the last group; they decode to S_0 and are never used
    for (int u = c & (NBR_GROUP - 1); u < NBR_GROUP; ++u) pk.push(0u);
    out[(size_t)(c / NBR_GROUP) * NBR_SLICE] = pk.get();
  }
  if (i < k.n) cnt[i] = !valid ? NBR_INVALID : ((ovf || c > cap) ? cap + 1 : c);
}

int grid_hash(sphx_context* c);
int mlp_prepare_start(sphx_context* c);
int grid_sort(sphx_context* c);
int grid_reorder(sphx_context* c);
int mlp_solve(sphx_context* c, int* iters, double* err);
int totals_enqueue(sphx_context* c, bool ring);
int totals_enqueue_final(sphx_context* c, int nblocks);

static int need_sorted(sphx_context* c, const char* who) {
  SPHX_CHECK(c, who, "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(c->sorted_valid, who, "grid not built (call sphx_grid_build first)");
  return 0;
}

// Capacity from the particle spacing s = (m/rho0)^(1/3): 1.6x the
// neighbours of a lattice within 2h, plus slack, rounded to 8, <= 256.
static int nbr_capacity(const sphx_ParamCarrier& P) {
  double s = std::cbrt((double)P.mass / (double)P.restdensity) / (double)P.simscale;
  double ratio = s > 0 ? 2.0 * P.smoothradius / P.simscale / s : 4.0;
  double nn = 4.18879 * ratio * ratio * ratio;
  int cap = (int)(1.6 * nn + 8.0);
  cap = (cap + 7) & ~7;
  return std::max(16, std::min(256, cap));
}

int nbr_list(sphx_context* c, NbrList* out) {
  int n = c->num;
  if (!c->nbr_valid) {
    int cap = nbr_capacity(c->P);
    size_t slices = (size_t)(n + NBR_SLICE - 1) / NBR_SLICE;
    SPHX_TRY(c->nbr_idx.alloc(std::max<size_t>(slices, 1) * (cap / NBR_GROUP) * NBR_SLICE));
    SPHX_TRY(c->nbr_cnt.alloc(std::max(n, 1)));
    SPHX_TRY(c->ps.alloc(c->cap));
    c->nbr_cap = cap;
    if (n > 0) {
      KP k = make_kp(c);
      timer_begin(c, "nbr_build");
      k_nbr_build<<<blocks_for(n, NB_THREADS_BUILD), NB_THREADS_BUILD, 0, c->stream>>>(
          k, c->key.p, c->cell_start.p, c->POS(), c->nbr_idx.p, c->nbr_cnt.p, c->ps.p, c->mqa.p,
          c->mqb.p, cap);
      timer_end(c);
      SPHX_HIP(hipGetLastError());
    }
    c->nbr_valid = true;
  }
  out->idx = c->nbr_idx.p;
  out->cnt = c->nbr_cnt.p;
  out->cap = c->nbr_cap;
  return 0;
}

// Dynamic LDS of the Jacobi pass: the per-scene sums of an ensemble.
static size_t scene_lds(const SceneCtl& sc) { return (size_t)3 * sc.n * sizeof(double); }

static DpArgs dp_args(sphx_context* c) {
  return DpArgs{c->key.p, c->cell_start.p, c->POS(), c->ps.p, c->dii.p, c->press_l.p,
                c->dijpj.p, c->mqa.p, c->mqb.p};
}

static IterArgs iter_args(sphx_context* c) {
  return IterArgs{c->key.p, c->cell_start.p, c->POS(), c->cst.p, c->dijpj.p, c->mqa.p,
                  c->mqb.p, c->gh.p, c->press_l.p, c->ps.p, c->residue.p};
}

static int ensure_red(sphx_context* c) {
  size_t nb = (size_t)blocks_for(c->cap, 64);   // smallest tuned workgroup
  return c->red.alloc(3 * nb + 16);
}

int solve_reset(sphx_context* c) {
  SceneCtl sc = c->scenes();
  k_solve_reset<<<blocks_for(std::max(sc.n, 1), 256), 256, 0, c->stream>>>(c->solve_state.p, sc);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// Publish a solve's outcome (pressure field, press_l1, pressure
// acceleration); enqueued after every solve and single iteration.
static int publish_solve(sphx_context* c, const KP& k) {
  if (k.n == 0) return 0;
  timer_begin(c, "iisph_publish");
  k_ii_publish<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, c->nbr_cnt.p, c->press_l.p, c->dii.p, c->mqa.p, c->mqb.p, c->PHYS(), c->press_l1.p,
      c->ACCEL());
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// one DP + Iter + finalize; respect_done=1 inside the pipelined solve
static int enqueue_iteration(sphx_context* c, const KP& k, int respect_done) {
  const int tdp = c->tune[T_DP], tit = c->tune[T_ITER];
  int nb = blocks_for(k.n, tit);
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  if (!c->gh_valid) {
    k_ii_gh<<<blocks_for(k.n, NB_THREADS), NB_THREADS, 0, c->stream>>>(
        k, L, c->key.p, c->cell_start.p, c->POS(), c->PHYS(), c->gh.p);
    SPHX_HIP(hipGetLastError());
    c->gh_valid = true;
  }
  IterCtl ctl{c->O.min_iters, c->O.max_iters, c->O.tol};
  SceneCtl sc = c->scenes();
  timer_begin(c, "iisph_dp");
  k_ii_dp<<<blocks_for(k.n, tdp), tdp, 0, c->stream>>>(k, L, c->solve_state.p, dp_args(c), sc);
  timer_end(c);
  timer_begin(c, "iisph_iter");
  k_ii_iter<<<nb, tit, scene_lds(sc), c->stream>>>(k, L, c->solve_state.p, iter_args(c),
                                                   c->red.p, sc);
  timer_end(c);
  if (sc.n)
    k_ii_finalize_scenes<<<1, 256, 0, c->stream>>>(c->solve_state.p, c->solve_err.p, sc, ctl,
                                                   respect_done);
  else
    k_ii_finalize<<<1, 1024, 0, c->stream>>>(c->solve_state.p, c->solve_err.p, c->red.p, nb,
                                             k.n, ctl, respect_done);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// The guard's launches (see k_guard_prep): prep, a probe DP + Iter gated on
// gs[0] instead of the done flag, pick, restore.
static int enqueue_guard(sphx_context* c, const KP& k) {
  auto& m = c->mlp;
  SPHX_TRY(m.pkeep.alloc((size_t)c->cap));
  const int tdp = c->tune[T_DP], tit = c->tune[T_ITER];
  const int nb = blocks_for(k.n, tit);
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  IterCtl ctl{c->O.min_iters, c->O.max_iters, c->O.tol};
  SceneCtl none{0, nullptr, nullptr, nullptr};
  int* gs = m.gstate.p;
  k_guard_prep<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, c->solve_err.p, (double)m.guard_factor * (double)c->O.tol, gs, c->POS(), c->key.p,
      c->PHYS(), m.pprev.p, c->press_l.p, c->ps.p, m.pkeep.p);
  timer_begin(c, "iisph_dp");
  k_ii_dp<<<blocks_for(k.n, tdp), tdp, 0, c->stream>>>(k, L, gs, dp_args(c), none);
  timer_end(c);
  timer_begin(c, "iisph_iter");
  k_ii_iter<<<nb, tit, 0, c->stream>>>(k, L, gs, iter_args(c), c->red.p, none);
  timer_end(c);
  k_guard_pick<<<1, 1024, 0, c->stream>>>(gs, c->solve_state.p, c->solve_err.p, c->red.p, nb,
                                          k.n, ctl);
  k_guard_restore<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, gs, c->POS(), c->key.p, c->PHYS(), m.pkeep.p, c->press_l.p, c->ps.p);
  SPHX_HIP(hipGetLastError());
  return 0;
}

// One iteration of a solve; the solve's first one is followed by the guard
// when the learned start armed it.
static int enqueue_first_or_next(sphx_context* c, const KP& k, bool first) {
  SPHX_TRY(enqueue_iteration(c, k, 1));
  if (first && c->mlp.guard_armed) {
    c->mlp.guard_armed = false;
    if (c->nscenes == 0) SPHX_TRY(enqueue_guard(c, k));
  }
  return 0;
}

// The solve's errors and state written straight into pinned host memory:
// one small launch in place of two copies on the stream.
__global__ void k_state_out(const double* __restrict__ err, const int* __restrict__ state,
                            double* __restrict__ host) {
  if (threadIdx.x < 2) host[threadIdx.x] = err[threadIdx.x];
  else if (threadIdx.x < 6) reinterpret_cast<int*>(host + 4)[threadIdx.x - 2] = state[threadIdx.x - 2];
}

static int read_state(sphx_context* c) {
  k_state_out<<<1, 64, 0, c->stream>>>(c->solve_err.p, c->solve_state.p, c->host_pinned);
  SPHX_HIP(hipGetLastError());
  int st[4];
  SPHX_HIP(hipStreamSynchronize(c->stream));
  std::memcpy(st, c->host_pinned + 4, 4 * sizeof(int));
  c->host_pinned[2] = st[0];
  // iterations of this solve plus those of a solve a device fallback redid
  c->host_pinned[3] = st[1] + st[2];
  c->host_pinned[6] = st[3];
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

int sphx_iisph_density(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_iisph_density"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  c->mf_rec_ok = false;   // densities change
  c->mf_stress_ok = false;
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  timer_begin(c, "iisph_density");
  k_ii_density<<<blocks_for(k.n, c->tune[T_DENSITY]), c->tune[T_DENSITY], 0, c->stream>>>(
      k, L, c->key.p, c->cell_start.p, c->POS(), c->PHYS());
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

int sphx_iisph_predict_dii(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_iisph_predict_dii"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  timer_begin(c, "iisph_dii");
  k_ii_dii<<<blocks_for(k.n, c->tune[T_DII]), c->tune[T_DII], 0, c->stream>>>(
      k, L, c->key.p, c->cell_start.p, c->POS(), c->VEL(), c->PHYS(), c->VELEVAL(), c->dii.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"

namespace sphx {

// sphx_iisph_density followed by sphx_iisph_predict_dii, as one pass; for
// the composite steps, whose calls between the two touch neither x, v.xyz,
// m nor rho0.
int iisph_density_dii(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_iisph_density"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  c->mf_rec_ok = false;   // densities change
  c->mf_stress_ok = false;
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  timer_begin(c, "iisph_density_dii");
  k_ii_density_dii<<<blocks_for(k.n, c->tune[T_DENS_DII]), c->tune[T_DENS_DII], 0, c->stream>>>(
      k, L, c->key.p, c->cell_start.p, c->POS(), c->VEL(), c->PHYS(), c->VELEVAL(), c->dii.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

}  // namespace sphx

extern "C" {

int sphx_iisph_density_dii(sphx_context* c) { return iisph_density_dii(c); }

// Per-particle neighbour counts of the step's list (builds it if needed):
// counts[i] = number of j != i within 2h, or -1 where the list overflowed
// and the loops scan the 27 cells instead.  *cap = the list's capacity.
int sphx_nbr_counts(sphx_context* c, int* counts, int* cap) {
  SPHX_TRY(need_sorted(c, "sphx_nbr_counts"));
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  if (cap) *cap = L.cap;
  if (counts && c->num > 0) {
    SPHX_HIP(hipMemcpy(counts, L.cnt, c->num * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < c->num; ++i) {
      if (counts[i] > L.cap) counts[i] = -1;
      else if (counts[i] == NBR_INVALID) counts[i] = 0;
    }
  }
  return 0;
}

int sphx_iisph_predict_aii(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_iisph_predict_aii"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  // the learned start's features come from this pass when the network is
  // in use (the feature stage then finds them ready)
  auto& m = c->mlp;
  const bool feat = m.ready && (m.warm || m.record || m.online_every > 0);
  FeatOut fo{};
  if (feat) {
    SPHX_TRY(m.X.alloc((size_t)c->cap * MLP_NF));
    SPHX_TRY(m.pprev.alloc((size_t)c->cap));
    fo = FeatOut{MLPP{m.p_scale, m.u_scale, m.drift_blend},
                 c->nphase > 0 ? c->mf_alpha[c->cur].p : nullptr, m.hist[c->cur].p, m.X.p,
                 m.pprev.p, 1, 0};
  }
  // the learned start too, when the network is evaluated this step; the
  // feature rows in memory are then only the recorder's input (the slots
  // it keeps, with the phase it will use after this step's solve)
  float4* dcorr = (feat && m.drift && c->nphase > 0) ? m.hist[c->cur].p : nullptr;
  const bool start = feat && m.fuse_start && (m.warm || dcorr);
  StartOut so{};
  if (start) {
    SPHX_TRY(mlp_prepare_start(c));
    so = StartOut{m.wmf.p, m.warm, dcorr};
    const int stride = std::max(1, m.record_stride);
    fo.xstride = stride;
    fo.xphase = stride > 1 ? (int)(m.solves % stride) : 0;
  }
  timer_begin(c, "iisph_aii");
  auto* kern = start ? k_ii_aii<true, true> : (feat ? k_ii_aii<true, false> : k_ii_aii<false, false>);
  // with the network: four waves share each staged copy of the weights
  const int tb = start ? NB_THREADS : c->tune[T_AII];
  kern<<<blocks_for(k.n, tb), tb, 0, c->stream>>>(
      k, L, c->key.p, c->cell_start.p, c->POS(), c->VELEVAL(), c->PHYS(), c->dii.p,
      c->rho_adv.p, c->press_l.p, c->ps.p, c->gh.p, c->cst.p, fo, so);
  timer_end(c);
  c->gh_valid = true;
  m.feat_ok = feat;
  m.start_ok = start;
  SPHX_HIP(hipGetLastError());
  return 0;
}

int sphx_iisph_predict_advection(sphx_context* c) {
  SPHX_TRY(sphx_iisph_predict_dii(c));
  return sphx_iisph_predict_aii(c);
}

int sphx_iisph_pressure_iteration(sphx_context* c, double* mean_signed, double* mean_abs_rel) {
  SPHX_TRY(need_sorted(c, "sphx_iisph_pressure_iteration"));
  SPHX_TRY(ensure_red(c));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  SPHX_TRY(solve_reset(c));
  SPHX_TRY(enqueue_iteration(c, k, 0));
  SPHX_TRY(publish_solve(c, k));
  if (mean_signed || mean_abs_rel) {
    SPHX_TRY(read_state(c));
    if (mean_abs_rel) *mean_abs_rel = c->host_pinned[0];
    if (mean_signed) *mean_signed = c->host_pinned[1];
  }
  return 0;
}

int sphx_iisph_pressure_solve(sphx_context* c, int* iters, double* err) {
  SPHX_TRY(need_sorted(c, "sphx_iisph_pressure_solve"));
  return mlp_solve(c, iters, err);
}

}  // extern "C"

namespace sphx {

// Workgroups of the persistent solve: as many as are co-resident on the
// device (occupancy of k_ii_solve_coop times the CUs), at most one per
// tile, a multiple of 8 for the XCD-aware tile walk.
static int coop_grid(sphx_context* c, int ntiles, int* G) {
  if (c->coop_max == 0) {
    int per_cu = 0, cus = 0;
    SPHX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_ii_solve_coop), NB_THREADS,
        (size_t)3 * SPHX_MAX_SCENES * sizeof(double)));
    SPHX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    SPHX_CHECK(per_cu > 0 && cus > 0, "persistent solve", "no co-resident workgroup fits");
    c->coop_max = per_cu * cus;
  }
  int g = std::min(c->coop_max, ntiles);
  if (g >= 8) g -= g % 8;
  *G = std::max(g, 1);
  return 0;
}

// One cooperative launch of the persistent solve (continuing the solve
// state as it stands; a done solve returns at once).
static int launch_coop(sphx_context* c, const KP& k) {
  NbrList L;
  SPHX_TRY(nbr_list(c, &L));
  if (!c->gh_valid) {
    k_ii_gh<<<blocks_for(k.n, NB_THREADS), NB_THREADS, 0, c->stream>>>(
        k, L, c->key.p, c->cell_start.p, c->POS(), c->PHYS(), c->gh.p);
    SPHX_HIP(hipGetLastError());
    c->gh_valid = true;
  }
  IterCtl ctl{c->O.min_iters, c->O.max_iters, c->O.tol};
  SceneCtl sc = c->scenes();
  int ntiles = blocks_for(k.n, NB_THREADS);
  int G = 1;
  SPHX_TRY(coop_grid(c, ntiles, &G));
  KP kk = k;
  DpArgs da = dp_args(c);
  IterArgs ia = iter_args(c);
  int* st = c->solve_state.p;
  double* er = c->solve_err.p;
  double* part = c->red.p;
  void* args[] = {&kk, &L, &da, &ia, &st, &er, &part, &sc, &ctl, &ntiles};
  timer_begin(c, "iisph_solve");
  SPHX_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_ii_solve_coop), dim3(G),
                                      dim3(NB_THREADS), args, scene_lds(sc), c->stream));
  timer_end(c);
  return 0;
}

// The solve on the device, no host read: with `reset` the whole solve
// (reset, persistent solve), else the continuation of a solve whose first
// iterations were launched one by one (a done solve returns at once); then,
// when a learned start is in use, the device-side fallback (restart of what
// missed the tolerance from the plain start) and a second persistent
// launch, which returns at once when nothing was restarted.
int enqueue_solve_coop(sphx_context* c, bool reset) {
  SPHX_TRY(ensure_red(c));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  if (reset) {
    c->mlp.guard_armed = false;   // a whole persistent solve carries no guard
    SPHX_TRY(solve_reset(c));
  }
  SPHX_TRY(launch_coop(c, k));
  if (c->mlp.ready && c->mlp.warm) {
    k_fallback_prep<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
        k, c->solve_state.p, c->solve_err.p, c->scenes(), c->O.tol, c->POS(), c->key.p,
        c->PHYS(), c->mlp.pprev.p, c->press_l.p, c->ps.p);
    SPHX_HIP(hipGetLastError());
    SPHX_TRY(launch_coop(c, k));
  }
  if (c->solve_log.p) {
    k_solve_log<<<1, 64, 0, c->stream>>>(c->solve_state.p, c->solve_err.p, c->solve_log.p,
                                         c->solve_log_count.p);
    SPHX_HIP(hipGetLastError());
  }
  return publish_solve(c, k);
}

int solve_log_setup(sphx_context* c) {
  SPHX_TRY(c->solve_log.alloc(2 * SOLVE_LOG));
  SPHX_TRY(c->solve_log_count.alloc(1));
  SPHX_HIP(hipMemsetAsync(c->solve_log_count.p, 0, sizeof(int), c->stream));
  return 0;
}

int solve_log_read(sphx_context* c, int* iters, double* err, int max, int* count) {
  *count = 0;
  if (!c->solve_log.p) return 0;
  SPHX_HIP(hipStreamSynchronize(c->stream));
  int n = 0;
  SPHX_HIP(hipMemcpy(&n, c->solve_log_count.p, sizeof(int), hipMemcpyDeviceToHost));
  std::vector<double> r(2 * SOLVE_LOG);
  SPHX_HIP(hipMemcpy(r.data(), c->solve_log.p, r.size() * sizeof(double), hipMemcpyDeviceToHost));
  const int kept = std::min(n, SOLVE_LOG), take = std::min(kept, max);
  for (int s = 0; s < take; ++s) {
    const int slot = (n - take + s) % SOLVE_LOG;
    if (iters) iters[s] = (int)r[2 * slot];
    if (err) err[s] = r[2 * slot + 1];
  }
  *count = take;
  return 0;
}

// PressureSolve (fluid_system_host.cu:525-552) from the start the a_ii pass
// (or the learned start) left in press_l.
int iisph_solve_plain(sphx_context* c, int* iters, double* err) {
  SPHX_TRY(ensure_red(c));
  KP k = make_kp(c);
  if (iters) *iters = 0;
  if (err) *err = 0.0;
  if (k.n == 0) return 0;
  if (c->coop) {
    SPHX_TRY(enqueue_solve_coop(c, true));
    SPHX_TRY(read_state(c));
    int it = (int)c->host_pinned[3];
    c->last_iters = it;
    c->coop_fallbacks = (int)c->host_pinned[6];
    if (iters) *iters = it;
    if (err) *err = c->host_pinned[0];
    return timer_flush(c);
  }
  SPHX_TRY(solve_reset(c));
  int enq = 0;
  // Iterations queued before each host check.  A host round trip idles the
  // GPU for longer than an iteration queued past convergence costs (its
  // kernels exit on the device flag), so the first check comes after one
  // more iteration than the previous solve needed; after a long solve, at a
  // third of that (at least 16), so that a solve that needs far fewer does
  // not queue dozens of idle launches.  Later checks add a quarter of the
  // iterations so far (at least two) until two checks give the error's decay
  // rate, then the iterations that rate predicts (at most doubling).
  const int first = std::max(std::max(c->O.min_iters, 2),
                             std::min(c->last_iters + 1, std::max(16, (c->last_iters + 3) / 3)));
  int target = first, k_prev = 0;
  double e_prev = 0.0;
  for (;;) {
    target = std::min(target, c->O.max_iters);
    while (enq < target) SPHX_TRY(enqueue_first_or_next(c, k, enq++ == 0));
    SPHX_TRY(read_state(c));
    if (c->host_pinned[2] != 0.0 || enq >= c->O.max_iters) break;
    const double e = c->host_pinned[0], tol = (double)c->O.tol;
    int more = std::max(2, enq / 4);
    if (k_prev > 0 && e > tol && e < e_prev && e_prev > 0.0) {
      const double rate = std::pow(e / e_prev, 1.0 / (double)(enq - k_prev));
      const double need = std::log(tol / e) / std::log(rate);
      if (std::isfinite(need)) more = std::min(std::max(2, (int)std::ceil(need) + 1), std::max(2, enq));
    }
    k_prev = enq;
    e_prev = e;
    target = enq + more;
  }
  SPHX_TRY(publish_solve(c, k));
  int it = (int)c->host_pinned[3];
  c->last_iters = it;
  if (iters) *iters = it;
  if (err) *err = c->host_pinned[0];
  return timer_collect(c);
}

// The outcome of the solve on the stream: iterations (with those a device
// fallback redid) and the error; synchronises.
int read_solve(sphx_context* c, int* iters, double* err) {
  SPHX_TRY(read_state(c));
  const int it = (int)c->host_pinned[3];
  c->last_iters = it;
  c->coop_fallbacks = (int)c->host_pinned[6];
  if (iters) *iters = it;
  if (err) *err = c->host_pinned[0];
  return 0;
}

// Capturable part of a solve: reset and n iterations, no host reads
// (iterations past convergence exit on the device flag).
int enqueue_solve_fixed(sphx_context* c, int n) {
  SPHX_TRY(ensure_red(c));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  SPHX_TRY(solve_reset(c));
  for (int t = 0; t < n; ++t) SPHX_TRY(enqueue_first_or_next(c, k, t == 0));
  return 0;
}

// Finish a solve begun by enqueue_solve_fixed: check on the host and go on,
// two iterations at a time, until it converged or reached max_iters.
int continue_solve(sphx_context* c, int* iters, double* err) {
  KP k = make_kp(c);
  if (iters) *iters = 0;
  if (err) *err = 0.0;
  if (k.n == 0) return 0;
  SPHX_TRY(read_state(c));
  while (c->host_pinned[2] == 0.0 && (int)c->host_pinned[3] < c->O.max_iters) {
    for (int t = 0; t < 2; ++t) SPHX_TRY(enqueue_iteration(c, k, 1));
    SPHX_TRY(read_state(c));
  }
  SPHX_TRY(publish_solve(c, k));
  int it = (int)c->host_pinned[3];
  c->last_iters = it;
  if (iters) *iters = it;
  if (err) *err = c->host_pinned[0];
  return 0;
}

}  // namespace sphx

extern "C" {

int sphx_iisph_integrate(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_iisph_integrate"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  positions_changed(c);
  // FP64 mass / momentum / energy totals of the new state, kept per step:
  // a single scene's are summed by the integration pass itself
  const int nb = blocks_for(k.n, NB_THREADS);
  const bool fused = c->tot_track && c->nscenes <= 1 && c->tot_part.p && c->tot_scenes == 1 &&
                     (size_t)nb * SPHX_TOTALS_Q <= c->tot_part.n;
  timer_begin(c, "iisph_integrate");
  k_ii_integrate<<<nb, NB_THREADS, 0, c->stream>>>(
      k, c->key.p, c->POS(), c->VEL(), c->VELEVAL(), c->ACCEL(),
      c->nphase > 0 ? c->mf_force[c->cur].p : nullptr, c->PHYS(), fused ? c->tot_part.p : nullptr);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  if (fused) SPHX_TRY(totals_enqueue_final(c, nb));
  else if (c->tot_track) SPHX_TRY(totals_enqueue(c, true));
  return 0;
}

int sphx_iisph_step(sphx_context* c, int* iters, double* err) {
  SPHX_CHECK(c, "sphx_iisph_step", "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_TRY(grid_hash(c));
  SPHX_TRY(grid_sort(c));
  SPHX_TRY(grid_reorder(c));
  SPHX_TRY(iisph_density_dii(c));
  SPHX_TRY(sphx_iisph_predict_aii(c));
  SPHX_TRY(sphx_iisph_pressure_solve(c, iters, err));
  return sphx_iisph_integrate(c);
}

// Batched ensemble of nscenes scenes (scene index in bits 8-23 of the
// type word): each scene's pressure solve converges on its own.
int sphx_ensemble_setup(sphx_context* c, int nscenes) {
  SPHX_CHECK(c, "sphx_ensemble_setup", "null context");
  SPHX_CHECK(nscenes >= 0 && nscenes <= SPHX_MAX_SCENES, "sphx_ensemble_setup",
             "nscenes must be in [0, 256]");
  SPHX_TRY(ensure_device(c));
  if (nscenes > 0) {
    SPHX_TRY(c->sc_acc.alloc(3 * (size_t)nscenes));
    SPHX_TRY(c->sc_err.alloc(nscenes));
    SPHX_TRY(c->sc_state.alloc(2 * (size_t)nscenes));
    SPHX_HIP(hipMemset(c->sc_acc.p, 0, 3 * (size_t)nscenes * sizeof(double)));
    SPHX_HIP(hipMemset(c->sc_err.p, 0, (size_t)nscenes * sizeof(double)));
    SPHX_HIP(hipMemset(c->sc_state.p, 0, 2 * (size_t)nscenes * sizeof(int)));
  }
  c->nscenes = nscenes;
  return 0;
}

// Iterations and final error of each scene's last pressure solve.
int sphx_scene_stats(sphx_context* c, int* iters, double* err) {
  SPHX_CHECK(c, "sphx_scene_stats", "null context");
  SPHX_CHECK(c->nscenes > 0, "sphx_scene_stats", "not an ensemble");
  SPHX_TRY(ensure_device(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  std::vector<int> st(2 * (size_t)c->nscenes);
  SPHX_HIP(hipMemcpy(st.data(), c->sc_state.p, st.size() * sizeof(int), hipMemcpyDeviceToHost));
  if (iters)
    for (int s = 0; s < c->nscenes; ++s) iters[s] = st[2 * s + 1];
  if (err) SPHX_HIP(hipMemcpy(err, c->sc_err.p, c->nscenes * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int sphx_download_iisph(sphx_context* c, float* dii, float* dijpj, float* aii, float* rho_adv,
                        float* press_l, float* press_l1, float* residue) {
  SPHX_CHECK(c, "sphx_download_iisph", "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  int n = c->num;
  if (n == 0) return 0;
  std::vector<float4> t(n);
  if (dii) {
    SPHX_HIP(hipMemcpy(t.data(), c->dii.p, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) { dii[3 * i] = t[i].x; dii[3 * i + 1] = t[i].y; dii[3 * i + 2] = t[i].z; }
  }
  if (aii) {
    SPHX_HIP(hipMemcpy(t.data(), c->cst.p, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) aii[i] = t[i].x;
  }
  if (dijpj) {
    SPHX_HIP(hipMemcpy(t.data(), c->dijpj.p, n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) { dijpj[3 * i] = t[i].x; dijpj[3 * i + 1] = t[i].y; dijpj[3 * i + 2] = t[i].z; }
  }
  if (rho_adv) SPHX_HIP(hipMemcpy(rho_adv, c->rho_adv.p, n * sizeof(float), hipMemcpyDeviceToHost));
  if (press_l) SPHX_HIP(hipMemcpy(press_l, c->press_l.p, n * sizeof(float), hipMemcpyDeviceToHost));
  if (press_l1) SPHX_HIP(hipMemcpy(press_l1, c->press_l1.p, n * sizeof(float), hipMemcpyDeviceToHost));
  if (residue) SPHX_HIP(hipMemcpy(residue, c->residue.p, n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int sphx_upload_iisph(sphx_context* c, const float* dii, const float* dijpj, const float* aii,
                      const float* rho_adv, const float* press_l) {
  SPHX_CHECK(c, "sphx_upload_iisph", "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_HIP(hipStreamSynchronize(c->stream));
  int n = c->num;
  if (n == 0) return 0;
  std::vector<float4> t(n), ph(n);
  SPHX_HIP(hipMemcpy(ph.data(), c->PHYS(), n * sizeof(float4), hipMemcpyDeviceToHost));
  if (dii) {
    // the DP pass's record {d_ii, m}
    for (int i = 0; i < n; ++i) t[i] = make_float4(dii[3 * i], dii[3 * i + 1], dii[3 * i + 2], ph[i].x);
    SPHX_HIP(hipMemcpy(c->dii.p, t.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  }
  if (dijpj) {
    for (int i = 0; i < n; ++i) t[i] = make_float4(dijpj[3 * i], dijpj[3 * i + 1], dijpj[3 * i + 2], 0.f);
    SPHX_HIP(hipMemcpy(c->dijpj.p, t.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  }
  if (rho_adv) SPHX_HIP(hipMemcpy(c->rho_adv.p, rho_adv, n * sizeof(float), hipMemcpyHostToDevice));
  if (aii || rho_adv) {
    // the Jacobi pass's constants {a_ii, rho0 - rho_adv, m / rho^2, rho0}
    std::vector<float> ra(n);
    SPHX_HIP(hipMemcpy(t.data(), c->cst.p, n * sizeof(float4), hipMemcpyDeviceToHost));
    SPHX_HIP(hipMemcpy(ra.data(), c->rho_adv.p, n * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
      const float4 q = ph[i];
      t[i] = make_float4(aii ? aii[i] : t[i].x, q.w - ra[i], q.x / q.y / q.y, q.w);
    }
    SPHX_HIP(hipMemcpy(c->cst.p, t.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  }
  if (press_l) {
    SPHX_HIP(hipMemcpy(c->press_l.p, press_l, n * sizeof(float), hipMemcpyHostToDevice));
    // keep s_j = m_j p_j / rho_j^2 consistent with the uploaded pressures
    NbrList L;
    SPHX_TRY(nbr_list(c, &L));
    SPHX_HIP(hipStreamSynchronize(c->stream));
    std::vector<float4> s(n);
    SPHX_HIP(hipMemcpy(s.data(), c->ps.p, n * sizeof(float4), hipMemcpyDeviceToHost));
    std::vector<float4> pos(n);
    SPHX_HIP(hipMemcpy(pos.data(), c->POS(), n * sizeof(float4), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
      int t;
      std::memcpy(&t, &pos[i].w, 4);
      s[i].w = type_bits(t) == SPHX_TYPE_BOUNDARY ? 0.f : ph[i].x * press_l[i] / ph[i].y / ph[i].y;
    }
    SPHX_HIP(hipMemcpy(c->ps.p, s.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  }
  return 0;
}

}  // extern "C"
