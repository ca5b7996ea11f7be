// WCSPH neighbour loops: boundary volume, density + equation of state,
// pressure/viscosity acceleration, symplectic-Euler advance.
// Reference: fluid_system_kern.cu:183-534 (ComputeBoundaryVolume,
// ComputeDensityPressure, ComputeForce, AdvanceParticles) launched by
// fluid_system_host.cu:337-438,510-513.
//
// Kernel family 0 is the reference's poly6/spiky pair with support h
// (fluid_system.cpp:309-312); family 1 swaps in the cubic spline with
// support 2h (kern.cu:637-671) for the dam-break configuration.
#include "nbr.hpp"

namespace sphx {

__device__ __forceinline__ float poly6_c3(float r2, float dsq) {
  float c = r2 - dsq;
  return c * c * c;
}

// ComputeBoundaryVolume (kern.cu:249-277): V_b = 1 / (k * sum_b w)
__global__ void k_boundary_volume(KP k, const uint32_t* __restrict__ key,
                                  const int* __restrict__ cs, const float4* __restrict__ pos,
                                  float4* __restrict__ phys) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) != SPHX_TYPE_BOUNDARY) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  float sum = 0.f;
  const bool cubic = k.wcsph_kernel == 1;
  for_each_neighbor(k, cs, gc, [&](int j) {
    float4 pj = pos[j];
    if (type_of(pj) != SPHX_TYPE_BOUNDARY) return;
    float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
    float dsq = dx * dx + dy * dy + dz * dz;
    if (cubic) {
      sum += cubic_w(sqrtf(dsq), k.inv_h, k.kspline);
    } else if (dsq < k.h2) {
      sum += poly6_c3(k.h2, dsq);
    }
  });
  if (!cubic) sum *= k.kpoly6;
  if (sum == 0.f) sum = 1.f;
  phys[i].y = 1.f / sum;
}

// ComputeDensityPressure (kern.cu:214-318).  Stores 1/rho in dens.
__global__ void k_density_pressure(KP k, const uint32_t* __restrict__ key,
                                   const int* __restrict__ cs, const float4* __restrict__ pos,
                                   float4* __restrict__ phys) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  float4 qi = phys[i];
  const float rest = qi.w;
  const bool cubic = k.wcsph_kernel == 1;
  float sum = 0.f;
  for_each_neighbor(k, cs, gc, [&](int j) {
    float4 pj = pos[j];
    float4 qj = phys[j];
    int tj = type_of(pj);
    // boundary: rho0_i * V_j; any other type uses its own mass (A5)
    float mj = (tj == SPHX_TYPE_BOUNDARY) ? rest * qj.y : qj.x;
    float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
    float dsq = dx * dx + dy * dy + dz * dz;
    if (cubic) {
      sum += cubic_w(sqrtf(dsq), k.inv_h, k.kspline) * mj;
    } else if (dsq < k.h2) {
      sum += poly6_c3(k.h2, dsq) * mj;
    }
  });
  if (!cubic) sum *= k.kpoly6;
  if (sum == 0.f) sum = 1.f;
  float B = k.eos_B > 0.f ? k.eos_B : 2.5f * rest;
  float x = sum / rest;
  float xp;
  if (k.eos_gamma == 7.f) {
    float x2 = x * x, x4 = x2 * x2;
    xp = x4 * x2 * x;
  } else {
    xp = powf(x, k.eos_gamma);
  }
  float p = B * (xp - 1.f);
  if (p < 0.f) p = 0.f;
  qi.y = 1.f / sum;
  qi.z = p;
  phys[i] = qi;
}

// ComputeForce + contributeForce_new (kern.cu:325-424).
__global__ void k_force(KP k, const uint32_t* __restrict__ key, const int* __restrict__ cs,
                        const float4* __restrict__ pos, const float4* __restrict__ vel,
                        const float4* __restrict__ phys, float4* __restrict__ accel) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) == SPHX_TYPE_BOUNDARY) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  const float4 vi = vel[i];       // the reference reads vel as "iveleval" (:410)
  const float4 qi = phys[i];
  const float idens = qi.y, ipress = qi.z, irest = qi.w, ivisc = vi.w;
  const bool cubic = k.wcsph_kernel == 1;
  const float rsup2 = cubic ? 4.f * k.h2 : k.h2;
  float fx = 0.f, fy = 0.f, fz = 0.f;
  for_each_neighbor(k, cs, gc, [&](int j) {
    float4 pj = pos[j];
    float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
    float dsq = dx * dx + dy * dy + dz * dz;
    if (!(dsq < rsup2 && dsq > 0.f)) return;
    int tj = type_of(pj);
    if (tj != SPHX_TYPE_FLUID && tj != SPHX_TYPE_BOUNDARY) return;
    float r = sqrtf(dsq);
    float cm1;   // grad W = dist * cm1
    if (cubic) {
      cm1 = cubic_gradf(r, 1.f / r, k.inv_h, k.kspline);
    } else {
      float c = k.h - r;
      cm1 = k.kspikydiff * c * c / r;
    }
    float4 qj = phys[j];
    float4 vj = vel[j];
    float vx = vi.x - vj.x, vy = vi.y - vj.y, vz = vi.z - vj.z;
    if (tj == SPHX_TYPE_FLUID) {
      float cm = cm1 * qj.x * qj.y;
      float pm = -0.5f * cm * qj.z * idens;
      float vm = cm * (ivisc + vj.w) * idens;
      float s = pm;
      if (k.art_alpha > 0.f) {
        float vdotx = vx * dx + vy * dy + vz * dz;
        if (vdotx < 0.f) {
          float mu = k.h * vdotx / (dsq + 0.01f * k.h2);
          float rhobar = 0.5f * (1.f / idens + 1.f / qj.y);
          float Pi = -k.art_alpha * k.c0 * mu / rhobar;
          s += -qj.x * Pi * cm1;
        }
      }
      fx += dx * s + vx * vm;
      fy += dy * s + vy * vm;
      fz += dz * s + vz * vm;
    } else {
      float pm = -cm1 * irest * qj.y * ipress * idens * idens;
      fx += dx * pm; fy += dy * pm; fz += dz * pm;
      float pij = vx * dx + vy * dy + vz * dz;
      if (pij < 0.f) {
        pij = pij / (dsq + k.h2 * 0.01f);
        pij = pij * k.bvisc * k.h / 2.f / irest;
        pij = cm1 * irest * qj.y * pij;
        fx += dx * pij; fy += dy * pij; fz += dz * pij;
      }
    }
  });
  accel[i] = make_float4(fx, fy, fz, 0.f);
}

// AdvanceParticles (kern.cu:426-534)
__global__ void k_advance(KP k, const uint32_t* __restrict__ key, float4* __restrict__ pos,
                          float4* __restrict__ vel, float4* __restrict__ veleval,
                          const float4* __restrict__ accel) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 p = pos[i];
  if (type_of(p) == SPHX_TYPE_BOUNDARY) return;
  float4 v = vel[i];
  if ((int)key[i] == k.gridtotal) {
    p.x = p.y = p.z = -1000.f;
    pos[i] = p;
    vel[i] = make_float4(0.f, 0.f, 0.f, v.w);
    return;
  }
  float4 a = accel[i];
  float ax = a.x + k.g.x, ay = a.y + k.g.y, az = a.z + k.g.z;
  float nx = ax * k.dt + v.x, ny = ay * k.dt + v.y, nz = az * k.dt + v.z;
  p.x += nx * k.dt / k.ss;
  p.y += ny * k.dt / k.ss;
  p.z += nz * k.dt / k.ss;
  pos[i] = p;
  veleval[i] = make_float4((v.x + nx) * 0.5f, (v.y + ny) * 0.5f, (v.z + nz) * 0.5f, 0.f);
  vel[i] = make_float4(nx, ny, nz, v.w);
}

int grid_hash(sphx_context* c);
int grid_sort(sphx_context* c);
int grid_reorder(sphx_context* c);

static int need_sorted(sphx_context* c, const char* who) {
  SPHX_CHECK(c, who, "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(c->sorted_valid, who, "grid not built (call sphx_grid_build first)");
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

int sphx_boundary_volume(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_boundary_volume"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  timer_begin(c, "boundary_volume");
  k_boundary_volume<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(k, c->key.p, c->cell_start.p,
                                                                  c->POS(), c->PHYS());
  timer_end(c);
  c->gh_valid = false;   // boundary volumes enter the pressure pass's G_i
  SPHX_HIP(hipGetLastError());
  return 0;
}

int sphx_wcsph_density_pressure(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_wcsph_density_pressure"));
  c->mf_stress_ok = false;   // phys.y changes
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  timer_begin(c, "wcsph_density");
  k_density_pressure<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(k, c->key.p, c->cell_start.p,
                                                                   c->POS(), c->PHYS());
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

int sphx_wcsph_force(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_wcsph_force"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  timer_begin(c, "wcsph_force");
  k_force<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(k, c->key.p, c->cell_start.p, c->POS(),
                                                        c->VEL(), c->PHYS(), c->ACCEL());
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

int sphx_wcsph_advance(sphx_context* c) {
  SPHX_TRY(need_sorted(c, "sphx_wcsph_advance"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  positions_changed(c);
  timer_begin(c, "wcsph_advance");
  k_advance<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(k, c->key.p, c->POS(), c->VEL(),
                                                          c->VELEVAL(), c->ACCEL());
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

int sphx_wcsph_step(sphx_context* c) {
  SPHX_CHECK(c, "sphx_wcsph_step", "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_TRY(grid_hash(c));
  SPHX_TRY(grid_sort(c));
  SPHX_TRY(grid_reorder(c));
  SPHX_TRY(sphx_wcsph_density_pressure(c));
  SPHX_TRY(sphx_wcsph_force(c));
  return sphx_wcsph_advance(c);
}

}  // extern "C"
