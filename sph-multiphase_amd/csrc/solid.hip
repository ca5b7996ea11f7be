// SPH elastic solid and surface tension: the remaining neighbour loops of
// the reference's particle step that share the WCSPH layout.
//   ComputeSolidTensor + contributeVGrad      sph_solid.cu:480-555
//   ComputeSolidForce + ContributeSolidForce  sph_solid.cu:557-635
//   SurfaceTension + contributeSurfaceTensionYT fluid_system_kern.cu:569-625
// They visit the 27-cell neighbourhood over the dense cell table in the
// reference's order, with one accumulator across cells as the reference.
// Deformation gradients and stresses are cold per-particle matrices (upload
// order, reached through perm): deformGrad at mat[36 q], stress at
// mat[36 q + 9], row-major.
#include "nbr.hpp"

namespace sphx {

// v_j - v_i scaled and outer-multiplied with x_ij (contributeVGrad), then
// F += (grad v F) dt, strain F F^T - I, stress solidK * strain.
__global__ void k_solid_tensor(KP k, const uint32_t* __restrict__ key,
                               const int* __restrict__ cs, const float4* __restrict__ pos,
                               const float4* __restrict__ vel, const float4* __restrict__ phys,
                               const int* __restrict__ perm, float* __restrict__ mat) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) != SPHX_TYPE_ELASTIC) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  const float4 vi = vel[i];
  float g[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for_each_neighbor(k, cs, gc, [&](int j) {
    const float4 pj = pos[j];
    const float4 vj = vel[j];
    float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
    float dsq = dx * dx + dy * dy + dz * dz;
    if (dsq < k.h2 && dsq > 0.f) {
      float r = sqrtf(dsq);
      float c = k.h - r;
      const float4 qj = phys[j];
      float cm = k.kspikydiff * c * c / r * qj.x * qj.y;
      float jx = (vj.x - vi.x) * cm, jy = (vj.y - vi.y) * cm, jz = (vj.z - vi.z) * cm;
      g[0] += jx * dx; g[1] += jx * dy; g[2] += jx * dz;
      g[3] += jy * dx; g[4] += jy * dy; g[5] += jy * dz;
      g[6] += jz * dx; g[7] += jz * dy; g[8] += jz * dz;
    }
  });
  float* m = mat + (size_t)perm[i] * 36;
  float F[9], t[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) F[q] = m[q];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      t[a * 3 + b] = g[a * 3] * F[b] + g[a * 3 + 1] * F[3 + b] + g[a * 3 + 2] * F[6 + b];
#pragma unroll
  for (int q = 0; q < 9; ++q) F[q] += t[q] * k.dt;
  // strain F F^T - I (mat3transpose + mat3prod), stress solidK * strain
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      t[a * 3 + b] = F[a * 3] * F[b * 3] + F[a * 3 + 1] * F[b * 3 + 1] + F[a * 3 + 2] * F[b * 3 + 2];
  t[0] -= 1.f; t[4] -= 1.f; t[8] -= 1.f;
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    m[q] = F[q];
    m[9 + q] = t[q] * k.solidK;
  }
}

// accel_i = sum over elastic j of kspikydiff (h - r)^2 / r m_j dens_j
// dens_i (sigma_i + sigma_j) x_ij; overwrites accel (sph_solid.cu:627).
__global__ void k_solid_force(KP k, const uint32_t* __restrict__ key,
                              const int* __restrict__ cs, const float4* __restrict__ pos,
                              const float4* __restrict__ phys, const int* __restrict__ perm,
                              const float* __restrict__ mat, float4* __restrict__ accel) {
  int i = xcd_index();
  if (i >= k.n) return;
  float4 pi = pos[i];
  if (type_of(pi) != SPHX_TYPE_ELASTIC) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  const float idens = phys[i].y;
  float si[9];
  const float* mi = mat + (size_t)perm[i] * 36 + 9;
#pragma unroll
  for (int q = 0; q < 9; ++q) si[q] = mi[q];
  float rx = 0.f, ry = 0.f, rz = 0.f;
  for_each_neighbor(k, cs, gc, [&](int j) {
    const float4 pj = pos[j];
    float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
    float dsq = dx * dx + dy * dy + dz * dz;
    if (dsq < k.h2 && dsq > 0.f && type_of(pj) == SPHX_TYPE_ELASTIC) {
      float r = sqrtf(dsq);
      float c = k.h - r;
      const float4 qj = phys[j];
      float cm = k.kspikydiff * c * c / r * qj.x * qj.y * idens;
      const float* mj = mat + (size_t)perm[j] * 36 + 9;
      float s[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) s[q] = si[q] + mj[q];
      float tx = s[0] * dx + s[1] * dy + s[2] * dz;
      float ty = s[3] * dx + s[4] * dy + s[5] * dz;
      float tz = s[6] * dx + s[7] * dy + s[8] * dz;
      rx += tx * cm; ry += ty * cm; rz += tz * cm;
    }
  });
  accel[i] = make_float4(rx, ry, rz, 0.f);
}

// accel_i += sum_j C cos(1.5 pi r / tR) / r x_ij over 0 < |x_ij|^2 <= tR with
// tR = K h.  The reference compares the squared distance with the radius
// (:596) and evaluates the cosine term in double; both are kept.  Every
// particle type takes part (the reference's type test is commented out).
__global__ void k_surface_tension(KP k, const uint32_t* __restrict__ key,
                                  const int* __restrict__ cs, const float4* __restrict__ pos,
                                  float4* __restrict__ accel) {
  int i = xcd_index();
  if (i >= k.n) return;
  uint32_t gc = key[i];
  if ((int)gc == k.gridtotal) return;
  const float4 pi = pos[i];
  const float tR = sqrtf(k.st_K * k.st_K * k.h2);
  const double C = k.st_C;
  float rx = 0.f, ry = 0.f, rz = 0.f;
  for_each_neighbor(k, cs, gc, [&](int j) {
    const float4 pj = pos[j];
    float dx = (pi.x - pj.x) * k.ss, dy = (pi.y - pj.y) * k.ss, dz = (pi.z - pj.z) * k.ss;
    float dsq = dx * dx + dy * dy + dz * dz;
    if (dsq != 0.f && dsq <= tR) {
      float r = sqrtf(dsq);
      float fm = (float)(C * cos(1.5 * 3.14159 * (double)r / (double)tR) / (double)r);
      rx += dx * fm; ry += dy * fm; rz += dz * fm;
    }
  });
  float4 a = accel[i];
  accel[i] = make_float4(a.x + rx, a.y + ry, a.z + rz, a.w);
}

int mpm_sync_cold(sphx_context* c);
void mpm_invalidate(sphx_context* c);

static int need_grid(sphx_context* c, const char* who) {
  SPHX_CHECK(c, who, "null context");
  SPHX_TRY(ensure_device(c));
  SPHX_CHECK(c->sorted_valid, who, "grid not built (call sphx_grid_build first)");
  return 0;
}

}  // namespace sphx

using namespace sphx;

extern "C" {

int sphx_solid_tensor(sphx_context* c) {
  SPHX_TRY(need_grid(c, "sphx_solid_tensor"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  SPHX_TRY(mpm_sync_cold(c));   // MPM keeps stresses hot; the cold mats are the truth here
  timer_begin(c, "solid_tensor");
  k_solid_tensor<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, c->key.p, c->cell_start.p, c->POS(), c->VEL(), c->PHYS(), c->PERM(), c->cold_mat.p);
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  mpm_invalidate(c);
  return 0;
}

int sphx_solid_force(sphx_context* c) {
  SPHX_TRY(need_grid(c, "sphx_solid_force"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  SPHX_TRY(mpm_sync_cold(c));
  timer_begin(c, "solid_force");
  k_solid_force<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, c->key.p, c->cell_start.p, c->POS(), c->PHYS(), c->PERM(), c->cold_mat.p, c->ACCEL());
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

int sphx_surface_tension(sphx_context* c) {
  SPHX_TRY(need_grid(c, "sphx_surface_tension"));
  KP k = make_kp(c);
  if (k.n == 0) return 0;
  timer_begin(c, "surface_tension");
  k_surface_tension<<<blocks_for(k.n, 256), 256, 0, c->stream>>>(
      k, c->key.p, c->cell_start.p, c->POS(), c->ACCEL());
  timer_end(c);
  SPHX_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
