// Feature map of the learned warm start (24 values per fluid particle),
// shared by the a_ii pass (which computes it in the same neighbour loop
// when the network is in use) and the stand-alone feature kernel.  Same map
// as the oracle's orc_mlp_features.
#pragma once

#include "nbr.hpp"

namespace sphx {

constexpr int MLP_NF = 24, MLP_H = 64, MLP_NO = 4;
constexpr float MLP_TSCALE = 10.f;   // target = (p* - p) / P * 10

struct MLPP {
  float p_scale;      // P (Pa)
  float u_scale;      // drift normalisation (m/s)
  float drift_blend;  // weight of the drift correction applied next step
};

// Neighbour sums of the features: kernel-weighted pressure mean, max and
// min, fluid and boundary counts, weighted density error, SPH pressure
// gradient.  add() takes every listed neighbour (r < 2h).
struct FeatAcc {
  float wsum = 0.f, wp = 0.f, pmax, pmin, wsrc = 0.f, gpx = 0.f, gpy = 0.f, gpz = 0.f;
  int nf = 0, nb = 0;
  __device__ __forceinline__ explicit FeatAcc(float p_i) : pmax(p_i), pmin(p_i) {}
  __device__ __forceinline__ void add(const KP& k, float dx, float dy, float dz, float r,
                                      float ir, const float4& pj, const float4& qj, float p_i) {
    if (!(r < 2.f * k.h && r > 0.f)) return;
    if (type_of(pj) == SPHX_TYPE_BOUNDARY) { ++nb; return; }
    ++nf;
    float w = cubic_w(r, k.inv_h, k.kspline);
    float pjp = fmaxf(0.f, qj.z);
    wsum += w;
    wp += w * pjp;
    pmax = fmaxf(pmax, pjp);
    pmin = fminf(pmin, pjp);
    // quotients through v_rcp_f32 (1 ulp) rather than IEEE division: these
    // are network inputs, not solver arithmetic
    wsrc += w * (qj.y - qj.w) * __builtin_amdgcn_rcpf(qj.w);
    float v = qj.x * __builtin_amdgcn_rcpf(qj.y) * cubic_gradf(r, ir, k.inv_h, k.kspline) * (pjp - p_i);
    gpx += v * dx; gpy += v * dy; gpz += v * dz;
  }
};

// The 24 features of fluid particle i from its own state and the sums.
__device__ __forceinline__ void feat_write(const KP& k, const MLPP& mp, const FeatAcc& a,
                                           const float4& qi, float radv, const float4& di,
                                           float aii, const float4& v, const float4* alpha,
                                           const float4* hist, int i, float* x) {
  const float rho0 = qi.w, rho = qi.y;
  const float p_i = fmaxf(0.f, qi.z);
  const float gn = sqrtf(k.g.x * k.g.x + k.g.y * k.g.y + k.g.z * k.g.z);
  const float3 gh = gn > 0.f ? f3(k.g.x / gn, k.g.y / gn, k.g.z / gn) : f3(0.f, -1.f, 0.f);
  const float P = mp.p_scale;
  float pav = a.wsum > 0.f ? a.wp / a.wsum : p_i;
  float pj_est = fabsf(aii) > 1e-12f ? fmaxf(0.f, (rho0 - radv) / aii) : 0.f;
  x[0] = p_i / P;
  x[1] = (radv - rho0) / rho0 * 100.f;
  x[2] = (rho - rho0) / rho0 * 100.f;
  x[3] = -aii * P / rho0;
  x[4] = pj_est / P;
  x[5] = pav / P;
  x[6] = a.pmax / P;
  x[7] = a.pmin / P;
  x[8] = (pav - p_i) / P;
  x[9] = a.nf / 64.f;
  x[10] = a.nb / 64.f;
  float4 al = alpha ? alpha[i] : make_float4(1.f, 0.f, 0.f, rho0);
  x[11] = alpha ? al.y : 0.f;
  x[12] = alpha ? al.w / rho0 : 1.f;
  x[13] = di.x * P / k.h; x[14] = di.y * P / k.h; x[15] = di.z * P / k.h;
  x[16] = v.x * k.dt / k.h; x[17] = v.y * k.dt / k.h; x[18] = v.z * k.dt / k.h;
  float den = rho0 * (gn > 0.f ? gn : 1.f);
  x[19] = (a.gpx * gh.x + a.gpy * gh.y + a.gpz * gh.z) / den;
  x[20] = hist[i].w / P * MLP_TSCALE;   // pressure increment of the last solve
  x[21] = a.gpz / den;
  x[22] = a.wsum > 0.f ? a.wsrc / a.wsum * 100.f : 0.f;
  x[23] = 1.f;
}

}  // namespace sphx
