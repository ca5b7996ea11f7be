// Oracle (TEST INFRASTRUCTURE ONLY): SPH elastic solid and surface tension.
//
//   ComputeSolidTensor + contributeVGrad   sph_solid.cu:480-555
//   ComputeSolidForce + ContributeSolidForce sph_solid.cu:557-635
//   SurfaceTension + contributeSurfaceTensionYT fluid_system_kern.cu:569-625
//
// Same loop order as the reference (27 cells in neighborid order, members
// ascending, one accumulator across cells) and the same expression order;
// T = float is the faithful FP32 restatement, T = double the FP64 twin.
// Matrices are row-major [n][9] (cmat3, catpaw/geometry.h:142-215).
#include <cmath>
#include <cstring>

#include "oracle.hpp"

namespace {

constexpr uint32_t UNDEF = 0xFFFFFFFFu;

template <class T> inline T ld(const void* a, size_t i) { return static_cast<const T*>(a)[i]; }
template <class T> inline T& at(void* a, size_t i) { return static_cast<T*>(a)[i]; }

// mat3prod (geometry.h:195-201): c = a b
template <class T> void mat3prod(const T* a, const T* b, T* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      c[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

// Velocity gradient, deformation gradient update F += (grad v F) dt, strain
// F F^T - I and stress solidK * strain of every elastic particle.
template <class T>
void solid_tensor(const sphx_ParamCarrier& P, OrcState& s, void* F, void* stress) {
  const T h = P.smoothradius, r2 = h * h, ss = P.simscale;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] != SPHX_TYPE_ELASTIC) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    T ip[3], iv[3];
    for (int q = 0; q < 3; ++q) {
      ip[q] = ld<T>(s.pos, 3 * (size_t)i + q);
      iv[q] = ld<T>(s.vel, 3 * (size_t)i + q);
    }
    T g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        T d[3], jv[3];
        for (int q = 0; q < 3; ++q) {
          jv[q] = ld<T>(s.vel, 3 * (size_t)j + q);
          d[q] = (ip[q] - ld<T>(s.pos, 3 * (size_t)j + q)) * ss;
        }
        T dsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (dsq < r2 && dsq > 0) {
          T r = std::sqrt(dsq);
          T cc = h - r;
          T cm = T(P.kspikydiff) * cc * cc / r * ld<T>(s.mass, j) * ld<T>(s.dens, j);
          for (int q = 0; q < 3; ++q) jv[q] = (jv[q] - iv[q]) * cm;
          for (int a = 0; a < 3; ++a)
            for (int bb = 0; bb < 3; ++bb) g[a * 3 + bb] += jv[a] * d[bb];
        }
      }
    }
    T* Fi = static_cast<T*>(F) + 9 * (size_t)i;
    T tmp[9], tr[9], st[9];
    mat3prod(g, Fi, tmp);
    for (int k = 0; k < 9; ++k) Fi[k] += tmp[k] * T(P.dt);
    for (int a = 0; a < 3; ++a)
      for (int bb = 0; bb < 3; ++bb) tr[a * 3 + bb] = Fi[bb * 3 + a];
    mat3prod(Fi, tr, st);
    st[0] -= 1; st[4] -= 1; st[8] -= 1;
    T* Si = static_cast<T*>(stress) + 9 * (size_t)i;
    for (int k = 0; k < 9; ++k) Si[k] = st[k] * T(P.solidK);
  }
}

// accel_i = sum over elastic j of kspikydiff (h - r)^2 / r m_j dens_j dens_i
// (sigma_i + sigma_j) x_ij, overwriting accel (sph_solid.cu:627).
template <class T>
void solid_force(const sphx_ParamCarrier& P, OrcState& s, const void* stress) {
  const T h = P.smoothradius, r2 = h * h, ss = P.simscale;
  const T* S = static_cast<const T*>(stress);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] != SPHX_TYPE_ELASTIC) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    T ip[3];
    for (int q = 0; q < 3; ++q) ip[q] = ld<T>(s.pos, 3 * (size_t)i + q);
    const T idens = ld<T>(s.dens, i);
    T res[3] = {0, 0, 0};
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        T d[3];
        for (int q = 0; q < 3; ++q) d[q] = (ip[q] - ld<T>(s.pos, 3 * (size_t)j + q)) * ss;
        T dsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (dsq < r2 && dsq > 0) {
          T r = std::sqrt(dsq);
          T cc = h - r;
          T jdens = ld<T>(s.dens, j);
          if (s.type[j] == SPHX_TYPE_ELASTIC) {
            T cm = T(P.kspikydiff) * cc * cc / r * ld<T>(s.mass, j) * jdens * idens;
            T sij[9];
            for (int k = 0; k < 9; ++k) sij[k] = S[9 * (size_t)i + k] + S[9 * (size_t)j + k];
            T t[3];
            for (int a = 0; a < 3; ++a) t[a] = sij[3 * a] * d[0] + sij[3 * a + 1] * d[1] + sij[3 * a + 2] * d[2];
            for (int a = 0; a < 3; ++a) res[a] += t[a] * cm;
          }
        }
      }
    }
    for (int q = 0; q < 3; ++q) at<T>(s.accel, 3 * (size_t)i + q) = res[q];
  }
}

// accel_i += sum_j C cos(1.5 pi r / tR) / r x_ij over 0 < |x_ij|^2 <= tR,
// tR = K h.  The reference compares the squared distance with the radius
// tR (:596) and evaluates the cosine term in double (double literals); both
// are kept.
template <class T>
void surface_tension(const sphx_ParamCarrier& P, OrcState& s) {
  const float sr2f = P.smoothradius * P.smoothradius;
  const float tR2f = P.surfaceTensionK * P.surfaceTensionK * sr2f;
  const T tR = (T)std::sqrt(tR2f);
  const T ss = P.simscale;
  const double C = P.surfaceTensionFluidC;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    T ip[3];
    for (int q = 0; q < 3; ++q) ip[q] = ld<T>(s.pos, 3 * (size_t)i + q);
    T res[3] = {0, 0, 0};
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        T d[3];
        for (int q = 0; q < 3; ++q) d[q] = (ip[q] - ld<T>(s.pos, 3 * (size_t)j + q)) * ss;
        T dsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (dsq != 0 && dsq <= tR) {
          T r = std::sqrt(dsq);
          T fm = (T)(C * std::cos(1.5 * 3.14159 * (double)r / (double)tR) / (double)r);
          for (int q = 0; q < 3; ++q) res[q] += d[q] * fm;
        }
      }
    }
    for (int q = 0; q < 3; ++q) at<T>(s.accel, 3 * (size_t)i + q) += res[q];
  }
}

}  // namespace

extern "C" {

void orc_f32_solid_tensor(const sphx_ParamCarrier* P, OrcState* s, float* F, float* stress) {
  solid_tensor<float>(*P, *s, F, stress);
}
void orc_f64_solid_tensor(const sphx_ParamCarrier* P, OrcState* s, double* F, double* stress) {
  solid_tensor<double>(*P, *s, F, stress);
}
void orc_f32_solid_force(const sphx_ParamCarrier* P, OrcState* s, const float* stress) {
  solid_force<float>(*P, *s, stress);
}
void orc_f64_solid_force(const sphx_ParamCarrier* P, OrcState* s, const double* stress) {
  solid_force<double>(*P, *s, stress);
}
void orc_f32_surface_tension(const sphx_ParamCarrier* P, OrcState* s) { surface_tension<float>(*P, *s); }
void orc_f64_surface_tension(const sphx_ParamCarrier* P, OrcState* s) { surface_tension<double>(*P, *s); }

}  // extern "C"
