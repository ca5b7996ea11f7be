// CPU oracle (test infrastructure only): drift-flux mixture stages.
// Restates SPH_Solver/fluids/multiphase.cu (formula specification; the file
// is not compiled in the reference) on the current buffer layout, with the
// cubic spline of support 2h and V_j = m_j / rho_j, cited per function.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>

extern "C" {
typedef struct OrcMF {
  int nphase;
  float rho[3], nu[3];
  float relax, sigma, cpress, cdiff;
  int boussinesq;
  float rho_ref;
  void* alpha;      /* T[n][4]: alpha_0..2, rho_m (fractions before the step) */
  void* alpha_new;  /* T[n][4] */
  void* drift;      /* T[n][3][3]: u_mk */
  void* mforce;     /* T[n][3]: g - a of the last step */
  void* amf;        /* T[n][3]: mixture acceleration added by mf_force */
} OrcMF;
}

namespace {

const uint32_t UNDEF = 0xFFFFFFFFu;

template <class T> inline T* P4(void* a, int i) { return static_cast<T*>(a) + 4 * (size_t)i; }
template <class T> inline T* P3(void* a, int i) { return static_cast<T*>(a) + 3 * (size_t)i; }
template <class T> inline T& S1(void* a, int i) { return static_cast<T*>(a)[i]; }

template <class T> T nabla_fac(T rmod, T h, T kspline) {
  T q = rmod / h;
  if (q > 2) return 0;
  else if (q > 1) return T(0.75f) * (2 - q) * (2 - q) * (-1) / h / rmod * kspline;
  else return (-3 * q + T(2.25f) * q * q) / h / rmod * kspline;
}

// Visit fluid neighbours j (not boundary) with 0 < r < 2h in the reference's
// 27-cell order; f(j, dist[3], r, cmterm) with cmterm = V_j * gradW factor.
template <class T, class F>
void fluid_pairs(const sphx_ParamCarrier& P, const OrcState& s, int i, F&& f) {
  const T h = P.smoothradius, ks = P.kspline;
  const T* pos = static_cast<const T*>(s.pos);
  uint32_t gc = s.mgcell[i];
  for (int c = 0; c < P.neighbornum; ++c) {
    int cell = (int)gc + P.neighborid[c];
    if (s.gridcnt[cell] == 0) continue;
    int b = s.gridoff[cell], e = b + s.gridcnt[cell];
    for (int j = b; j < e; ++j) {
      if (s.type[j] == SPHX_TYPE_BOUNDARY) continue;
      T d[3];
      for (int q = 0; q < 3; ++q) d[q] = (pos[3 * i + q] - pos[3 * j + q]) * T(P.simscale);
      T dsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (!(dsq < 4 * h * h && dsq > 0)) continue;
      T r = std::sqrt(dsq);
      T cm = nabla_fac<T>(r, h, ks) * static_cast<const T*>(s.mass)[j] /
             static_cast<const T*>(s.dens)[j];
      f(j, d, r, cm);
    }
  }
}

template <class T> inline bool skip(const OrcState& s, int i) {
  return s.type[i] == SPHX_TYPE_BOUNDARY || s.mgcell[i] == UNDEF;
}

// mfComputeDriftVel + contributeDriftVel (multiphase.cu:19-166)
template <class T>
void mf_drift(const sphx_ParamCarrier& P, OrcState& s, OrcMF& m) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (skip<T>(s, i)) continue;
    const T* ai = P4<T>(m.alpha, i);
    T ipress = S1<T>(s.pressure, i);
    T relax = T(m.relax) / S1<T>(s.visc, i);                 // :116
    T invdens = 1 / ai[3];
    T apre[3] = {0, 0, 0}, conc[3] = {0, 0, 0}, dv[3][3] = {}, ag[3][3] = {};
    for (int f = 0; f < m.nphase; ++f) {
      apre[f] = ai[f];
      conc[f] = apre[f] * T(m.rho[f]) * invdens;               // :122
    }
    fluid_pairs<T>(P, s, i, [&](int j, const T* d, T, T cm) {
      const T* aj = P4<T>(m.alpha, j);
      T pg[3], pgs[3] = {0, 0, 0};
      for (int f = 0; f < m.nphase; ++f) {
        pg[f] = cm * (-apre[f] * ipress + aj[f] * S1<T>(s.pressure, j));   // :62
        for (int q = 0; q < 3; ++q) pgs[q] += pg[f] * d[q] * conc[f];       // :64-65
        for (int q = 0; q < 3; ++q) ag[f][q] += (aj[f] - apre[f]) * cm * d[q];   // :67
      }
      for (int f = 0; f < m.nphase; ++f)
        for (int q = 0; q < 3; ++q) dv[f][q] -= relax * (pg[f] * d[q] - pgs[q]);   // :72
    });
    T dsum = 0, ssum[3] = {0, 0, 0}, st[3][3] = {};
    for (int f = 0; f < m.nphase; ++f) {                       // :139-152
      T t = apre[f];
      dsum += t * T(m.rho[f]) * T(m.rho[f]) * invdens;
      for (int q = 0; q < 3; ++q) {
        st[f][q] = t > T(0.0001) ? ag[f][q] / t : T(0);
        ssum[q] += st[f][q] * t * T(m.rho[f]) * invdens;
      }
    }
    const T* acc = P3<T>(m.mforce, i);
    for (int f = 0; f < m.nphase; ++f) {                       // :131-163
      T tdiff = (T(m.rho[f]) - dsum) * relax;
      T* u = static_cast<T*>(m.drift) + ((size_t)i * 3 + f) * 3;
      for (int q = 0; q < 3; ++q)
        u[q] = T(m.cpress) * dv[f][q] + acc[q] * tdiff -
               T(m.cdiff) * T(m.sigma) * (st[f][q] - ssum[q]);
    }
  }
}

// mfComputeAlphaAdvance + contributeAlphaChange (multiphase.cu:169-272),
// advected with the particle velocity
template <class T>
void mf_alpha(const sphx_ParamCarrier& P, OrcState& s, OrcMF& m) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    T* an = P4<T>(m.alpha_new, i);
    const T* ai = P4<T>(m.alpha, i);
    if (skip<T>(s, i)) {
      for (int q = 0; q < 4; ++q) an[q] = ai[q];
      continue;
    }
    const T* vi = P3<T>(s.vel, i);
    const T* ui = static_cast<const T*>(m.drift) + (size_t)i * 9;
    T ach[3] = {0, 0, 0};
    fluid_pairs<T>(P, s, i, [&](int j, const T* d, T, T cm) {
      const T* vj = P3<T>(s.vel, j);
      const T* aj = P4<T>(m.alpha, j);
      const T* uj = static_cast<const T*>(m.drift) + (size_t)j * 9;
      T vmrd = (vj[0] - vi[0]) * d[0] + (vj[1] - vi[1]) * d[1] + (vj[2] - vi[2]) * d[2];
      for (int f = 0; f < m.nphase; ++f) {
        ach[f] -= T(0.5) * cm * (aj[f] + ai[f]) * vmrd;                     // :207
        T vk[3];
        for (int q = 0; q < 3; ++q) vk[q] = aj[f] * uj[3 * f + q] + ai[f] * ui[3 * f + q];
        ach[f] -= cm * (vk[0] * d[0] + vk[1] * d[1] + vk[2] * d[2]);        // :212
      }
    });
    for (int f = 0; f < 3; ++f) {
      T dch = ach[f] * T(P.dt);
      if (dch < T(-0.99)) dch = T(-0.99);                                   // :258-261
      an[f] = f < m.nphase ? dch + ai[f] : T(0);
    }
    an[3] = ai[3];
  }
}

// mfComputeCorrection (multiphase.cu:275-349), on alpha_new in place; the
// mixture viscosity goes to the particle's visc (:348)
template <class T>
void mf_correction(const sphx_ParamCarrier&, OrcState& s, OrcMF& m) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (skip<T>(s, i)) continue;
    T* a = P4<T>(m.alpha_new, i);
    const T* ap = P4<T>(m.alpha, i);
    T sum = 0;
    for (int f = 0; f < m.nphase; ++f) {
      if (a[f] < 0) a[f] = 0;
      sum += a[f];
    }
    int flag = sum > 0;
    int maxcat = 0;
    for (int f = 1; f < m.nphase; ++f)
      if (ap[f] > ap[maxcat]) maxcat = f;
    T inv = 1 / (flag ? sum : T(1));
    T rm = 0, nm = 0;
    for (int f = 0; f < m.nphase; ++f) {
      a[f] = flag ? a[f] * inv : (f == maxcat ? T(1) : T(0));
      rm += a[f] * T(m.rho[f]);
      nm += a[f] * T(m.nu[f]);
    }
    for (int f = m.nphase; f < 3; ++f) a[f] = 0;
    a[3] = rm;
    S1<T>(s.visc, i) = nm;
  }
}

// drift stress (contributeForce_projectu, multiphase.cu:555-562) with the
// fractions before the step, plus Boussinesq buoyancy; added to veleval
template <class T>
void mf_force(const sphx_ParamCarrier& P, OrcState& s, OrcMF& m) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (skip<T>(s, i)) continue;
    const T* ai = P4<T>(m.alpha, i);
    const T* ui = static_cast<const T*>(m.drift) + (size_t)i * 9;
    T idens = 1 / S1<T>(s.dens, i);
    T fsum[3] = {0, 0, 0};
    fluid_pairs<T>(P, s, i, [&](int j, const T* d, T, T cm) {
      const T* aj = P4<T>(m.alpha, j);
      const T* uj = static_cast<const T*>(m.drift) + (size_t)j * 9;
      for (int f = 0; f < m.nphase; ++f) {
        T sj = cm * (uj[3 * f] * d[0] + uj[3 * f + 1] * d[1] + uj[3 * f + 2] * d[2]) * aj[f];
        T si = cm * (ui[3 * f] * d[0] + ui[3 * f + 1] * d[1] + ui[3 * f + 2] * d[2]) * ai[f];
        for (int q = 0; q < 3; ++q)
          fsum[q] += (sj * uj[3 * f + q] + si * ui[3 * f + q]) * T(m.rho[f]) * idens;
      }
    });
    if (m.boussinesq) {
      T b = (P4<T>(m.alpha_new, i)[3] - T(m.rho_ref)) / T(m.rho_ref);
      fsum[0] += b * T(P.gravity.x);
      fsum[1] += b * T(P.gravity.y);
      fsum[2] += b * T(P.gravity.z);
    }
    T* ve = P3<T>(s.veleval, i);
    T* am = P3<T>(m.amf, i);
    for (int q = 0; q < 3; ++q) {
      ve[q] += fsum[q] * T(P.dt);
      am[q] = fsum[q];
    }
  }
}

}  // namespace

extern "C" {
void orc_f32_mf_drift(const sphx_ParamCarrier* P, OrcState* s, OrcMF* m) { mf_drift<float>(*P, *s, *m); }
void orc_f64_mf_drift(const sphx_ParamCarrier* P, OrcState* s, OrcMF* m) { mf_drift<double>(*P, *s, *m); }
void orc_f32_mf_alpha(const sphx_ParamCarrier* P, OrcState* s, OrcMF* m) { mf_alpha<float>(*P, *s, *m); }
void orc_f64_mf_alpha(const sphx_ParamCarrier* P, OrcState* s, OrcMF* m) { mf_alpha<double>(*P, *s, *m); }
void orc_f32_mf_correction(const sphx_ParamCarrier* P, OrcState* s, OrcMF* m) { mf_correction<float>(*P, *s, *m); }
void orc_f64_mf_correction(const sphx_ParamCarrier* P, OrcState* s, OrcMF* m) { mf_correction<double>(*P, *s, *m); }
void orc_f32_mf_force(const sphx_ParamCarrier* P, OrcState* s, OrcMF* m) { mf_force<float>(*P, *s, *m); }
void orc_f64_mf_force(const sphx_ParamCarrier* P, OrcState* s, OrcMF* m) { mf_force<double>(*P, *s, *m); }
}
