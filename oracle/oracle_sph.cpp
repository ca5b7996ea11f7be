// CPU oracle (test infrastructure only): grid build, WCSPH and IISPH.
// Restates SPH_Solver/fluids/fluid_system_kern.cu and
// SPH_Solver/fluids/fluid_system_host.cu (cited per function).
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

namespace {

template <class T>
struct V3 {
  T x, y, z;
};
template <class T> inline V3<T> ld3(const void* a, int i) {
  const T* p = static_cast<const T*>(a) + 3 * (size_t)i;
  return {p[0], p[1], p[2]};
}
template <class T> inline void st3(void* a, int i, V3<T> v) {
  T* p = static_cast<T*>(a) + 3 * (size_t)i;
  p[0] = v.x; p[1] = v.y; p[2] = v.z;
}
template <class T> inline T& at(void* a, int i) { return static_cast<T*>(a)[i]; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

const uint32_t UNDEF = 0xFFFFFFFFu;

template <class T>
struct Ctx {
  const sphx_ParamCarrier& P;
  const sphx_SolverOptions& O;
  OrcState& s;
  T h, ss;
  // member range of cell (gridcnt == 0 -> empty), as contributeX() do
  bool range(int cell, int& b, int& e) const {
    if (s.gridcnt[cell] == 0) return false;
    b = s.gridoff[cell];
    e = b + s.gridcnt[cell];
    return true;
  }
};

// kern / nablakern (fluid_system_kern.cu:637-671)
template <class T> T kern(V3<T> r, T h, T kspline) {
  T q = std::sqrt(dot(r, r)) / h;
  if (q > 2) return 0;
  else if (q > 1) return T(0.25f) * std::pow(2 - q, T(3.0f)) * kspline;
  else return (1 - T(1.5f) * q * q + T(0.75f) * std::pow(q, T(3.0f))) * kspline;
}
// nablakern returns r * fac; nabla_fac gives fac for |r| = rmod
template <class T> T nabla_fac(T rmod, T h, T kspline) {
  T q = rmod / h;
  if (q > 2) return 0;
  else if (q > 1) return T(0.75f) * (2 - q) * (2 - q) * (-1) / h / rmod * kspline;
  else return (-3 * q + T(2.25f) * q * q) / h / rmod * kspline;
}
template <class T> V3<T> nablakern(V3<T> r, T h, T kspline) {
  T rmod = std::sqrt(dot(r, r));
  T q = rmod / h;
  if (q > 2) return {0, 0, 0};
  return r * nabla_fac<T>(rmod, h, kspline);
}

// ------------------------------------------------------------ grid build --
// InitialSort (fluid_system_kern.cu:36-64), always FP32 as the reference.
void initial_sort(const sphx_ParamCarrier& P, OrcState& s) {
  const float* pos = static_cast<const float*>(s.pos);
  for (int i = 0; i < s.n; ++i) {
    float gx = (pos[3 * i] - P.gridmin.x) * P.gridIdfac.x;
    float gy = (pos[3 * i + 1] - P.gridmin.y) * P.gridIdfac.y;
    float gz = (pos[3 * i + 2] - P.gridmin.z) * P.gridIdfac.z;
    int cx = int(gx), cy = int(gy), cz = int(gz);
    int gs = (cy * P.gridres.z + cz) * P.gridres.x + cx;
    if (cx >= 1 && cx < P.gridres.x - 1 && cy >= 1 && cy < P.gridres.y - 1 && cz >= 1 &&
        cz < P.gridres.z - 1) {
      s.mgcell[i] = (uint32_t)gs;
    } else {
      s.mgcell[i] = UNDEF;
    }
    s.midsort[i] = (uint32_t)i;
  }
}

}  // namespace

extern "C" {

void orc_initial_sort(const sphx_ParamCarrier* P, OrcState* s) { initial_sort(*P, *s); }

// thrust::sort_by_key (fluid_system_host.cu:164-166): ascending, stable
// (std::stable_sort is the closed form of a stable key sort).
void orc_sort_by_key(OrcState* s) {
  std::vector<uint32_t> idx(s->n);
  std::iota(idx.begin(), idx.end(), 0u);
  std::stable_sort(idx.begin(), idx.end(),
                   [&](uint32_t a, uint32_t b) { return s->mgcell[a] < s->mgcell[b]; });
  std::vector<uint32_t> k(s->n), v(s->n);
  for (int i = 0; i < s->n; ++i) {
    k[i] = s->mgcell[idx[i]];
    v[i] = s->midsort[idx[i]];
  }
  std::memcpy(s->mgcell, k.data(), s->n * sizeof(uint32_t));
  std::memcpy(s->midsort, v.data(), s->n * sizeof(uint32_t));
}

// memsets (host.cu:149-151) + CalcFirstCnt (kern.cu:67-82) + GetCnt (:84-95)
void orc_cell_ranges(OrcState* s) {
  std::fill(s->gridoff, s->gridoff + s->gridtotal, 0);
  std::fill(s->gridcnt, s->gridcnt + s->gridtotal, 0);
  const uint32_t* c = s->mgcell;
  int n = s->n;
  for (int i = 0; i < n; ++i) {
    if (i == 0 || c[i] != c[i - 1])
      if (c[i] != UNDEF) s->gridoff[c[i]] = i;
    if (i != 0 && c[i] != c[i - 1] && c[i - 1] != UNDEF) s->gridcnt[c[i - 1]] = i;
    if (i == n - 1 && c[i] != UNDEF) s->gridcnt[c[i]] = i + 1;
  }
  std::vector<uint32_t> ndx(n, UNDEF);
  for (int i = 0; i < n; ++i) {
    if (c[i] != UNDEF) {
      ndx[i] = i - s->gridoff[c[i]];
    }
  }
  // GetCnt: cnt = tail - head for the first particle of each cell
  for (int i = 0; i < n; ++i)
    if (c[i] != UNDEF && ndx[i] == 0) s->gridcnt[c[i]] -= s->gridoff[c[i]];
  if (s->mgndx) std::memcpy(s->mgndx, ndx.data(), n * sizeof(uint32_t));
}

// RearrangeData (kern.cu:98-121) on the fields of the state (FP32).  The
// sorted slot takes a copy of particle midsort[s]; an out-of-grid slot keeps
// its pre-sort contents with pos = (-1000)^3.  MFidTable: mode 0 = the
// intended inverse permutation (bornid of the particle now in s -> s);
// mode 1 = the reference's literal keying by the pre-sort occupant (A8).
void orc_rearrange(OrcState* s, int mfid_mode) {
  int n = s->n;
  auto cp3 = [&](void* a) {
    std::vector<float> old(static_cast<float*>(a), static_cast<float*>(a) + 3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      if (s->mgcell[i] == UNDEF) continue;
      uint32_t ex = s->midsort[i];
      for (int k = 0; k < 3; ++k) static_cast<float*>(a)[3 * i + k] = old[3 * ex + k];
    }
  };
  auto cp1 = [&](auto* a) {
    using E = std::remove_pointer_t<decltype(a)>;
    std::vector<E> old(a, a + n);
    for (int i = 0; i < n; ++i) {
      if (s->mgcell[i] == UNDEF) continue;
      a[i] = old[s->midsort[i]];
    }
  };
  std::vector<int> born_before(s->bornid, s->bornid + n);
  cp3(s->pos);
  cp1(s->type);
  cp3(s->vel);
  cp3(s->veleval);
  cp3(s->accel);
  cp1(static_cast<float*>(s->pressure));
  cp1(static_cast<float*>(s->dens));
  cp1(static_cast<float*>(s->restdens));
  cp1(static_cast<float*>(s->mass));
  cp1(static_cast<float*>(s->visc));
  cp1(s->bornid);
  float* pos = static_cast<float*>(s->pos);
  for (int i = 0; i < n; ++i) {
    if (s->mgcell[i] != UNDEF) {
      int b = mfid_mode == 1 ? born_before[i] : s->bornid[i];
      if (s->MFidTable && b >= 0 && b < n) s->MFidTable[b] = i;
    } else {
      pos[3 * i] = pos[3 * i + 1] = pos[3 * i + 2] = -1000.f;
    }
  }
}

}  // extern "C"

namespace {

// ------------------------------------------------------------------ WCSPH --
// ComputeBoundaryVolume + contributeDensity_boundary (kern.cu:183-212,249-277)
template <class T>
void boundary_volume(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s) {
  const bool cubic = O.wcsph_kernel == 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] != SPHX_TYPE_BOUNDARY) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    T sum = 0;
    V3<T> p = ld3<T>(s.pos, i);
    T r2 = T(P.smoothradius) * T(P.smoothradius);
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        if (s.type[j] != SPHX_TYPE_BOUNDARY) continue;
        V3<T> d = (p - ld3<T>(s.pos, j)) * T(P.simscale);
        if (cubic) {
          sum += kern<T>(d, T(P.smoothradius), T(P.kspline));
        } else {
          T dsq = d.x * d.x + d.y * d.y + d.z * d.z;
          if (dsq < r2) {
            T cc = r2 - dsq;
            sum += cc * cc * cc;
          }
        }
      }
    }
    if (!cubic) sum = sum * T(P.kpoly6);
    if (sum == 0) sum = 1;
    at<T>(s.dens, i) = 1 / sum;
  }
}

// ComputeDensityPressure + contributeDensity (kern.cu:214-318)
template <class T>
void density_pressure(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s) {
  const bool cubic = O.wcsph_kernel == 1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == SPHX_TYPE_BOUNDARY) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    T sum = 0;
    T dens = at<T>(s.restdens, i);
    V3<T> p = ld3<T>(s.pos, i);
    T r2 = T(P.smoothradius) * T(P.smoothradius);
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        T massj;
        if (s.type[j] == 0) massj = at<T>(s.mass, j);
        else if (s.type[j] == 1) massj = at<T>(s.restdens, i) * at<T>(s.dens, j);
        else massj = at<T>(s.mass, j);   // uninitialised in the reference (A5)
        V3<T> d = (p - ld3<T>(s.pos, j)) * T(P.simscale);
        T dsq = d.x * d.x + d.y * d.y + d.z * d.z;
        if (cubic) {
          sum += kern<T>(d, T(P.smoothradius), T(P.kspline)) * massj;
        } else if (dsq < r2) {
          T cc = r2 - dsq;
          sum += cc * cc * cc * massj;
        }
      }
    }
    if (!cubic) sum = sum * T(P.kpoly6);
    if (sum == 0) sum = 1;
    at<T>(s.dens, i) = 1 / sum;
    // 2.5 * dens * (pow(sum/dens, 7) - 1) in double (kern.cu:310)
    double B = O.eos_B > 0 ? (double)O.eos_B : 2.5 * (double)dens;
    T x = sum / dens;
    T pw = std::pow(x, T(O.eos_gamma));
    T pr = (T)(B * (double)(pw - 1));
    if (pr < 0) pr = 0;
    at<T>(s.pressure, i) = pr;
  }
}

// ComputeForce + contributeForce_new (kern.cu:325-424), per-cell partials
template <class T>
void force(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s) {
  const bool cubic = O.wcsph_kernel == 1;
  const T h = P.smoothradius;
  const T r2 = h * h;
  const T rsup2 = cubic ? 4 * r2 : r2;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == 1) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    V3<T> ipos = ld3<T>(s.pos, i);
    V3<T> iveleval = ld3<T>(s.vel, i);
    T ipress = at<T>(s.pressure, i), idens = at<T>(s.dens, i), ivisc = at<T>(s.visc, i);
    T irest = at<T>(s.restdens, i);
    V3<T> fsum = {0, 0, 0};
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      V3<T> force = {0, 0, 0};
      if (s.gridcnt[cell] != 0) {
        int b = s.gridoff[cell], e = b + s.gridcnt[cell];
        for (int j = b; j < e; ++j) {
          V3<T> dist = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
          T dsq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
          if (!(dsq < rsup2 && dsq > 0)) continue;
          T r = std::sqrt(dsq);
          T cm1;
          if (cubic) {
            cm1 = nabla_fac<T>(r, h, T(P.kspline));
          } else {
            T cc = h - r;
            cm1 = T(P.kspikydiff) * cc * cc / r;
          }
          T cm = cm1 * at<T>(s.mass, j) * at<T>(s.dens, j);
          if (s.type[j] == 0) {
            T pm = (T)(-0.5 * (double)cm * (double)at<T>(s.pressure, j) * (double)idens);
            force = force + dist * pm;
            V3<T> vmr = iveleval - ld3<T>(s.vel, j);
            T vm = cm * (ivisc + at<T>(s.visc, j)) * idens;
            force = force + vmr * vm;
            if (O.art_visc_alpha > 0) {
              T vdx = dot(vmr, dist);
              if (vdx < 0) {
                T mu = h * vdx / (dsq + T(0.01) * r2);
                T rhobar = T(0.5) * (1 / idens + 1 / at<T>(s.dens, j));
                T Pi = -T(O.art_visc_alpha) * T(O.sound_speed) * mu / rhobar;
                force = force + dist * (-at<T>(s.mass, j) * Pi * cm1);
              }
            }
          } else if (s.type[j] == 1) {
            T pm = -cm1 * irest * at<T>(s.dens, j) * ipress * idens * idens;
            force = force + dist * pm;
            V3<T> vmr = iveleval - ld3<T>(s.vel, j);
            T pij = vmr.x * dist.x + vmr.y * dist.y + vmr.z * dist.z;
            if (pij < 0) {
              pij = pij / (dist.x * dist.x + dist.y * dist.y + dist.z * dist.z + r2 * T(0.01f));
              pij = pij * T(P.bvisc) * h / 2 / irest;
              pij = cm1 * irest * at<T>(s.dens, j) * pij;
              force = force + dist * pij;
            }
          }
        }
      }
      fsum = fsum + force;
    }
    st3<T>(s.accel, i, fsum);
  }
}

// AdvanceParticles (kern.cu:426-534)
template <class T>
void advance(const sphx_ParamCarrier& P, OrcState& s) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == 1) continue;
    if (s.mgcell[i] == UNDEF) {
      st3<T>(s.pos, i, {-1000, -1000, -1000});
      st3<T>(s.vel, i, {0, 0, 0});
      continue;
    }
    T dt = P.dt, ss = P.simscale;
    V3<T> accel = ld3<T>(s.accel, i);
    accel = accel + V3<T>{T(P.gravity.x), T(P.gravity.y), T(P.gravity.z)};
    V3<T> vel = ld3<T>(s.vel, i);
    V3<T> vnext = accel * dt + vel;
    st3<T>(s.pos, i, ld3<T>(s.pos, i) + vnext * dt * (1 / ss));
    st3<T>(s.veleval, i, (vel + vnext) * T(0.5));
    st3<T>(s.vel, i, vnext);
  }
}

// ------------------------------------------------------------------ IISPH --
// ComputeDensityIISPH + contributeIIDensity (kern.cu:1067-1115)
template <class T>
void ii_density(const sphx_ParamCarrier& P, OrcState& s) {
  const T h = P.smoothradius, ks = P.kspline;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == SPHX_TYPE_BOUNDARY) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    T sum = 0;
    V3<T> ipos = ld3<T>(s.pos, i);
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        T massj;
        if (s.type[j] == 0) massj = at<T>(s.mass, j);
        else if (s.type[j] == 1) massj = at<T>(s.restdens, i) * at<T>(s.dens, j);
        else massj = at<T>(s.mass, j);
        V3<T> xij = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
        sum += kern<T>(xij, h, ks) * massj;
      }
    }
    at<T>(s.dens, i) = sum;
  }
}

template <class T> inline bool supp(const sphx_SolverOptions& O, T dist, T h) {
  T lim = O.full_support ? 2 * h : h;
  return dist < lim && dist > 0;
}

// ComputeDii + contributeAcc_adv + contributeDii (kern.cu:673-798)
template <class T>
void ii_dii(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s) {
  const T h = P.smoothradius, ks = P.kspline, h2 = h * h;
  const T dt = P.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == SPHX_TYPE_BOUNDARY) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    V3<T> ipos = ld3<T>(s.pos, i), ivel = ld3<T>(s.vel, i);
    T irest = at<T>(s.restdens, i);
    V3<T> acc_adv = {0, 0, 0};
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      V3<T> accel = {0, 0, 0};
      for (int j = b; j < e; ++j) {
        V3<T> xij = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
        T dist = std::sqrt(dot(xij, xij));
        if (!(dist < 2 * h && dist > 0)) continue;
        V3<T> nw = nablakern<T>(xij, h, ks);
        if (s.type[j] == 1) {
          V3<T> vij = ivel - ld3<T>(s.vel, j);
          T pij = dot(xij, vij);
          if (pij < 0) {
            pij = pij / (dist * dist + h2 * T(0.01f));
            pij = pij * T(P.bvisc) * h / 2 / irest;
            pij = irest * at<T>(s.dens, j) * pij;
            accel = accel + nw * pij;
          }
        }
      }
      acc_adv = acc_adv + accel;
    }
    acc_adv = acc_adv + V3<T>{T(P.gravity.x), T(P.gravity.y), T(P.gravity.z)};
    st3<T>(s.veleval, i, ivel + acc_adv * dt);
    V3<T> dii = {0, 0, 0};
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        V3<T> xij = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
        T dist = std::sqrt(dot(xij, xij));
        if (!supp(O, dist, h)) continue;
        V3<T> nw = nablakern<T>(xij, h, ks);
        if (s.type[j] != SPHX_TYPE_BOUNDARY) dii = dii - nw * at<T>(s.mass, j);
        else if (O.boundary_pressure) dii = dii - nw * (irest * at<T>(s.dens, j));  // :763-764
      }
    }
    T rho = at<T>(s.dens, i);
    dii = dii * (1 / rho) * (1 / rho) * dt * dt;
    st3<T>(s.dii, i, dii);
  }
}

// ComputeAii + contributeDensChange + contributeAii (kern.cu:801-908)
template <class T>
void ii_aii(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s) {
  const T h = P.smoothradius, ks = P.kspline, dt = P.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == SPHX_TYPE_BOUNDARY) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    V3<T> ipos = ld3<T>(s.pos, i), ivel = ld3<T>(s.veleval, i);
    T irest = at<T>(s.restdens, i), mi = at<T>(s.mass, i), rhoi = at<T>(s.dens, i);
    T denschange = 0;
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      T sum = 0;
      for (int j = b; j < e; ++j) {
        V3<T> xij = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
        T dist = std::sqrt(dot(xij, xij));
        if (!supp(O, dist, h)) continue;
        V3<T> vij = ivel - ld3<T>(s.veleval, j);
        V3<T> nw = nablakern<T>(xij, h, ks);
        if (s.type[j] != SPHX_TYPE_BOUNDARY) sum += dot(vij, nw) * at<T>(s.mass, j) * dt;
        else if (O.boundary_pressure) sum += dot(vij, nw) * at<T>(s.dens, j) * irest * dt;  // :833-834
      }
      denschange += sum;
    }
    at<T>(s.rho_adv, i) = rhoi + denschange;
    at<T>(s.press_l, i) = (T)((double)O.warm_start * (double)at<T>(s.pressure, i));   // :899
    T aii = 0;
    V3<T> dii = ld3<T>(s.dii, i);
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      T sum = 0;
      for (int j = b; j < e; ++j) {
        if (s.type[j] == SPHX_TYPE_BOUNDARY) {
          if (!O.boundary_pressure) continue;
          V3<T> xij = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
          T dist = std::sqrt(dot(xij, xij));
          if (!supp(O, dist, h)) continue;
          V3<T> nw = nablakern<T>(xij, h, ks);
          sum += dot(dii, nw) * irest * at<T>(s.dens, j);
          continue;
        }
        V3<T> xij = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
        T dist = std::sqrt(dot(xij, xij));
        if (!supp(O, dist, h)) continue;
        V3<T> nw = nablakern<T>(xij, h, ks);
        V3<T> dji = nw * mi * (1 / rhoi) * (1 / rhoi) * dt * dt;
        V3<T> dtmp = dii - dji;
        sum += dot(dtmp, nw) * at<T>(s.mass, j);
      }
      aii += sum;
    }
    at<T>(s.aii, i) = aii;
  }
}

// Pressure_DP + contributeDP (kern.cu:910-958)
template <class T>
void ii_dp(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s) {
  const T h = P.smoothradius, ks = P.kspline, dt = P.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == SPHX_TYPE_BOUNDARY) continue;
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) continue;
    V3<T> ipos = ld3<T>(s.pos, i);
    V3<T> dp = {0, 0, 0};
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        if (s.type[j] == SPHX_TYPE_BOUNDARY) continue;
        V3<T> xij = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
        T dist = std::sqrt(xij.x * xij.x + xij.y * xij.y + xij.z * xij.z);
        if (!supp(O, dist, h)) continue;
        V3<T> nw = nablakern<T>(xij, h, ks);
        T rj = at<T>(s.dens, j);
        dp = dp + nw * (at<T>(s.press_l, j) * at<T>(s.mass, j) / rj / rj * (-1));
      }
    }
    dp = dp * dt * dt;
    st3<T>(s.dijpj, i, dp);
  }
}

// Pressure_Iter + contributeP_updateterm (kern.cu:960-1050).  Writes
// press_l1 (the reference copies it into press_l on the host, :536).
template <class T>
void ii_iter(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s) {
  const T h = P.smoothradius, ks = P.kspline, dt = P.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == SPHX_TYPE_BOUNDARY) {
      at<T>(s.residue, i) = 0;
      continue;
    }
    uint32_t gc = s.mgcell[i];
    if (gc == UNDEF) {
      at<T>(s.residue, i) = 0;   // left stale by the reference
      continue;
    }
    V3<T> ipos = ld3<T>(s.pos, i);
    V3<T> dijpj = ld3<T>(s.dijpj, i);
    T mi = at<T>(s.mass, i), rhoi = at<T>(s.dens, i), irest = at<T>(s.restdens, i);
    T pli = at<T>(s.press_l, i);
    T updateterm = 0;
    for (int c = 0; c < P.neighbornum; ++c) {
      int cell = (int)gc + P.neighborid[c];
      if (s.gridcnt[cell] == 0) continue;
      int b = s.gridoff[cell], e = b + s.gridcnt[cell];
      for (int j = b; j < e; ++j) {
        V3<T> xij = (ipos - ld3<T>(s.pos, j)) * T(P.simscale);
        T dist = std::sqrt(xij.x * xij.x + xij.y * xij.y + xij.z * xij.z);
        if (!supp(O, dist, h)) continue;
        V3<T> nw = nablakern<T>(xij, h, ks);
        if (s.type[j] != SPHX_TYPE_BOUNDARY) {
          V3<T> djj = ld3<T>(s.dii, j);
          V3<T> dji = nw * mi * (1 / rhoi) * (1 / rhoi) * dt * dt;
          V3<T> djkpk = ld3<T>(s.dijpj, j) - dji * pli;
          V3<T> tmp = dijpj - djj * at<T>(s.press_l, j) - djkpk;
          updateterm += at<T>(s.mass, j) * dot(tmp, nw);
        } else if (O.boundary_pressure) {
          updateterm += irest * at<T>(s.dens, j) * dot(dijpj, nw);   // :1000-1002
        }
      }
    }
    T omega = O.omega;
    T aii = at<T>(s.aii, i);
    T p1;
    if (std::fabs((double)aii) < 1e-6) p1 = pli;   // EPSILON = 1e-6 (A6)
    else p1 = (1 - omega) * pli + omega / aii * (irest - at<T>(s.rho_adv, i) - updateterm);
    if (O.clamp_pressure && p1 < 0) p1 = 0;   // commented out at kern.cu:1035-1036
    at<T>(s.press_l1, i) = p1;
    at<T>(s.residue, i) = -irest + (at<T>(s.rho_adv, i) + updateterm + aii * pli);
    V3<T> f = ld3<T>(s.dii, i) * pli + dijpj;
    st3<T>(s.accel, i, f * (1 / dt) * (1 / dt));
    at<T>(s.pressure, i) = p1;
  }
}

// IntegrateIISPH (kern.cu:1052-1064)
template <class T>
void ii_integrate(const sphx_ParamCarrier& P, OrcState& s) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.n; ++i) {
    if (s.type[i] == SPHX_TYPE_BOUNDARY) continue;
    if (s.mgcell[i] == UNDEF) continue;
    V3<T> v = ld3<T>(s.veleval, i) + ld3<T>(s.accel, i) * T(P.dt);
    st3<T>(s.vel, i, v);
    st3<T>(s.pos, i, ld3<T>(s.pos, i) + v * T(P.dt) * (1 / T(P.simscale)));
  }
}

// One pressure iteration as PressureSolve's loop body (host.cu:530-541):
// DP, Iter, press_l <- press_l1, mean signed residue; also returns the FP64
// mean |residue|/rho0 over fluid particles.
template <class T>
void ii_iteration(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s,
                  double* mean_signed, double* mean_abs_rel) {
  ii_dp<T>(P, O, s);
  ii_iter<T>(P, O, s);
  double sum = 0, ab = 0;
  long cnt = 0;
  for (int i = 0; i < s.n; ++i) {
    double r = (double)at<T>(s.residue, i);
    sum += r;
    if (s.type[i] != SPHX_TYPE_BOUNDARY && s.mgcell[i] != UNDEF) {
      // with the clamp the solve is a complementarity problem (p > 0 and
      // rho = rho0, or p = 0 and rho <= rho0): error |r| where the residual's
      // pressure p_l is positive, compression only where it is zero
      double e = O.clamp_pressure ? ((double)at<T>(s.press_l, i) > 0.0 ? std::fabs(r)
                                                                      : std::max(r, 0.0))
                                  : std::fabs(r);
      ab += e / (double)at<T>(s.restdens, i);
      ++cnt;
    }
  }
  std::memcpy(s.press_l, s.press_l1, sizeof(T) * s.n);
  if (mean_signed) *mean_signed = s.n ? sum / s.n : 0.0;
  if (mean_abs_rel) *mean_abs_rel = cnt ? ab / cnt : 0.0;
}

template <class T>
int ii_solve(const sphx_ParamCarrier& P, const sphx_SolverOptions& O, OrcState& s,
             double* err) {
  int it = 0;
  double e = 0;
  for (;;) {
    ii_iteration<T>(P, O, s, nullptr, &e);
    ++it;
    if ((it >= O.min_iters && e <= O.tol) || it >= O.max_iters) break;
  }
  if (err) *err = e;
  return it;
}

}  // namespace

extern "C" {

void orc_f32_boundary_volume(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { boundary_volume<float>(*P, *O, *s); }
void orc_f64_boundary_volume(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { boundary_volume<double>(*P, *O, *s); }
void orc_f32_density_pressure(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { density_pressure<float>(*P, *O, *s); }
void orc_f64_density_pressure(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { density_pressure<double>(*P, *O, *s); }
void orc_f32_force(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { force<float>(*P, *O, *s); }
void orc_f64_force(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { force<double>(*P, *O, *s); }
void orc_f32_advance(const sphx_ParamCarrier* P, OrcState* s) { advance<float>(*P, *s); }
void orc_f64_advance(const sphx_ParamCarrier* P, OrcState* s) { advance<double>(*P, *s); }
void orc_f32_ii_density(const sphx_ParamCarrier* P, OrcState* s) { ii_density<float>(*P, *s); }
void orc_f64_ii_density(const sphx_ParamCarrier* P, OrcState* s) { ii_density<double>(*P, *s); }
void orc_f32_ii_dii(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { ii_dii<float>(*P, *O, *s); }
void orc_f64_ii_dii(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { ii_dii<double>(*P, *O, *s); }
void orc_f32_ii_aii(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { ii_aii<float>(*P, *O, *s); }
void orc_f64_ii_aii(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { ii_aii<double>(*P, *O, *s); }
void orc_f32_ii_dp(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { ii_dp<float>(*P, *O, *s); }
void orc_f64_ii_dp(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { ii_dp<double>(*P, *O, *s); }
void orc_f32_ii_iter(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { ii_iter<float>(*P, *O, *s); }
void orc_f64_ii_iter(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s) { ii_iter<double>(*P, *O, *s); }
void orc_f32_ii_iteration(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s, double* ms, double* mar) { ii_iteration<float>(*P, *O, *s, ms, mar); }
void orc_f64_ii_iteration(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s, double* ms, double* mar) { ii_iteration<double>(*P, *O, *s, ms, mar); }
int orc_f32_ii_solve(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s, double* err) { return ii_solve<float>(*P, *O, *s, err); }
int orc_f64_ii_solve(const sphx_ParamCarrier* P, const sphx_SolverOptions* O, OrcState* s, double* err) { return ii_solve<double>(*P, *O, *s, err); }
void orc_f32_ii_integrate(const sphx_ParamCarrier* P, OrcState* s) { ii_integrate<float>(*P, *s); }
void orc_f64_ii_integrate(const sphx_ParamCarrier* P, OrcState* s) { ii_integrate<double>(*P, *s); }

}  // extern "C"
