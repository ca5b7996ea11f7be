// CPU oracle (test infrastructure only) of the MPM transfer:
//   P2G (APIC / PIC)   SPH_Solver/fluids/MpmSolver.cu:823-889 / :197-236
//   particle stress    :477-592 (rate form, Drucker-Prager :310-371, ProjectStress :373-378)
//   node update        :658-704 (+ GridCollisionDetection :595-656)
//   G2P (APIC / PIC)   :891-951 / :706-756
//   B-splines          :155-195
// The reference gathers node sums over the 27 SPH cells around each node;
// for an SPH cell at least as wide as the spline support that is exactly
// the particles with a non-zero weight, so the sums are restated here per
// particle with FP64 node accumulators (order-independent to rounding), the
// particles spread over OpenMP threads.
// kernel 0 is the reference's cubic spline (gradients in cell units, as
// the reference); kernel 1 the quadratic spline of the sand configuration.
// physical = 1 scales gradients to metres and uses the momentum-conserving
// node force m_p/(rho_p m_I) sigma grad N instead of the reference's
// sigma grad N / rho_p (:700).
#include "oracle.hpp"

#include <cmath>
#include <cstring>
#include <vector>

extern "C" {
typedef struct OrcMPM {
  int kernel;
  int physical;
  void* node_v;   /* T[N][3] */
  void* node_m;   /* T[N] */
  void* B;        /* T[n][9] */
  void* stress;   /* T[n][9] */
  void* G;        /* T[n] */
  void* K;        /* T[n] */
} OrcMPM;
}

namespace {

const uint32_t UNDEF = 0xFFFFFFFFu;

template <class T> T& at(void* a, size_t i) { return static_cast<T*>(a)[i]; }

// MpmSolver.cu:155-165 (constants as written there)
template <class T> T cubic(T x) {
  x = std::fabs(x);
  if (x > 2) return 0;
  if (x > 1) return T(-0.166666f) * x * x * x + x * x - 2 * x + T(1.33333f);
  return T(0.5f) * x * x * x - x * x + T(0.666666f);
}
// :173-186
template <class T> T dcubic(T x) {
  if (x <= -2) return 0;
  if (x <= -1) return T(0.5f) * x * x + 2 * x + 2;
  if (x <= 0) return T(-1.5f) * x * x - 2 * x;
  if (x < 1) return T(1.5f) * x * x - 2 * x;
  if (x < 2) return T(-0.5f) * x * x + 2 * x - 2;
  return 0;
}
template <class T> T quad(T x) {
  x = std::fabs(x);
  if (x < T(0.5)) return T(0.75) - x * x;
  if (x < T(1.5)) { T t = T(1.5) - x; return T(0.5) * t * t; }
  return 0;
}
template <class T> T dquad(T x) {
  T a = std::fabs(x);
  if (a < T(0.5)) return -2 * x;
  if (a < T(1.5)) return (x > 0 ? -1 : 1) * (T(1.5) - a);
  return 0;
}

struct Grid {
  double xmin[3];
  double dx;
  int res[3];
  double ss;
};

Grid grid_of(const sphx_ParamCarrier& P) {
  Grid g;
  g.xmin[0] = P.mpmXmin.x; g.xmin[1] = P.mpmXmin.y; g.xmin[2] = P.mpmXmin.z;
  g.dx = P.mpmcellsize;
  g.res[0] = P.mpmRes.x; g.res[1] = P.mpmRes.y; g.res[2] = P.mpmRes.z;
  g.ss = P.simscale;
  return g;
}

// GetMpmIdx (:35-38): y-major node id
inline size_t node_id(const Grid& g, int x, int y, int z) {
  return ((size_t)y * g.res[2] + z) * g.res[0] + x;
}

// Stencil of a particle: first node per axis and width (4 cubic, 3 quad).
template <class T>
void stencil(const Grid& g, int kernel, const T* p, int base[3], int& width) {
  for (int a = 0; a < 3; ++a) {
    T cs = (T)g.dx;
    T rel = (p[a] - (T)g.xmin[a]) / cs;
    if (kernel == 0) base[a] = (int)rel - 1;                 // xid-1 .. xid+2 (:418-425)
    else base[a] = (int)std::floor(rel - T(0.5));            // quadratic: 3 nodes
  }
  width = kernel == 0 ? 4 : 3;
}

template <class T> T N1(int kernel, T x) { return kernel == 0 ? cubic(x) : quad(x); }
template <class T> T dN1(int kernel, T x) { return kernel == 0 ? dcubic(x) : dquad(x); }

template <class T>
void node_pos(const Grid& g, int x, int y, int z, T out[3]) {
  int id[3] = {x, y, z};
  for (int a = 0; a < 3; ++a) out[a] = (T)g.xmin[a] + (T)id[a] * (T)g.dx;   // initMpm :51-53
}

// ------------------------------------------------------------------ P2G --
template <class T>
void p2g(const sphx_ParamCarrier& P, OrcState& s, OrcMPM& M, int apic) {
  Grid g = grid_of(P);
  size_t N = (size_t)g.res[0] * g.res[1] * g.res[2];
  std::vector<double> mv(3 * N, 0.0), m(N, 0.0);
  const T cs = (T)g.dx, ss = (T)g.ss;
  T iDp = cs * cs * ss * ss / (M.kernel == 0 ? T(3.0f) : T(4.0f));   // :844-845
  iDp = T(1) / iDp;
  // particles in parallel; the FP64 node sums are order-independent to
  // rounding, so they take atomic adds
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s.n; ++p) {
    if (s.type[p] == SPHX_TYPE_BOUNDARY || s.mgcell[p] == UNDEF) continue;
    const T* xp = &at<T>(s.pos, 3 * (size_t)p);
    const T* vp = &at<T>(s.vel, 3 * (size_t)p);
    const T mp = at<T>(s.mass, p);
    const T* Bp = &at<T>(M.B, 9 * (size_t)p);
    int base[3], w;
    stencil(g, M.kernel, xp, base, w);
    for (int a = 0; a < w; ++a)
      for (int b = 0; b < w; ++b)
        for (int c = 0; c < w; ++c) {
          int x = base[0] + a, y = base[1] + b, z = base[2] + c;
          if (x < 0 || x >= g.res[0] || y < 0 || y >= g.res[1] || z < 0 || z >= g.res[2]) continue;
          T xi[3];
          node_pos(g, x, y, z, xi);
          T r[3] = {(xi[0] - xp[0]) / cs, (xi[1] - xp[1]) / cs, (xi[2] - xp[2]) / cs};
          T wt = N1(M.kernel, r[0]) * N1(M.kernel, r[1]) * N1(M.kernel, r[2]);
          T vb[3] = {vp[0], vp[1], vp[2]};
          if (apic) {
            T d[3] = {(xi[0] - xp[0]) * ss, (xi[1] - xp[1]) * ss, (xi[2] - xp[2]) * ss};
            for (int q = 0; q < 3; ++q)
              vb[q] += (Bp[3 * q] * iDp) * d[0] + (Bp[3 * q + 1] * iDp) * d[1] + (Bp[3 * q + 2] * iDp) * d[2];
          }
          size_t id = node_id(g, x, y, z);
          const double dm = (double)(wt * mp);
#pragma omp atomic
          m[id] += dm;
          for (int q = 0; q < 3; ++q) {
            const double dq = (double)(vb[q] * mp * wt);
#pragma omp atomic
            mv[3 * id + q] += dq;
          }
        }
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)N; ++i) {
    if (m[i] > 1e-9) {
      for (int q = 0; q < 3; ++q) at<T>(M.node_v, 3 * i + q) = (T)(mv[3 * i + q] / m[i]);
      at<T>(M.node_m, i) = (T)m[i];
    } else {
      for (int q = 0; q < 3; ++q) at<T>(M.node_v, 3 * i + q) = 0;
      at<T>(M.node_m, i) = 0;
    }
  }
}

// ------------------------------------------------------- Drucker-Prager --
// :310-371 with the yield path completed as the formula of :361-368 (the
// reference's version returns a reference to a local and falls off the end
// on that path).
template <class T>
void drucker_prager(const sphx_ParamCarrier& P, T G, T K, const T eps[9], const T sig[9],
                    T plastic[9]) {
  for (int i = 0; i < 9; ++i) plastic[i] = 0;
  T s[9];
  std::memcpy(s, sig, sizeof(s));
  T trs = (s[0] + s[4] + s[8]) / T(3.0f);
  s[0] -= trs; s[4] -= trs; s[8] -= trs;
  T j2 = 0;
  for (int i = 0; i < 9; ++i) j2 += T(0.5) * s[i] * s[i];
  if (j2 <= T(0.0000001f)) return;
  j2 = std::sqrt(j2);
  T i1 = sig[0] + sig[4] + sig[8];
  T a_phi = (T)P.a_phi, a_psi = (T)P.a_psi, k_c = (T)P.k_c;
  T f = j2 + a_phi * i1 - k_c;
  if (f <= 0) return;
  T tr_e = eps[0] + eps[4] + eps[8];
  T se = 0;
  for (int i = 0; i < 9; ++i) se += s[i] * eps[i];
  T lambda = 3 * a_phi * K * tr_e + G / j2 * se;
  lambda /= (9 * a_phi * a_psi * K + G);
  for (int ii = 0; ii < 3; ++ii)
    for (int ij = 0; ij < 3; ++ij) {
      plastic[ii * 3 + ij] = G / j2 * s[ii * 3 + ij] * lambda;
      if (ii == ij) plastic[ii * 3 + ij] += 3 * a_psi * K * lambda;
    }
}

// ------------------------------------------------------- particle stress --
template <class T>
void stress(const sphx_ParamCarrier& P, OrcState& s, OrcMPM& M) {
  Grid g = grid_of(P);
  const T cs = (T)g.dx, rc = cs * (T)g.ss;
  const T dt = (T)P.dt;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s.n; ++p) {
    if (s.type[p] == SPHX_TYPE_BOUNDARY) continue;
    const T* xp = &at<T>(s.pos, 3 * (size_t)p);
    int base[3], w;
    stencil(g, M.kernel, xp, base, w);
    T vg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    T rho = 0;
    for (int a = 0; a < w; ++a)
      for (int b = 0; b < w; ++b)
        for (int c = 0; c < w; ++c) {
          int x = base[0] + a, y = base[1] + b, z = base[2] + c;
          if (x < 0 || x >= g.res[0] || y < 0 || y >= g.res[1] || z < 0 || z >= g.res[2]) continue;
          T xi[3];
          node_pos(g, x, y, z, xi);
          T r[3] = {(xp[0] - xi[0]) / cs, (xp[1] - xi[1]) / cs, (xp[2] - xi[2]) / cs};
          T n0 = N1(M.kernel, r[0]), n1 = N1(M.kernel, r[1]), n2 = N1(M.kernel, r[2]);
          T wt = n0 * n1 * n2;
          T nw[3] = {dN1(M.kernel, r[0]) * n1 * n2, dN1(M.kernel, r[1]) * n0 * n2,
                     dN1(M.kernel, r[2]) * n0 * n1};
          size_t id = node_id(g, x, y, z);
          const T* vi = &at<T>(M.node_v, 3 * id);
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) vg[3 * i + j] += vi[i] * nw[j];
          rho += at<T>(M.node_m, id) * wt / rc / rc / rc;
        }
    if (M.physical)
      for (int i = 0; i < 9; ++i) vg[i] /= rc;
    at<T>(s.dens, p) = rho;
    T de[9], om[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        de[3 * i + j] = (vg[3 * i + j] + vg[3 * j + i]) * T(0.5f);
        om[3 * i + j] = (vg[3 * i + j] - vg[3 * j + i]) * T(0.5f);
      }
    T trE = (de[0] + de[4] + de[8]) / T(3.0f);
    T* sig = &at<T>(M.stress, 9 * (size_t)p);
    const T G = at<T>(M.G, p), K = at<T>(M.K, p);
    T ds[9];
    for (int ii = 0; ii < 3; ++ii)
      for (int ij = 0; ij < 3; ++ij) {
        T v = 2 * G * de[3 * ii + ij];
        for (int ik = 0; ik < 3; ++ ik)
          v += om[3 * ii + ik] * sig[3 * ik + ij] - sig[3 * ii + ik] * om[3 * ik + ij];
        if (ii == ij) v += K * trE * T(3.0f) - 2 * G * trE;
        ds[3 * ii + ij] = v;
      }
    T pl[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (s.type[p] == SPHX_TYPE_GRANULAR) drucker_prager<T>(P, G, K, de, sig, pl);
    for (int i = 0; i < 9; ++i) {
      sig[i] += ds[i] * dt;
      sig[i] -= pl[i] * dt;
    }
    if (s.type[p] == SPHX_TYPE_GRANULAR && sig[0] + sig[4] + sig[8] > 0)   // ProjectStress
      for (int i = 0; i < 9; ++i) sig[i] = 0;
  }
}

// ---------------------------------------------------------- node update --
template <class T>
void collide(const sphx_ParamCarrier& P, const T x[3], T v[3]) {
  const T ss = (T)P.simscale, eps = T(1e-6f);   // EPSILON at this use site (A6)
  if (-(x[1] - (T)P.softminx.y) * ss > eps && v[1] < 0) { v[0] = v[1] = v[2] = 0; }
  if (-(x[0] - (T)P.softminx.x) * ss > eps && v[0] < 0) v[0] = 0;
  if (-(x[2] - (T)P.softminx.z) * ss > eps && v[2] < 0) v[2] = 0;
  if ((x[0] - (T)P.softmaxx.x) * ss > eps && -v[0] < 0) v[0] = 0;
  if ((x[2] - (T)P.softmaxx.z) * ss > eps && -v[2] < 0) v[2] = 0;
}

template <class T>
void node_update(const sphx_ParamCarrier& P, OrcState& s, OrcMPM& M) {
  Grid g = grid_of(P);
  size_t N = (size_t)g.res[0] * g.res[1] * g.res[2];
  std::vector<double> f(3 * N, 0.0);
  const T cs = (T)g.dx, rc = cs * (T)g.ss;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s.n; ++p) {
    if (s.type[p] == SPHX_TYPE_BOUNDARY || s.mgcell[p] == UNDEF) continue;
    const T* xp = &at<T>(s.pos, 3 * (size_t)p);
    const T* sig = &at<T>(M.stress, 9 * (size_t)p);
    const T rp = at<T>(s.dens, p), mp = at<T>(s.mass, p);
    int base[3], w;
    stencil(g, M.kernel, xp, base, w);
    for (int a = 0; a < w; ++a)
      for (int b = 0; b < w; ++b)
        for (int c = 0; c < w; ++c) {
          int x = base[0] + a, y = base[1] + b, z = base[2] + c;
          if (x < 0 || x >= g.res[0] || y < 0 || y >= g.res[1] || z < 0 || z >= g.res[2]) continue;
          size_t id = node_id(g, x, y, z);
          T mi = at<T>(M.node_m, id);
          if (mi < T(0.000000001f)) continue;
          T xi[3];
          node_pos(g, x, y, z, xi);
          T r[3] = {(xi[0] - xp[0]) / cs, (xi[1] - xp[1]) / cs, (xi[2] - xp[2]) / cs};
          T n0 = N1(M.kernel, r[0]), n1 = N1(M.kernel, r[1]), n2 = N1(M.kernel, r[2]);
          T nw[3] = {dN1(M.kernel, r[0]) * n1 * n2, dN1(M.kernel, r[1]) * n0 * n2,
                     dN1(M.kernel, r[2]) * n0 * n1};
          T sn[3];
          for (int q = 0; q < 3; ++q) sn[q] = sig[3 * q] * nw[0] + sig[3 * q + 1] * nw[1] + sig[3 * q + 2] * nw[2];
          T scale = M.physical ? mp / (rp * mi * rc) : T(1) / rp;
          for (int q = 0; q < 3; ++q) {
            const double fq = (double)(sn[q] * scale);
#pragma omp atomic
            f[3 * id + q] += fq;
          }
        }
  }
  const T dt = (T)P.dt;
  const T gv[3] = {(T)P.gravity.x, (T)P.gravity.y, (T)P.gravity.z};
#pragma omp parallel for schedule(static)
  for (int y = 0; y < g.res[1]; ++y)
    for (int z = 0; z < g.res[2]; ++z)
      for (int x = 0; x < g.res[0]; ++x) {
        size_t id = node_id(g, x, y, z);
        if (at<T>(M.node_m, id) < T(0.000000001f)) continue;
        T* v = &at<T>(M.node_v, 3 * id);
        for (int q = 0; q < 3; ++q) v[q] += ((T)f[3 * id + q] + gv[q]) * dt;
        T xi[3];
        node_pos(g, x, y, z, xi);
        collide<T>(P, xi, v);
      }
}

// ------------------------------------------------------------------ G2P --
template <class T>
void g2p(const sphx_ParamCarrier& P, OrcState& s, OrcMPM& M, int apic) {
  Grid g = grid_of(P);
  const T cs = (T)g.dx, ss = (T)g.ss, dt = (T)P.dt;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < s.n; ++p) {
    if (s.type[p] == SPHX_TYPE_BOUNDARY) continue;
    T* xp = &at<T>(s.pos, 3 * (size_t)p);
    int base[3], w;
    stencil(g, M.kernel, xp, base, w);
    T v[3] = {0, 0, 0}, B[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < w; ++a)
      for (int b = 0; b < w; ++b)
        for (int c = 0; c < w; ++c) {
          int x = base[0] + a, y = base[1] + b, z = base[2] + c;
          if (x < 0 || x >= g.res[0] || y < 0 || y >= g.res[1] || z < 0 || z >= g.res[2]) continue;
          T xi[3];
          node_pos(g, x, y, z, xi);
          T r[3] = {(xp[0] - xi[0]) / cs, (xp[1] - xi[1]) / cs, (xp[2] - xi[2]) / cs};
          T wt = N1(M.kernel, r[0]) * N1(M.kernel, r[1]) * N1(M.kernel, r[2]);
          const T* vi = &at<T>(M.node_v, 3 * node_id(g, x, y, z));
          T d[3] = {(xi[0] - xp[0]) * ss, (xi[1] - xp[1]) * ss, (xi[2] - xp[2]) * ss};
          for (int q = 0; q < 3; ++q) v[q] += vi[q] * wt;
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) B[3 * i + j] += wt * (vi[i] * d[j]);
        }
    T* vp = &at<T>(s.vel, 3 * (size_t)p);
    for (int q = 0; q < 3; ++q) vp[q] = v[q];
    if (apic) std::memcpy(&at<T>(M.B, 9 * (size_t)p), B, sizeof(B));
    for (int q = 0; q < 3; ++q) xp[q] += vp[q] * dt / ss;
  }
}

// ------------------------------------------------------ MpmGetMomentum --
// :107-153.  The node's SPH cell is initMpm's (:39-73): FP32 node position
// mpmXmin + id * mpmcellsize, cell (x - gridmin) * gridIdfac truncated,
// valid for 1 <= g <= res - 1.  The 27 cells are visited in neighborid
// order (y, then z, then x); cells past the grid's last layer are skipped
// (the reference reads past the row there, A11).  Invalid node cell: 0.
template <class T>
void mpm_momentum(const sphx_ParamCarrier& P, OrcState& s, OrcMPM& M) {
  const int rx = P.mpmRes.x, ry = P.mpmRes.y, rz = P.mpmRes.z;
  const int gx = P.gridres.x, gy = P.gridres.y, gz = P.gridres.z;
  const float fx0 = P.mpmXmin.x, fy0 = P.mpmXmin.y, fz0 = P.mpmXmin.z, fc = P.mpmcellsize;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < rx * ry * rz; ++i) {
    const int y = i / rx / rz, x = i % rx, z = i / rx % rz;
    const float px = fx0 + (float)x * fc, py = fy0 + (float)y * fc, pz = fz0 + (float)z * fc;
    const int cx = (int)((px - P.gridmin.x) * P.gridIdfac.x);
    const int cy = (int)((py - P.gridmin.y) * P.gridIdfac.y);
    const int cz = (int)((pz - P.gridmin.z) * P.gridIdfac.z);
    T* v = &at<T>(M.node_v, 3 * (size_t)i);
    T sum[3] = {0, 0, 0}, wsum = 0;
    const T cs = (T)fc;
    if (cx >= 1 && cx <= gx - 1 && cy >= 1 && cy <= gy - 1 && cz >= 1 && cz <= gz - 1) {
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          for (int c = -1; c <= 1; ++c) {
            const int yy = cy + a, zz = cz + b, xx = cx + c;
            if (yy >= gy || zz >= gz || xx >= gx) continue;
            const int cell = (yy * gz + zz) * gx + xx;
            if (s.gridcnt[cell] == 0) continue;
            const int jb = s.gridoff[cell], je = jb + s.gridcnt[cell];
            for (int j = jb; j < je; ++j) {
              const T* pj = &at<T>(s.pos, 3 * (size_t)j);
              const T d[3] = {(T)px - pj[0], (T)py - pj[1], (T)pz - pj[2]};
              if (std::fabs(d[0]) >= cs || std::fabs(d[1]) >= cs || std::fabs(d[2]) >= cs) continue;
              const T w = (1 - std::fabs(d[0]) / cs) * (1 - std::fabs(d[1]) / cs) *
                          (1 - std::fabs(d[2]) / cs);
              const T* vj = &at<T>(s.vel, 3 * (size_t)j);
              for (int q = 0; q < 3; ++q) sum[q] += vj[q] * w;
              wsum += w;
            }
          }
    }
    if (wsum > T(0.000001f))
      for (int q = 0; q < 3; ++q) v[q] = sum[q] / wsum;
    else
      for (int q = 0; q < 3; ++q) v[q] = 0;
  }
}

}  // namespace

extern "C" {
#define ORC_MPM(T, pfx)                                                                        \
  void pfx##mpm_p2g(const sphx_ParamCarrier* P, OrcState* s, OrcMPM* M, int apic) {            \
    p2g<T>(*P, *s, *M, apic);                                                                   \
  }                                                                                            \
  void pfx##mpm_stress(const sphx_ParamCarrier* P, OrcState* s, OrcMPM* M) {                   \
    stress<T>(*P, *s, *M);                                                                      \
  }                                                                                            \
  void pfx##mpm_node_update(const sphx_ParamCarrier* P, OrcState* s, OrcMPM* M) {              \
    node_update<T>(*P, *s, *M);                                                                 \
  }                                                                                            \
  void pfx##mpm_g2p(const sphx_ParamCarrier* P, OrcState* s, OrcMPM* M, int apic) {            \
    g2p<T>(*P, *s, *M, apic);                                                                   \
  }                                                                                            \
  void pfx##mpm_momentum(const sphx_ParamCarrier* P, OrcState* s, OrcMPM* M) {                 \
    mpm_momentum<T>(*P, *s, *M);                                                                \
  }                                                                                            \
  double pfx##mpm_weight(int kernel, double x) { return (double)N1<T>(kernel, (T)x); }          \
  double pfx##mpm_dweight(int kernel, double x) { return (double)dN1<T>(kernel, (T)x); }       \
  void pfx##mpm_drucker_prager(const sphx_ParamCarrier* P, double G, double K, const double* e, \
                               const double* sg, double* out) {                                \
    T ee[9], ss[9], pl[9];                                                                      \
    for (int i = 0; i < 9; ++i) { ee[i] = (T)e[i]; ss[i] = (T)sg[i]; }                          \
    drucker_prager<T>(*P, (T)G, (T)K, ee, ss, pl);                                              \
    for (int i = 0; i < 9; ++i) out[i] = (double)pl[i];                                         \
  }
ORC_MPM(float, orc_f32_)
ORC_MPM(double, orc_f64_)
}
