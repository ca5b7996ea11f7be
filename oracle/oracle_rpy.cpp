// CPU oracle (test infrastructure only) of the Stokesian mobility:
//   PairwiseMobMatrix        SPH_Solver/stokesian/Mobility.cu:37-82
//   Kern_getmobu_walkthrough :207-255
//   MVmultiply / getMobU     :85-109
// and the polydisperse Rotne-Prager-Yamakawa mobility (non-overlapping
// spheres, r >= a_i + a_j) that the MFMA path applies:
//   M_ii = I / (6 pi mu a_i)
//   M_ij = 1/(8 pi mu r) [ (1 + (a_i^2 + a_j^2)/(3 r^2)) I
//                          + (1 - (a_i^2 + a_j^2)/r^2) r^ r^ ]
// whose equal-radius limit, times 6 pi mu a, is the reference's A1N.
#include <cmath>
#include <cstring>

namespace {

// epsilon_ijk as setSignature (:20-35) stores it (index i*9 + j*3 + k)
double eps3(int i, int j, int k) {
  if ((i == 0 && j == 1 && k == 2) || (i == 2 && j == 0 && k == 1) || (i == 1 && j == 2 && k == 0))
    return 1;
  if ((i == 0 && j == 2 && k == 1) || (i == 2 && j == 1 && k == 0) || (i == 1 && j == 0 && k == 2))
    return -1;
  return 0;
}

// PairwiseMobMatrix: equal spheres a = 1, units 1/(6 pi mu a); only A1N and
// B1N are used by the walkthrough.
template <class T>
void pair_ref(const T drv[3], T A1N[9], T B1N[9]) {
  T dr = std::sqrt(drv[0] * drv[0] + drv[1] * drv[1] + drv[2] * drv[2]);
  T eu[3] = {drv[0] / dr, drv[1] / dr, drv[2] / dr};
  T dr3 = dr * dr * dr;
  T X1NA = T(1.5) / dr - T(1) / dr3;
  T Y1NA = T(0.75) / dr + T(0.5) / dr3;
  T Y1NB = T(-3) / (T(4) * dr * dr);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      A1N[i * 3 + j] = X1NA * eu[i] * eu[j] + Y1NA * (T(i == j) - eu[i] * eu[j]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T dot = 0;
      for (int k = 0; k < 3; ++k) dot = dot + (T)eps3(j, i, k) * eu[k];
      B1N[i * 3 + j] = -Y1NB * dot;
    }
}

}  // namespace

extern "C" {

void orc_rpy_pair_ref(const double* drv, double* A1N, double* B1N) { pair_ref<double>(drv, A1N, B1N); }

// Kern_getmobu_walkthrough over n spheres (pos [n][3], force [n][3]);
// unew [n][3] holds the background velocity on entry (Kern_interpolate-
// Velocity) and U on exit; omega [n][3]; positions advance by U dt into
// pos_out (the reference moves them in place while other threads read
// them, A10).
void orc_rpy_walkthrough(int n, const double* pos, const double* force, double dt,
                         double* unew, double* omega, double* pos_out) {
  for (int i = 0; i < n; ++i) {
    double u[3] = {unew[3 * i], unew[3 * i + 1], unew[3 * i + 2]};
    double w[3] = {0, 0, 0};
    for (int k = 0; k < n; ++k) {
      const double* f = force + 3 * k;
      if (k != i) {
        double d[3] = {pos[3 * k] - pos[3 * i], pos[3 * k + 1] - pos[3 * i + 1],
                       pos[3 * k + 2] - pos[3 * i + 2]};
        double A[9], B[9];
        pair_ref<double>(d, A, B);
        for (int a = 0; a < 3; ++a) {
          u[a] += A[3 * a] * f[0] + A[3 * a + 1] * f[1] + A[3 * a + 2] * f[2];
          w[a] += B[3 * a] * f[0] + B[3 * a + 1] * f[1] + B[3 * a + 2] * f[2];
        }
      } else {
        for (int a = 0; a < 3; ++a) u[a] += f[a];
      }
    }
    for (int a = 0; a < 3; ++a) {
      unew[3 * i + a] = u[a];
      omega[3 * i + a] = w[a];
      pos_out[3 * i + a] = pos[3 * i + a] + u[a] * dt;
    }
  }
}

// Polydisperse RPY applied to K right-hand sides: U[3n][K] = M F[3n][K]
// (row 3p + a), FP64.
void orc_rpy_apply(int n, const double* pos, const double* rad, double mu, const double* F,
                   int K, double* U) {
  const double pi = 3.141592653589793;
  std::memset(U, 0, sizeof(double) * 3 * (size_t)n * K);
  for (int p = 0; p < n; ++p) {
    double self = 1.0 / (6 * pi * mu * rad[p]);
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < K; ++c) U[(3 * (size_t)p + a) * K + c] += self * F[(3 * (size_t)p + a) * K + c];
    for (int q = 0; q < n; ++q) {
      if (q == p) continue;
      double d[3] = {pos[3 * q] - pos[3 * p], pos[3 * q + 1] - pos[3 * p + 1], pos[3 * q + 2] - pos[3 * p + 2]};
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = std::sqrt(r2);
      double s2 = rad[p] * rad[p] + rad[q] * rad[q];
      double c0 = 1.0 / (8 * pi * mu * r);
      double ci = c0 * (1 + s2 / (3 * r2)), cr = c0 * (1 - s2 / r2) / r2;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          double m = (a == b ? ci : 0.0) + cr * d[a] * d[b];
          const double* f = F + (3 * (size_t)q + b) * K;
          double* u = U + (3 * (size_t)p + a) * K;
          for (int c = 0; c < K; ++c) u[c] += m * f[c];
        }
    }
  }
}

}  // extern "C"
