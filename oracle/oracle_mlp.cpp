// CPU oracle (test infrastructure only): the learned warm start's feature
// map, the 24-64-64-4 MLP forward pass and its loss gradient, in FP64.
// The network has no counterpart in the reference; it replaces the warm
// start p0 = 0.5 p(t-1) of SPH_Solver/fluids/fluid_system_kern.cu:899 and
// corrects the drift velocity of SPH_Solver/fluids/multiphase.cu:131-163.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

extern "C" {
typedef struct OrcMF OrcMF;
}

namespace {

const uint32_t UNDEF = 0xFFFFFFFFu;
constexpr int NF = 24, H = 64, NO = 4;
constexpr int OFF_W1 = 0, OFF_B1 = NF * H, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + H * NO, NPARAM = OFF_B3 + NO;

double kern(double r, double h, double ks) {
  double q = r / h;
  if (q > 2) return 0;
  if (q > 1) return 0.25 * (2 - q) * (2 - q) * (2 - q) * ks;
  return (1 - 1.5 * q * q + 0.75 * q * q * q) * ks;
}
double nabla_fac(double r, double h, double ks) {
  double q = r / h;
  if (q > 2) return 0;
  if (q > 1) return 0.75 * (2 - q) * (2 - q) * (-1) / h / r * ks;
  return (-3 * q + 2.25 * q * q) / h / r * ks;
}

}  // namespace

extern "C" {

int orc_mlp_nparam(void) { return NPARAM; }

// Feature map of fluid particle i (after ComputeAii): X[n][24] in FP64.
// alpha4: [n][4] (alpha_0..2, rho_m) or NULL for a single phase; dp_last:
// [n] pressure increment of each particle's last solve, or NULL (zero).
void orc_mlp_features(const sphx_ParamCarrier* P, const OrcState* s, const float* alpha4,
                      const float* dp_last, double p_scale, double* X) {
  const double h = P->smoothradius, ks = P->kspline, ss = P->simscale, dt = P->dt;
  const double g[3] = {P->gravity.x, P->gravity.y, P->gravity.z};
  const double gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  const double gh[3] = {gn > 0 ? g[0] / gn : 0, gn > 0 ? g[1] / gn : -1, gn > 0 ? g[2] / gn : 0};
  const float* pos = static_cast<const float*>(s->pos);
  const float* ve = static_cast<const float*>(s->veleval);
  const float* dii = static_cast<const float*>(s->dii);
  auto F = [&](void* a, int i) { return (double)static_cast<const float*>(a)[i]; };
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->n; ++i) {
    double* x = X + (size_t)i * NF;
    std::memset(x, 0, sizeof(double) * NF);
    if (s->type[i] == SPHX_TYPE_BOUNDARY || s->mgcell[i] == UNDEF) continue;
    double rho0 = F(s->restdens, i), rho = F(s->dens, i), radv = F(s->rho_adv, i);
    double aii = F(s->aii, i);
    double pi = std::max(0.0, F(s->pressure, i));
    double wsum = 0, wp = 0, pmax = pi, pmin = pi, wsrc = 0, gp[3] = {0, 0, 0};
    int nf = 0, nb = 0;
    uint32_t gc = s->mgcell[i];
    for (int c = 0; c < P->neighbornum; ++c) {
      int cell = (int)gc + P->neighborid[c];
      if (s->gridcnt[cell] == 0) continue;
      int b = s->gridoff[cell], e = b + s->gridcnt[cell];
      for (int j = b; j < e; ++j) {
        if (j == i) continue;
        double d[3];
        for (int q = 0; q < 3; ++q) d[q] = ((double)pos[3 * i + q] - (double)pos[3 * j + q]) * ss;
        double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!(r < 2 * h && r > 0)) continue;
        if (s->type[j] == SPHX_TYPE_BOUNDARY) { ++nb; continue; }
        ++nf;
        double w = kern(r, h, ks);
        double pj = std::max(0.0, F(s->pressure, j));
        wsum += w;
        wp += w * pj;
        pmax = std::max(pmax, pj);
        pmin = std::min(pmin, pj);
        wsrc += w * (F(s->dens, j) - F(s->restdens, j)) / F(s->restdens, j);
        double vj = F(s->mass, j) / F(s->dens, j) * nabla_fac(r, h, ks) * (pj - pi);
        for (int q = 0; q < 3; ++q) gp[q] += vj * d[q];
      }
    }
    double pav = wsum > 0 ? wp / wsum : pi;
    double pj_est = std::fabs(aii) > 1e-12 ? std::max(0.0, (rho0 - radv) / aii) : 0.0;
    x[0] = pi / p_scale;
    x[1] = (radv - rho0) / rho0 * 100.0;
    x[2] = (rho - rho0) / rho0 * 100.0;
    x[3] = -aii * p_scale / rho0;
    x[4] = pj_est / p_scale;
    x[5] = pav / p_scale;
    x[6] = pmax / p_scale;
    x[7] = pmin / p_scale;
    x[8] = (pav - pi) / p_scale;
    x[9] = nf / 64.0;
    x[10] = nb / 64.0;
    x[11] = alpha4 ? (double)alpha4[4 * (size_t)i + 1] : 0.0;
    x[12] = alpha4 ? (double)alpha4[4 * (size_t)i + 3] / rho0 : 1.0;
    for (int q = 0; q < 3; ++q) x[13 + q] = (double)dii[3 * i + q] * p_scale / h;
    for (int q = 0; q < 3; ++q) x[16 + q] = (double)ve[3 * i + q] * dt / h;
    double den = rho0 * (gn > 0 ? gn : 1.0);
    x[19] = (gp[0] * gh[0] + gp[1] * gh[1] + gp[2] * gh[2]) / den;
    x[20] = dp_last ? (double)dp_last[i] / p_scale * 10.0 : 0.0;
    x[21] = gp[2] / den;
    x[22] = wsum > 0 ? wsrc / wsum * 100.0 : 0.0;
    x[23] = 1.0;
  }
}

// Forward pass: Y[n][4]; H1, H2 [n][64] (post-ReLU) may be NULL.
void orc_mlp_forward(const double* w, const double* X, int n, double* Y, double* H1,
                     double* H2) {
#pragma omp parallel for schedule(static)
  for (int s = 0; s < n; ++s) {
    const double* x = X + (size_t)s * NF;
    double h1[H], h2[H];
    for (int a = 0; a < H; ++a) {
      double v = w[OFF_B1 + a];
      for (int k = 0; k < NF; ++k) v += x[k] * w[OFF_W1 + k * H + a];
      h1[a] = v > 0 ? v : 0;
    }
    for (int b = 0; b < H; ++b) {
      double v = w[OFF_B2 + b];
      for (int a = 0; a < H; ++a) v += h1[a] * w[OFF_W2 + a * H + b];
      h2[b] = v > 0 ? v : 0;
    }
    for (int o = 0; o < NO; ++o) {
      double v = w[OFF_B3 + o];
      for (int b = 0; b < H; ++b) v += h2[b] * w[OFF_W3 + b * NO + o];
      Y[(size_t)s * NO + o] = v;
    }
    if (H1) std::memcpy(H1 + (size_t)s * H, h1, sizeof(h1));
    if (H2) std::memcpy(H2 + (size_t)s * H, h2, sizeof(h2));
  }
}

// Loss L = (1/n) sum_s sum_o lw[o] (y_so - t_so)^2 and its gradient with
// respect to all 6020 parameters (layout W1[24][64] b1 W2[64][64] b2
// W3[64][4] b3), by backpropagation.
double orc_mlp_loss_grad(const double* w, const double* X, const double* T, int n,
                         const double* lw, double* grad) {
  std::memset(grad, 0, sizeof(double) * NPARAM);
  double loss = 0;
  for (int s = 0; s < n; ++s) {
    const double* x = X + (size_t)s * NF;
    double h1[H], h2[H], y[NO];
    orc_mlp_forward(w, x, 1, y, h1, h2);
    double dy[NO];
    for (int o = 0; o < NO; ++o) {
      double e = y[o] - T[(size_t)s * NO + o];
      loss += lw[o] * e * e / n;
      dy[o] = 2.0 * lw[o] * e / n;
    }
    double dh2[H], dh1[H];
    for (int b = 0; b < H; ++b) {
      double v = 0;
      for (int o = 0; o < NO; ++o) {
        v += dy[o] * w[OFF_W3 + b * NO + o];
        grad[OFF_W3 + b * NO + o] += h2[b] * dy[o];
      }
      dh2[b] = h2[b] > 0 ? v : 0;
    }
    for (int o = 0; o < NO; ++o) grad[OFF_B3 + o] += dy[o];
    for (int a = 0; a < H; ++a) {
      double v = 0;
      for (int b = 0; b < H; ++b) {
        v += dh2[b] * w[OFF_W2 + a * H + b];
        grad[OFF_W2 + a * H + b] += h1[a] * dh2[b];
      }
      dh1[a] = h1[a] > 0 ? v : 0;
    }
    for (int b = 0; b < H; ++b) grad[OFF_B2 + b] += dh2[b];
    for (int k = 0; k < NF; ++k)
      for (int a = 0; a < H; ++a) grad[OFF_W1 + k * H + a] += x[k] * dh1[a];
    for (int a = 0; a < H; ++a) grad[OFF_B1 + a] += dh1[a];
  }
  return loss;
}

}  // extern "C"
