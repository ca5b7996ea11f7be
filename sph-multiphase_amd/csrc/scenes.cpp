// Host-side scene set-up: grid parameters, kernel constants and seeded
// synthetic particle sets.  No device code; usable without a GPU.
//
// sphx_setup_grid restates FluidSystem::SetupGrid (fluid_system.cpp:119-145)
// with the same mixed float/double expression types, so gridIdfac and the
// neighbour offsets are bit-identical to what the reference would carry to
// the device; sphx_setup_kernels restates the kernel constants of
// fluid_system.cpp:308-313.
#include <cmath>
#include <cstring>
#include <cstdint>
#include <algorithm>
#include <vector>

#include "../../include/sphx.h"

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // 24-bit uniform in [0, 1)
  float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
  double normal() {
    double u1 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = (next() >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

void init_particle(sphx_displayPack& d, sphx_calculationPack& q, float x, float y, float z,
                   int type, float mass, float rest, float visc, int born) {
  std::memset(&d, 0, sizeof(d));
  std::memset(&q, 0, sizeof(q));
  d.pos = {x, y, z};
  d.type = type;
  if (type == SPHX_TYPE_BOUNDARY) d.color = {1.f, 1.f, 1.f, 0.2f};
  else d.color = {0.2f, 0.5f, 0.8f, 0.5f};
  q.mass = mass;
  q.restdens = rest;
  q.visc = visc;
  q.bornid = born;
  q.deformGrad.data[0] = q.deformGrad.data[4] = q.deformGrad.data[8] = 1.f;
  q.X = {x, y, z};
}

}  // namespace

extern "C" {

void sphx_setup_grid(sphx_ParamCarrier* p) {
  float worldcellsize = p->cellsize / p->simscale;
  p->gridmin = p->volmin;
  p->gridmax = p->volmax;
  p->gridsize = {p->gridmax.x - p->gridmin.x, p->gridmax.y - p->gridmin.y,
                 p->gridmax.z - p->gridmin.z};
  p->gridres.x = (int)std::ceil(p->gridsize.x / worldcellsize);
  p->gridres.y = (int)std::ceil(p->gridsize.y / worldcellsize);
  p->gridres.z = (int)std::ceil(p->gridsize.z / worldcellsize);
  p->gridsize.x = p->gridres.x * worldcellsize;
  p->gridsize.y = p->gridres.y * worldcellsize;
  p->gridsize.z = p->gridres.z * worldcellsize;
  // 1.0 is a double: the quotient and product are formed in double and
  // rounded once to float, as in fluid_system.cpp:132-134
  p->gridIdfac.x = (float)(1.0 / p->gridsize.x * p->gridres.x);
  p->gridIdfac.y = (float)(1.0 / p->gridsize.y * p->gridres.y);
  p->gridIdfac.z = (float)(1.0 / p->gridsize.z * p->gridres.z);
  p->gridtotal = p->gridres.x * p->gridres.y * p->gridres.z;
  p->searchnum = 3;
  p->neighbornum = 27;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        p->neighborid[i * 9 + j * 3 + k] =
            (i - 1) * p->gridres.x * p->gridres.z + (j - 1) * p->gridres.x + (k - 1);
}

void sphx_setup_kernels(sphx_ParamCarrier* p) {
  float h = p->smoothradius;
  p->cellsize = h * 2;
  p->kpoly6 = (float)(315.0f / (64.0f * 3.141592 * std::pow(h, 9.0f)));
  p->kspiky = (float)(15 / (3.141592 * std::pow(h, 6.0f)));
  p->kspikydiff = (float)(-45.0f / (3.141592 * std::pow(h, 6.0f)));
  p->klaplacian = (float)(45.0f / (3.141592 * std::pow(h, 6.0f)));
  p->kspline = 1.0f / 3.141593f / std::pow(h, 3.0f);
}

int sphx_scene_dambreak(int nx, int ny, int nz, float s, float Lx, float Ly, float Lz, float h,
                        int layers, float jitter, uint64_t seed, int kernel_family,
                        sphx_ParamCarrier* params, sphx_displayPack* disp,
                        sphx_calculationPack* calc, sphx_SceneInfo* info) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !(s > 0) || !(h > 0) || layers < 0) return -1;
  const float rho0 = 1000.f;
  const float m = rho0 * s * s * s;
  // boundary lattice: `layers` layers of spacing s outside the box walls
  // (floor and four sides; open top), as the reference's open box
  // (AddShape.cpp:355-366)
  std::vector<float> bx, by, bz;
  int bnx = (int)std::lround(Lx / s), bny = (int)std::lround(Ly / s), bnz = (int)std::lround(Lz / s);
  for (int iy = -layers; iy < bny; ++iy)
    for (int iz = -layers; iz < bnz + layers; ++iz)
      for (int ix = -layers; ix < bnx + layers; ++ix) {
        bool inside = ix >= 0 && ix < bnx && iz >= 0 && iz < bnz && iy >= 0;
        if (inside) continue;
        bx.push_back((ix + 0.5f) * s);
        by.push_back((iy + 0.5f) * s);
        bz.push_back((iz + 0.5f) * s);
      }
  int nf = nx * ny * nz;
  int nb = (int)bx.size();
  if (info) {
    info->n_fluid = nf;
    info->n_boundary = nb;
    info->n_total = nf + nb;
    info->n_phase = 1;
    info->spacing = s;
  }
  if (params) {
    sphx_ParamCarrier& P = *params;
    std::memset(&P, 0, sizeof(P));
    P.acclimit = 1500.f;
    P.vlimit = 100.f;
    P.gravity = {0.f, -9.81f, 0.f};
    P.simscale = 1.f;
    P.dt = 1e-4f;
    P.smoothradius = h;
    P.radius = s * 0.5f;
    P.restdensity = rho0;
    P.mass = m;
    P.viscosity = 0.f;
    for (int k = 0; k < SPHX_MAX_FLUIDNUM; ++k) {
      P.massArr[k] = m;
      P.densArr[k] = rho0;
      P.viscArr[k] = 0.f;
    }
    P.bmass = m;
    P.bRestdensity = rho0;
    P.bvisc = 0.4f;   // 2 alpha c0 with alpha = 0.01, c0 = 20
    sphx_setup_kernels(&P);
    float cell = P.cellsize;
    int margin = (int)std::ceil(layers * s / cell) + 1;
    P.volmin = {-margin * cell, -margin * cell, -margin * cell};
    P.volmax = {Lx + margin * cell, Ly + margin * cell, Lz + margin * cell};
    P.softminx = {0.f, 0.f, 0.f};
    P.softmaxx = {Lx, Ly, Lz};
    sphx_setup_grid(&P);
    P.num = nf + nb;
    P.maxNum = nf + nb;
    (void)kernel_family;
  }
  if (disp && calc) {
    Rng rng(seed);
    int p = 0;
    for (int iy = 0; iy < ny; ++iy)
      for (int iz = 0; iz < nz; ++iz)
        for (int ix = 0; ix < nx; ++ix, ++p) {
          float jx = (rng.uniform() * 2.f - 1.f) * jitter;
          float jy = (rng.uniform() * 2.f - 1.f) * jitter;
          float jz = (rng.uniform() * 2.f - 1.f) * jitter;
          init_particle(disp[p], calc[p], (ix + 0.5f) * s + jx, (iy + 0.5f) * s + jy,
                        (iz + 0.5f) * s + jz, SPHX_TYPE_FLUID, m, rho0, 0.f, p);
        }
    for (int b = 0; b < nb; ++b, ++p)
      init_particle(disp[p], calc[p], bx[b], by[b], bz[b], SPHX_TYPE_BOUNDARY, m, rho0, 0.4f, p);
  }
  return nf + nb;
}

// Two-phase bubble tank: n_fluid particles on a cubic lattice of spacing
// s = (fill * Lx * Lz / n_fluid)^(1/3) filling a Lx x Ly x Lz tank from the
// floor in y-major order; n_bubbles spheres with centres uniform in the
// water and radii uniform in [rmin, rmax] mark air (alpha_air = 1).  All
// fluid particles carry the solver mass rho_ref * s^3 and rest density
// rho_ref (the pressure solve enforces the mixture's volume); the phase
// fractions go to `alpha` ([n_total][2]: water, air; boundary rows 0).
int sphx_scene_bubbles(int n_fluid, float Lx, float Ly, float Lz, float fill, float h,
                       int n_bubbles, float rmin, float rmax, int layers, uint64_t seed,
                       sphx_ParamCarrier* params, sphx_displayPack* disp,
                       sphx_calculationPack* calc, float* alpha, sphx_SceneInfo* info) {
  if (n_fluid <= 0 || !(Lx > 0) || !(Ly > 0) || !(Lz > 0) || !(fill > 0) || !(h > 0) ||
      layers < 0 || n_bubbles < 0 || rmax < rmin)
    return -1;
  const float rho_ref = 1000.f;
  double sd = std::cbrt((double)fill * Lx * Ly * Lz / n_fluid);
  float s = (float)sd;
  int nx = (int)std::floor(Lx / s), nz = (int)std::floor(Lz / s);
  if (nx < 1 || nz < 1) return -1;
  int ny = (n_fluid + nx * nz - 1) / (nx * nz);
  // boundary walls (floor + 4 sides, open top), `layers` layers outside
  std::vector<float> bx, by, bz;
  int bnx = nx, bnz = nz, bny = (int)std::ceil(Ly / s);
  for (int iy = -layers; iy < bny; ++iy)
    for (int iz = -layers; iz < bnz + layers; ++iz)
      for (int ix = -layers; ix < bnx + layers; ++ix) {
        bool inside = ix >= 0 && ix < bnx && iz >= 0 && iz < bnz && iy >= 0;
        if (inside) continue;
        bx.push_back((ix + 0.5f) * s);
        by.push_back((iy + 0.5f) * s);
        bz.push_back((iz + 0.5f) * s);
      }
  int nb = (int)bx.size();
  if (info) {
    info->n_fluid = n_fluid;
    info->n_boundary = nb;
    info->n_total = n_fluid + nb;
    info->n_phase = 2;
    info->spacing = s;
  }
  // Lattice sum f = sum_lattice s^3 W(r) of the cubic spline (support 2h):
  // mass rho_ref s^3 / f puts a resting lattice exactly at rho_ref, and the
  // wall particles (a continuation of the lattice) get the volume s^3 / f,
  // so fluid next to walls, edges and corners sees a complete neighbourhood.
  double fsum = 0.0;
  {
    const double kspl = 1.0 / 3.141593 / ((double)h * h * h);
    int R = (int)std::ceil(2.0 * h / s) + 1;
    for (int a = -R; a <= R; ++a)
      for (int b = -R; b <= R; ++b)
        for (int cc = -R; cc <= R; ++cc) {
          double q = std::sqrt((double)a * a + (double)b * b + (double)cc * cc) * s / h;
          double w = q > 2 ? 0.0 : (q > 1 ? 0.25 * (2 - q) * (2 - q) * (2 - q)
                                           : 1 - 1.5 * q * q + 0.75 * q * q * q);
          fsum += w * kspl * sd * sd * sd;
        }
  }
  const float m = (float)(rho_ref * sd * sd * sd / fsum);
  const float vb = (float)(sd * sd * sd / fsum);
  if (params) {
    sphx_ParamCarrier& P = *params;
    std::memset(&P, 0, sizeof(P));
    P.acclimit = 1500.f;
    P.vlimit = 100.f;
    P.gravity = {0.f, -9.81f, 0.f};
    P.simscale = 1.f;
    P.dt = 1e-3f;
    P.smoothradius = h;
    P.radius = s * 0.5f;
    P.restdensity = rho_ref;
    P.mass = m;
    P.viscosity = 1e-6f;
    P.massArr[0] = m; P.massArr[1] = m; P.massArr[2] = m;
    P.densArr[0] = 1000.f; P.densArr[1] = 1.2f; P.densArr[2] = 1000.f;
    P.viscArr[0] = 1e-6f; P.viscArr[1] = 1.5e-5f; P.viscArr[2] = 1e-6f;
    P.bmass = m;
    P.bRestdensity = rho_ref;
    P.bvisc = 0.4f;
    sphx_setup_kernels(&P);
    float cell = P.cellsize;
    int margin = (int)std::ceil(layers * s / cell) + 1;
    P.volmin = {-margin * cell, -margin * cell, -margin * cell};
    P.volmax = {nx * s + margin * cell, bny * s + margin * cell, nz * s + margin * cell};
    P.softminx = {0.f, 0.f, 0.f};
    P.softmaxx = {nx * s, bny * s, nz * s};
    sphx_setup_grid(&P);
    P.num = n_fluid + nb;
    P.maxNum = n_fluid + nb;
  }
  if (disp && calc) {
    Rng rng(seed);
    const float water_top = (float)ny * s;
    std::vector<float> cx(n_bubbles), cy(n_bubbles), cz(n_bubbles), rr(n_bubbles);
    for (int b = 0; b < n_bubbles; ++b) {
      cx[b] = rng.uniform() * nx * s;
      cy[b] = rng.uniform() * water_top;
      cz[b] = rng.uniform() * nz * s;
      rr[b] = rmin + rng.uniform() * (rmax - rmin);
    }
    // bucket bubbles by a coarse grid for the inside test
    const float cb = 2.f * rmax > s ? 2.f * rmax : s;
    int gx = (int)std::ceil(nx * s / cb) + 1, gy = (int)std::ceil(water_top / cb) + 1,
        gz = (int)std::ceil(nz * s / cb) + 1;
    std::vector<std::vector<int>> bucket((size_t)gx * gy * gz);
    for (int b = 0; b < n_bubbles; ++b) {
      int ix = std::min(gx - 1, (int)(cx[b] / cb)), iy = std::min(gy - 1, (int)(cy[b] / cb)),
          iz = std::min(gz - 1, (int)(cz[b] / cb));
      bucket[((size_t)iy * gz + iz) * gx + ix].push_back(b);
    }
    int p = 0;
    for (int iy = 0; iy < ny && p < n_fluid; ++iy)
      for (int iz = 0; iz < nz && p < n_fluid; ++iz)
        for (int ix = 0; ix < nx && p < n_fluid; ++ix, ++p) {
          float x = (ix + 0.5f) * s, y = (iy + 0.5f) * s, z = (iz + 0.5f) * s;
          init_particle(disp[p], calc[p], x, y, z, SPHX_TYPE_FLUID, m, rho_ref, 1e-6f, p);
          // hydrostatic initial pressure rho_ref g (top - y)
          calc[p].pressure = rho_ref * 9.81f * std::max(0.f, water_top - y);
          bool air = false;
          int bx0 = (int)(x / cb), by0 = (int)(y / cb), bz0 = (int)(z / cb);
          for (int dy = -1; dy <= 1 && !air; ++dy)
            for (int dz = -1; dz <= 1 && !air; ++dz)
              for (int dx = -1; dx <= 1 && !air; ++dx) {
                int qx = bx0 + dx, qy = by0 + dy, qz = bz0 + dz;
                if (qx < 0 || qy < 0 || qz < 0 || qx >= gx || qy >= gy || qz >= gz) continue;
                for (int b : bucket[((size_t)qy * gz + qz) * gx + qx]) {
                  float ddx = x - cx[b], ddy = y - cy[b], ddz = z - cz[b];
                  if (ddx * ddx + ddy * ddy + ddz * ddz < rr[b] * rr[b]) { air = true; break; }
                }
              }
          if (alpha) {
            alpha[2 * p] = air ? 0.f : 1.f;
            alpha[2 * p + 1] = air ? 1.f : 0.f;
          }
          if (air) disp[p].color = {0.9f, 0.9f, 1.f, 0.3f};
        }
    for (int b = 0; b < nb; ++b, ++p) {
      init_particle(disp[p], calc[p], bx[b], by[b], bz[b], SPHX_TYPE_BOUNDARY, m, rho_ref, 0.4f, p);
      calc[p].dens = vb;   // boundary volume (lattice volume, see above)
      if (alpha) { alpha[2 * p] = 0.f; alpha[2 * p + 1] = 0.f; }
    }
  }
  return n_fluid + nb;
}

}  // extern "C"

extern "C" {

// SetupMPMGrid (fluid_system.cpp:1004-1025), in float as cfloat3 / cint3:
// the node count per axis is truncated to int (cint3::Set(int,...)).
void sphx_setup_mpm_grid(sphx_ParamCarrier* p, float pSpacing) {
  sphx_float3 lo = {p->softminx.x - 10.f, p->softminx.y - 10.f, p->softminx.z - 10.f};
  sphx_float3 hi = {p->softmaxx.x + 10.f, p->softmaxx.y + 10.f, p->softmaxx.z + 10.f};
  float cs = (float)(pSpacing * 1.5);
  const bool ok = cs > 0.f && std::isfinite(cs);   // else an empty grid
  float tx = ok ? (hi.x - lo.x) / cs + 2.f : 0.f, ty = ok ? (hi.y - lo.y) / cs + 2.f : 0.f,
        tz = ok ? (hi.z - lo.z) / cs + 2.f : 0.f;
  // a count that is not a finite non-negative int is stored as 0
  auto to_int = [](float t) { return t >= 0.f && t < 2147483648.f ? (int)t : 0; };
  sphx_int3 res = {to_int(tx), to_int(ty), to_int(tz)};
  p->mpmXmin = lo;
  p->mpmXmax = {lo.x + (float)(res.x - 1) * cs, lo.y + (float)(res.y - 1) * cs,
                lo.z + (float)(res.z - 1) * cs};
  p->mpmcellsize = cs;
  p->mpmRes = res;
  // -1: more nodes than the ParamCarrier's int count holds (sphx_mpm_setup
  // refuses such a grid)
  const long long nodes = (long long)res.x * res.y * res.z;
  p->mpmNodeNum = nodes > 2147483647LL ? -1 : (int)nodes;
}

// SetDruckerPragerParam (fluid_system.cpp:882-892): friction and dilation
// angles in degrees, cohesion in k_c.
void sphx_set_drucker_prager(sphx_ParamCarrier* p) {
  float phi = p->a_phi * (3.141593f / 180.0f);
  float psi = p->a_psi * (3.141593f / 180.0f);
  float kc = p->k_c;
  p->a_phi = (float)(2 * std::sin(phi) / std::sqrt(3) / (3 - std::sin(phi)));
  p->a_psi = (float)(2 * std::sin(psi) / std::sqrt(3) / (3 - std::sin(psi)));
  p->k_c = (float)(3 * kc * std::cos(psi) / std::sqrt(3) / (3 - std::sin(phi)));
}

// Granular block for the MPM transfer: n particles uniform in the cube
// [lo, hi]^3 (metres, simscale 1) over a res^3 node grid of spacing dx from
// the origin; particle mass rho (hi-lo)^3 / n; G = E/(2(1+nu)),
// K = E/(3(1-2nu)) (fluid_system.cpp:290-291); Drucker-Prager friction
// angle phi (degrees), no cohesion or dilation; deformation gradients
// I + N(0, f_noise^2).  Soft walls 3 nodes inside the grid.  The SPH sort
// grid has cells of 4 dx (h = 2 dx).  Arrays may be NULL to query sizes.
int sphx_scene_sand(int n, float lo, float hi, float dx, int res, float E, float nu,
                    float phi_deg, float rho, float f_noise, uint64_t seed,
                    sphx_ParamCarrier* p, sphx_displayPack* disp, sphx_calculationPack* calc,
                    sphx_SceneInfo* info) {
  if (n <= 0 || !(hi > lo) || !(dx > 0.f) || res < 8) return -1;
  if (info) {
    info->n_fluid = n;
    info->n_boundary = 0;
    info->n_total = n;
    info->n_phase = 1;
    info->spacing = std::cbrt((hi - lo) * (hi - lo) * (hi - lo) / (float)n);
  }
  if (!p) return n;
  std::memset(p, 0, sizeof(*p));
  p->simscale = 1.f;
  p->dt = 5e-5f;
  p->gravity = {0.f, -9.81f, 0.f};
  p->restdensity = rho;
  p->mass = rho * (hi - lo) * (hi - lo) * (hi - lo) / (float)n;
  p->smoothradius = 2.f * dx;
  sphx_setup_kernels(p);
  float cell = p->cellsize / p->simscale;
  float ext = (float)res * dx;
  p->volmin = {-cell, -cell, -cell};
  p->volmax = {ext + cell, ext + cell, ext + cell};
  sphx_setup_grid(p);
  p->softminx = {3.f * dx, 3.f * dx, 3.f * dx};
  p->softmaxx = {(float)(res - 4) * dx, (float)(res - 4) * dx, (float)(res - 4) * dx};
  p->mpmXmin = {0.f, 0.f, 0.f};
  p->mpmcellsize = dx;
  p->mpmRes = {res, res, res};
  p->mpmXmax = {(float)(res - 1) * dx, (float)(res - 1) * dx, (float)(res - 1) * dx};
  p->mpmNodeNum = res * res * res;
  p->solidG = E / 2 / (1 + nu);
  p->solidK = E / 3 / (1 - 2 * nu);
  p->a_phi = phi_deg;
  p->a_psi = 0.f;
  p->k_c = 0.f;
  sphx_set_drucker_prager(p);
  p->num = n;
  p->maxNum = n;
  if (!disp || !calc) return n;
  Rng rng(seed);
  for (int i = 0; i < n; ++i) {
    float x = lo + (hi - lo) * rng.uniform();
    float y = lo + (hi - lo) * rng.uniform();
    float z = lo + (hi - lo) * rng.uniform();
    init_particle(disp[i], calc[i], x, y, z, SPHX_TYPE_GRANULAR, p->mass, rho, 0.f, i);
    calc[i].dens = rho;
    calc[i].solidG = p->solidG;
    calc[i].solidK = p->solidK;
    for (int k = 0; k < 9; ++k)
      calc[i].deformGrad.data[k] = (k % 4 == 0 ? 1.f : 0.f) + (float)(f_noise * rng.normal());
  }
  return n;
}

}  // extern "C"

extern "C" {

// Batched ensemble: nscenes two-phase tanks (as sphx_scene_bubbles) of
// n_fluid[s] particles, each a cube of side (n_fluid[s] / 2^20)^(1/3) m
// filled to `fill` (the spacing of the 2^20-particle tank), with
// bubbles_per_million * n_fluid[s] / 2^20 bubbles seeded per scene, laid
// out in disjoint, cell-aligned boxes of one grid (2-cell gaps, wider than
// the kernel support, so the scenes never interact).  The scene index goes
// into bits 8-23 of each particle's type word.  scene_first[s] receives the
// first particle of scene s (length nscenes + 1).  Returns the particle
// count (arrays may be NULL to query it).
int sphx_scene_ensemble(int nscenes, const int* n_fluid, float fill, float h,
                        float bubbles_per_million, float rmin, float rmax, int layers,
                        uint64_t seed, sphx_ParamCarrier* params, sphx_displayPack* disp,
                        sphx_calculationPack* calc, float* alpha, int* scene_first,
                        sphx_SceneInfo* info) {
  if (nscenes <= 0 || nscenes > 65535 || !n_fluid || !(h > 0)) return -1;
  std::vector<float> side(nscenes);
  std::vector<int> total(nscenes), fluid(nscenes);
  std::vector<sphx_ParamCarrier> sp(nscenes);
  float ext_max = 0.f;
  int ntot = 0, nfl = 0, nbd = 0;
  float spacing = 0.f;
  for (int s = 0; s < nscenes; ++s) {
    if (n_fluid[s] <= 0) return -1;
    side[s] = (float)std::cbrt((double)n_fluid[s] / (double)(1 << 20));
    int nb = (int)std::lround(bubbles_per_million * n_fluid[s] / (double)(1 << 20));
    sphx_SceneInfo si;
    int rc = sphx_scene_bubbles(n_fluid[s], side[s], side[s], side[s], fill, h, nb, rmin, rmax,
                                layers, seed + 7919ull * (uint64_t)s, &sp[s], nullptr, nullptr,
                                nullptr, &si);
    if (rc < 0) return -1;
    total[s] = si.n_total;
    fluid[s] = si.n_fluid;
    ntot += si.n_total;
    nfl += si.n_fluid;
    nbd += si.n_boundary;
    spacing = si.spacing;
    ext_max = std::max(ext_max, side[s] + 2.f * layers * si.spacing);
  }
  if (info) {
    info->n_fluid = nfl;
    info->n_boundary = nbd;
    info->n_total = ntot;
    info->n_phase = 2;
    info->spacing = spacing;
  }
  const float cell = 2.f * h;
  const int per = (int)std::ceil(std::cbrt((double)nscenes) - 1e-9);
  const int slot_cells = (int)std::ceil(ext_max / cell) + 2;   // 2-cell gap
  const float slot = slot_cells * cell;
  const int wall_cells = (int)std::ceil(layers * spacing / cell) + 1;
  if (params) {
    *params = sp[0];
    sphx_ParamCarrier& P = *params;
    float ext = per * slot + 2.f * cell;
    P.volmin = {-cell, -cell, -cell};
    P.volmax = {ext, ext, ext};
    P.softminx = {0.f, 0.f, 0.f};
    P.softmaxx = {ext - cell, ext - cell, ext - cell};
    sphx_setup_grid(&P);
    P.num = ntot;
    P.maxNum = ntot;
  }
  if (scene_first) {
    int o = 0;
    for (int s = 0; s < nscenes; ++s) { scene_first[s] = o; o += total[s]; }
    scene_first[nscenes] = o;
  }
  if (!disp || !calc) return ntot;
  int o = 0;
  for (int s = 0; s < nscenes; ++s) {
    int ix = s % per, iz = (s / per) % per, iy = s / (per * per);
    float ox = ix * slot + wall_cells * cell, oy = iy * slot + wall_cells * cell,
          oz = iz * slot + wall_cells * cell;
    int nb = (int)std::lround(bubbles_per_million * n_fluid[s] / (double)(1 << 20));
    sphx_SceneInfo si;
    sphx_ParamCarrier tmp;
    sphx_scene_bubbles(n_fluid[s], side[s], side[s], side[s], fill, h, nb, rmin, rmax, layers,
                       seed + 7919ull * (uint64_t)s, &tmp, disp + o, calc + o,
                       alpha ? alpha + 2 * (size_t)o : nullptr, &si);
    for (int i = 0; i < total[s]; ++i) {
      disp[o + i].pos.x += ox;
      disp[o + i].pos.y += oy;
      disp[o + i].pos.z += oz;
      disp[o + i].type |= s << 8;
      calc[o + i].bornid = o + i;
    }
    o += total[s];
  }
  return ntot;
}

}  // extern "C"
