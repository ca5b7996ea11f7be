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

}  // extern "C"
